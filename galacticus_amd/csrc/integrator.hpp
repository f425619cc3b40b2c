// integrator.hpp -- device Cash-Karp (4,5) stepper pieces and the scaled
// step controller.
//
// Controller: sc2_control_hadjust (source/external/gslODEInitVal2/cscal2.c:93-169)
// restated bit for bit (FMA contraction off so each product and sum rounds
// as written).  Tableau: Cash & Karp (1990), ACM TOMS 16:201, as used by the
// GSL 2.6 rkck stepper that the reference selects (interface/GSL/C/odeiv2.c:39-41).
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>

namespace gal {

// GSL_ODEIV_HADJ_* values
constexpr int kHadjDec = -1;
constexpr int kHadjNil = 0;
constexpr int kHadjInc = 1;
constexpr int kRkckOrder = 4;

namespace ck {
constexpr double a2 = 1.0 / 5.0, a3 = 3.0 / 10.0, a4 = 3.0 / 5.0, a5 = 1.0, a6 = 7.0 / 8.0;
constexpr double b21 = 1.0 / 5.0;
constexpr double b31 = 3.0 / 40.0, b32 = 9.0 / 40.0;
constexpr double b41 = 3.0 / 10.0, b42 = -9.0 / 10.0, b43 = 6.0 / 5.0;
constexpr double b51 = -11.0 / 54.0, b52 = 5.0 / 2.0, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0;
constexpr double b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0, b64 = 44275.0 / 110592.0,
                 b65 = 253.0 / 4096.0;
constexpr double c1 = 37.0 / 378.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0;
constexpr double ec1 = 37.0 / 378.0 - 2825.0 / 27648.0;
constexpr double ec3 = 250.0 / 621.0 - 18575.0 / 48384.0;
constexpr double ec4 = 125.0 / 594.0 - 13525.0 / 55296.0;
constexpr double ec5 = -277.0 / 14336.0;
constexpr double ec6 = 512.0 / 1771.0 - 1.0 / 4.0;
}  // namespace ck

// Accumulates the error-norm maximum one variable at a time.  D0 and r as
// cscal2.c:116-120; rmax starts at DBL_MIN and uses GSL_MAX_DBL, so a NaN
// ratio (zero scale and zero error) never replaces it.
struct ErrorNorm {
  double rmax = DBL_MIN;
  bool forbidden = false;
};

__device__ __forceinline__ void error_norm_add(ErrorNorm& e, double eps_abs, double eps_rel, double a_y, double a_dydt,
                                               double h_old, double y, double yerr, double yp, double scale_abs,
                                               bool non_negative) {
#pragma clang fp contract(off)
  const double D0 = eps_rel * (a_y * fabs(y) + a_dydt * fabs(h_old * yp)) + eps_abs * scale_abs;
  const double r = fabs(yerr) / fabs(D0);
  e.rmax = (r > e.rmax) ? r : e.rmax;
  if (non_negative && y < 0.0) e.forbidden = true;
}

// Step decision from the accumulated norm (cscal2.c:127-168).
__device__ __forceinline__ int hadjust_decide(const ErrorNorm& e, int ord, double& h) {
#pragma clang fp contract(off)
  const double S = 0.9;
  const double h_old = h;
  if (e.rmax > 1.1) {
    double r = S / pow(e.rmax, 1.0 / ord);
    if (r < 0.2) r = 0.2;
    h = r * h_old;
    return kHadjDec;
  } else if (e.forbidden) {
    h = 0.5 * h_old;
    return kHadjDec;
  } else if (e.rmax < 0.5) {
    double r = S / pow(e.rmax, 1.0 / (ord + 1.0));
    if (r > 4.9) r = 4.9;
    if (r < 1.0) r = 1.0;
    h = r * h_old;
    return kHadjInc;
  }
  return kHadjNil;
}

}  // namespace gal
