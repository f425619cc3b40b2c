// physics.hpp -- device RHS of the node ODE system (one node per lane).
//
// dy/dt = nodeOperatorMulti%differentialEvolution (nodes/operators/multi.F90:313-332)
// over the operators of parameters/quickTest.xml:220-289, in list order,
// after the analytic solutions of the basic component
// (nodes/operators/physics/cosmic_time.F90:81-95,
//  dark_matter_only_mass/interpolate.F90:215-235).  Each block cites the
// reference routine whose rate it computes.
//
// The formulas are the reference's; the evaluation is arranged for a GPU
// lane with a long dependent chain, where every instruction counts:
//  * The cooling radius -- a bracketed numerical root of
//    t_cool(r) = t_available in the reference (cooling_radius/simple.F90:365-381)
//    -- is evaluated in closed form: with the isothermal (virial) temperature
//    profile and a CIE cooling function, t_cool(r) = K / rho(r), and the
//    beta=2/3 profile inverts to r = r_c sqrt(rho0 t_avail / K - 1).
//  * Power laws whose exponent is a multiple of 1/2 (all four of
//    quickTest.xml: -1.5, 2, 3.5, 3.5) are products of the base and its
//    square root instead of exp(alpha ln x); other exponents use exp/log.
//  * a^3 and the virial overdensity are formed without cbrt:
//    Delta_m rho_mean = Delta_c rho_crit0 E^2(a), E^2 = Omega_m/a^3 + Omega_L.
//  * The jet power's two power laws and spin exponential share one exp:
//    (M/1e9)^0.9 (Mdot/Mdot_Edd)^1.2 = C M^-0.3 Mdot^1.2.
//  * Constants that depend only on the parameters are computed on the host
//    (DerivedConstants).
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/gal_ode.h"
#include "constants.hpp"

namespace gal {

constexpr int kNV = GAL_NVAR;
constexpr double kDoubleMax = 1.7976931348623157e308;   // huge(0d0)
constexpr int kNoHalfPower = -100000;

// Constants derived from the physics parameters once per set_physics call.
struct DerivedConstants {
  double om, ol;          // Omega_m, Omega_L
  double a3_norm;         // Omega_m/Omega_L: a^3 = a3_norm sinh^2(sinh_arg t)
  double sinh_arg;        // 1.5 sqrt(Omega_L) H0 [1/Gyr]
  double rvir_coeff;      // 3 / (4 pi rho_crit0)
  double a3_reion;        // (1/(1+z_reion))^3
  double f_baryon;        // Omega_b / Omega_m
  double inv_sqrt_200;    // 1/sqrt(200): dynamical-time velocity zero point
  double inv_200;
  double v_fb_disk, sqrt_v_fb_disk, v_fb_sph, sqrt_v_fb_sph;
  double ln_200, ln_v_fb_disk, ln_v_fb_sph;
  int n2_sf_disk, n2_sf_sph, n2_fb_disk, n2_fb_sph;  // 2*exponent, or kNoHalfPower
  double t_vir_per_v2;    // 0.5 mu m_u (1e3)^2 / k
  double cs2_per_t;       // (5/3) k / (mu m_u) / (1e3)^2  [(km/s)^2 / K]
  double cs2_sph;         // cs^2 at bondiHoyleAccretionTemperatureSpheroid
  double m_edd_per_mass;  // Eddington accretion rate per Msun [1/Gyr]
  double ln_jet_kerr;     // ln of the jet-power prefactor for spin > 0.8
  double ln_jet_schw;     // ln of the jet-power prefactor for spin <= 0.8
  double cool_k_pre;      // (dof/2) k / ergs / Gyr
  double unit_energy;     // Msun (km/s)^2 / Gyr -> erg/s
  double nh_coeff;        // Msun / m_H / (Mpc in cm)^3
  double inv_eff_sf_disk, inv_eff_sf_sph;        // 1/efficiency (0 when efficiency is 0)
  double inv_tmin_fb_disk, inv_tmin_fb_sph;      // 1/timescaleOutflowFractionalMinimum
  double inv_core;                               // 1/coreRadiusOverVirialRadius
  double inv_cs2_sph;                            // 1/cs^2 at the spheroid accretion temperature
  double bhl_sph_coeff;                          // enhancement * 4 pi G^2 (unit) / cs^3, spheroid mode
  double jeans_coeff;                            // G / cs^2 at the spheroid accretion temperature
  double h0_gyr;                                 // H0 [1/Gyr]
  double inv_one_minus_j_loss;                   // 1 / (1 - fractionLossAngularMomentum)
  double wind_pressure_ratio;                    // P_ISM / P_wind = this * M_gas / (r eff Mdot)
};

// Reciprocal to full double precision: hardware estimate plus two
// Newton-Raphson corrections (Markstein 1990).  Callers only pass finite,
// non-zero values or select the result away.
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Cooling table descriptor (table itself in global memory).
struct CoolingDesc {
  int n_t, n_z;
  double ln_t_min, d_ln_t, inv_d_ln_t;
  double ln_z_min, d_ln_z, inv_d_ln_z;
  double t_min, t_max, z_min, z_max;
};

struct NodePar {
  double t_ref, m_ref, m_rate, spin_rate, ram_force, v_host;
  bool satellite;
};

__device__ __forceinline__ double sq(double x) { return x * x; }

__device__ __forceinline__ NodePar load_node_par(const double* __restrict__ np, int ld, int i) {
  NodePar p;
  p.t_ref = np[GAL_NP_TIME_REFERENCE * (size_t)ld + i];
  p.m_ref = np[GAL_NP_MASS_REFERENCE * (size_t)ld + i];
  p.m_rate = np[GAL_NP_MASS_RATE * (size_t)ld + i];
  p.spin_rate = np[GAL_NP_SPIN_RATE * (size_t)ld + i];
  p.satellite = np[GAL_NP_IS_SATELLITE * (size_t)ld + i] != 0.0;
  p.ram_force = np[GAL_NP_RAM_PRESSURE_FORCE * (size_t)ld + i];
  p.v_host = np[GAL_NP_HOST_VELOCITY_VIRIAL * (size_t)ld + i];
  return p;
}

// isNonNegative attributes of the state variables (component _class.F90 XML).
__device__ __forceinline__ bool var_non_negative(int v) {
  return !(v == GAL_V_HOT_ANGULAR_MOMENTUM || v == GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM ||
           v == GAL_V_HOT_ENERGY_RADIATED || v == GAL_V_BH_RADIAL_POSITION || v == GAL_V_SPIN_ANGULAR_MOMENTUM);
}

// x^(n2/2) for integer n2 given sqrt(x); wave-uniform n2.
__device__ __forceinline__ double pow_half(double x, double sqrt_x, int n2) {
  const int m = n2 < 0 ? -n2 : n2;
  double r = (m & 1) ? sqrt_x : 1.0;
  double b = x;
  for (int k = m >> 1; k; k >>= 1) {
    if (k & 1) r *= b;
    b *= b;
  }
  return n2 < 0 ? rcp(r) : r;
}

// x^alpha: half-integer fast path, else exp(alpha ln x).
__device__ __forceinline__ double pow_param(double x, double sqrt_x, double ln_x_fn_arg, int n2, double alpha) {
  if (n2 != kNoHalfPower) return pow_half(x, sqrt_x, n2);
  return exp(alpha * ln_x_fn_arg);
}

// Bilinear interpolation of ln(Lambda) in (ln T, ln Z) with "fix"
// extrapolation: cieFileInterpolatingFactors / cieFileInterpolate
// (cooling/cooling_function/CIE_file.F90:665-735).  Grids are uniform, so
// the bracketing index is computed directly instead of by bisection.  The
// logarithms of the clamped temperature and metallicity are passed in so
// that they can be computed as early as their inputs are known.
__device__ __forceinline__ double clamp_temperature(const CoolingDesc& c, double temperature) {
  return fmin(fmax(temperature, c.t_min), c.t_max);
}
__device__ __forceinline__ double clamp_metallicity(const CoolingDesc& c, double z_solar) {
  return fmin(fmax(z_solar, c.z_min), c.z_max);
}
__device__ __forceinline__ double cooling_function_ln(const double* __restrict__ tab, const CoolingDesc& c, double lt,
                                                      double lz) {
  int it = (int)((lt - c.ln_t_min) * c.inv_d_ln_t);
  int iz = (int)((lz - c.ln_z_min) * c.inv_d_ln_z);
  it = it < 0 ? 0 : (it > c.n_t - 2 ? c.n_t - 2 : it);
  iz = iz < 0 ? 0 : (iz > c.n_z - 2 ? c.n_z - 2 : iz);
  const double t0 = c.ln_t_min + c.d_ln_t * it, t1 = c.ln_t_min + c.d_ln_t * (it + 1);
  const double z0 = c.ln_z_min + c.d_ln_z * iz, z1 = c.ln_z_min + c.d_ln_z * (iz + 1);
  const double ht = (lt - t0) / (t1 - t0);
  const double hz = (lz - z0) / (z1 - z0);
  const double* r0 = tab + iz * c.n_t + it;
  const double* r1 = r0 + c.n_t;
  const double v = r0[0] * (1.0 - ht) * (1.0 - hz) + r1[0] * (1.0 - ht) * hz + r0[1] * ht * (1.0 - hz) +
                   r1[1] * ht * hz;
  return exp(v);
}

// ISCO specific energy and angular momentum, prograde, gravitational units
// (black_holes/fundamentals.F90; Bardeen, Press & Teukolsky 1972).
__device__ __forceinline__ void isco(double a, double& energy, double& ang_mom) {
  const double cp = cbrt(1.0 + a), cm = cbrt(1.0 - a);
  const double a1 = 1.0 + cp * cm * (cp + cm);
  const double a2 = sqrt(3.0 * a * a + a1 * a1);
  // A1 <= 3 for |a| <= 1; rounding can push it just above 3 for tiny spins
  const double r = 3.0 + a2 - sqrt(fmax(3.0 - a1, 0.0) * (3.0 + a1 + 2.0 * a2));
  const double sr = sqrt(r);
  const double inv_den = rcp(r) * rsqrt(r * r - 3.0 * r + 2.0 * a * sr);
  energy = (r * r - 2.0 * r + a * sr) * inv_den;
  ang_mom = (r * r - 2.0 * a * sr + a * a) * sr * inv_den;
}

// Hernquist density of mass M with scale radius a at r (zero for
// non-positive arguments).
__device__ __forceinline__ double hernquist_density(double mass, double scale, double r) {
  const double rs = r + scale;
  const double rho = mass * scale * (1.0 / (2.0 * kPi)) * rcp(r * rs * rs * rs);
  return (mass > 0.0 && scale > 0.0 && r > 0.0) ? rho : 0.0;
}

// Halo quantities shared by the RHS, the scale rules and the post-step.
struct Halo {
  double mass, a3, r_vir, inv_r_vir, v_vir, inv_v, t_dyn, inv_t_dyn, lambda, r_gal;
  double r_vir_rate;   // darkMatterHaloScale%radiusVirialGrowthRate
};

__device__ __forceinline__ Halo halo(const double* __restrict__ P, const DerivedConstants& D, const NodePar& np,
                                     double t, double spin_j) {
  Halo h;
  const double mrate = np.satellite ? 0.0 : np.m_rate;
  h.mass = np.m_ref + mrate * (t - np.t_ref);
  // The virial scales use the mean density at the time the node was last
  // isolated (dark_matter_halos/scales/virial_density_contrast.F90:356-382):
  // a satellite's reference time is that time.
  const double t_iso = np.satellite ? np.t_ref : t;
  // a^3 = (Omega_m/Omega_L) sinh^2(x), sinh(x) = (e^x - e^-x)/2
  const double ex = exp(D.sinh_arg * t_iso);
  const double s = 0.5 * (ex - rcp(ex));
  h.a3 = D.a3_norm * s * s;
  const double om_a3 = D.om * rcp(h.a3);
  const double e2 = om_a3 + D.ol;
  const double inv_e2 = rcp(e2);
  const double om_a = om_a3 * inv_e2;
  const double x = om_a - 1.0;
  const double delta_c = 18.0 * kPi * kPi + 82.0 * x - 39.0 * x * x;
  h.r_vir = cbrt(D.rvir_coeff * h.mass * inv_e2 * rcp(delta_c));
  h.inv_r_vir = rcp(h.r_vir);
  const double u = kGravity * h.mass * h.inv_r_vir;
  h.inv_v = rsqrt(u);
  h.v_vir = u * h.inv_v;
  h.t_dyn = kMpcPerKmsToGyr * h.r_vir * h.inv_v;
  h.inv_t_dyn = h.v_vir * h.inv_r_vir * (1.0 / kMpcPerKmsToGyr);
  const double jscale = h.mass * h.v_vir * h.r_vir;
  h.lambda = jscale > 0.0 ? spin_j * rcp(jscale) : 0.0;
  h.r_gal = P[GAL_P_RADIUS_FACTOR] * h.r_vir * h.lambda;  // fixed.F90:318-325
  // radiusVirialGrowthRate (virial_density_contrast.F90:338-354):
  // (r_vir/3)(Mdot/M - dln(Delta_c E^2)/dt), dln E^2/dt = -3 H Omega_m(a),
  // dOmega_m(a)/dt = -3 H Omega_m(a)(1 - Omega_m(a)); zero for satellites.
  const double hub = D.h0_gyr * sqrt(e2);
  const double dx = -3.0 * hub * om_a * (1.0 - om_a);
  const double dln = (82.0 - 78.0 * x) * dx * rcp(delta_c) - 3.0 * hub * om_a;
  h.r_vir_rate = np.satellite ? 0.0 : h.r_vir * (1.0 / 3.0) * (mrate * rcp(h.mass) - dln);
  return h;
}

// ---- the RHS -------------------------------------------------------------
// Written as straight-line dataflow: every process rate is computed and
// then selected by its applicability condition, so that the independent
// chains (halo scales, ISCO, metallicity, star formation, hot-halo
// structure) can be interleaved by the scheduler.  A rate that does not
// apply is exactly zero, which leaves the accumulated sums unchanged.
//
// The operators fall into two halves that share only the halo scales, the
// hot halo's density profile and the state: the galaxy half (spin growth,
// satellite mass loss, star formation, stellar feedback, bar instability,
// CGM accretion, outflow reincorporation and the CGM outer radius) and the
// hot-halo half (ISCO, cooling function and radius, black-hole accretion,
// jet and winds, cooling and heating, radiated energy).  The outer radius
// needs only the density profile, and sits in the galaxy half because the
// hot-halo half is the longer of the two.  rhs_part<PART> computes one
// half (or both); every path forms the derivative as f_galaxy + f_halo, so a
// wave that computes one half while another wave computes the other gets the
// same bits as one lane computing both.
enum : int { kRhsGalaxy = 1, kRhsHalo = 2 };

template <int PART>
__device__ __forceinline__ void rhs_part(const double* __restrict__ P, const DerivedConstants& D,
                                         const double* __restrict__ tab, const CoolingDesc& cd, int mode,
                                         const NodePar& np, double t, const double (&y)[kNV], double (&f)[kNV]) {
  constexpr bool GAL = (PART & kRhsGalaxy) != 0, HOT = (PART & kRhsHalo) != 0;
#pragma unroll
  for (int i = 0; i < kNV; ++i) f[i] = 0.0;
  if (mode == GAL_PHYSICS_TEST) {
    // source/tests/ODE_solver/functions.F90:37-80
    if constexpr (GAL) {
      f[0] = sin(t);
      f[1] = y[2];
      f[2] = -1.0 * y[1];
    }
    return;
  }
  const bool box = mode == GAL_PHYSICS_BOX;
  // closed-box mode has no hot halo, black hole or CGM
  if constexpr (!GAL) {
    if (box) return;
  }
  // ---- chains that do not depend on the halo scales ---------------------
  const double m_bh = y[GAL_V_BH_MASS];
  const double spin = fmin(fmax(y[GAL_V_BH_SPIN], 0.0), 0.998);
  double e_isco, l_isco;
  isco(spin, e_isco, l_isco);
  const double m_hot = y[GAL_V_HOT_MASS], r_out = y[GAL_V_HOT_OUTER_RADIUS];
  const bool hot_ok = m_hot > 0.0 && r_out > 0.0;
  // Reciprocal masses for the ratios of a reservoir (metallicity, specific
  // angular momentum).  Below kRatioMassMin the ratio is taken as zero: rcp()
  // of a subnormal mass overflows (NaN after the Newton steps), and a
  // reservoir that enforceNonNegativity lets decay towards zero by halved
  // steps reaches subnormal values; every flow such a ratio multiplies scales
  // with the reservoir itself, so the flows dropped are below 1e-100 Msun/Gyr.
  constexpr double kRatioMassMin = 1.0e-100;
  const double inv_m_hot = (hot_ok && m_hot > kRatioMassMin) ? rcp(m_hot) : 0.0;
  const double zf = fmin(fmax(y[GAL_V_HOT_ABUNDANCES] * inv_m_hot, 0.0), 1.0);
  const double lz = log(clamp_metallicity(cd, zf * (1.0 / kMetallicitySolar)));
  const double mg_d = y[GAL_V_DISK_MASS_GAS], ms_d = y[GAL_V_DISK_MASS_STELLAR];
  const double j_d = y[GAL_V_DISK_ANGULAR_MOMENTUM], zg_d = y[GAL_V_DISK_ABUNDANCES_GAS];
  const double mg_s = y[GAL_V_SPHEROID_MASS_GAS], ms_s = y[GAL_V_SPHEROID_MASS_STELLAR];
  const double j_s = y[GAL_V_SPHEROID_ANGULAR_MOMENTUM], zg_s = y[GAL_V_SPHEROID_ABUNDANCES_GAS];
  const double inv_mg_d = mg_d > kRatioMassMin ? rcp(mg_d) : 0.0;
  const double inv_mg_s = mg_s > kRatioMassMin ? rcp(mg_s) : 0.0;
  const double mt_d = mg_d + ms_d, mt_s = mg_s + ms_s;
  const double inv_mt_d = mt_d > kRatioMassMin ? rcp(mt_d) : 0.0;
  const double inv_mt_s = mt_s > kRatioMassMin ? rcp(mt_s) : 0.0;

  // ---- halo analytics ----------------------------------------------------
  const Halo H = halo(P, D, np, t, y[GAL_V_SPIN_ANGULAR_MOMENTUM]);
  const double v_gal = H.v_vir, r_gal = H.r_gal;
  const double inv_sqrt_v = rsqrt(v_gal);
  const double sqrt_v = v_gal * inv_sqrt_v;
  const bool generic_pow = D.n2_sf_disk == kNoHalfPower || D.n2_sf_sph == kNoHalfPower ||
                           D.n2_fb_disk == kNoHalfPower || D.n2_fb_sph == kNoHalfPower;
  const double ln_v = generic_pow ? log(v_gal) : 0.0;
  const double tdyn_gal = kMpcPerKmsToGyr * r_gal * H.inv_v;
  const double inv_tdyn_gal = H.v_vir * rcp(r_gal) * (1.0 / kMpcPerKmsToGyr);
  const double t_vir = D.t_vir_per_v2 * H.v_vir * H.v_vir;
  const double lt = log(clamp_temperature(cd, t_vir));

  // ---- hot-halo structure (both halves: the galaxy half strips it) ------
  // beta=2/3 profile, core coreRadiusOverVirialRadius*r_vir, truncated at
  // outerRadius (hot_halo/mass_distribution/beta_profile.F90); virial
  // (isothermal) temperature (hot_halo/temperature_profile/virial.F90).
  const double r_core = P[GAL_P_CORE_RADIUS_OVER_VIRIAL] * H.r_vir;
  const double inv_r_core = H.inv_r_vir * D.inv_core;
  const double xo = r_out * inv_r_core;
  // shape integrals S = x - atan x and N = (x^2 - ln(1 + x^2))/2, by their
  // Taylor series below x = 0.01 where the differences cancel
  const double uo = xo * xo;
  const bool small_xo = xo < 0.01;
  const double shape_m = small_xo ? xo * uo * (1.0 / 3.0 - uo * (1.0 / 5.0 - uo * (1.0 / 7.0 - uo * (1.0 / 9.0))))
                                  : xo - atan(xo);
  const double inv_shape = hot_ok ? rcp(shape_m) : 0.0;
  const double rho0 = m_hot * inv_shape * (1.0 / (4.0 * kPi)) * inv_r_core * inv_r_core * inv_r_core;
  const double rho_out = rho0 * rcp(1.0 + xo * xo);

  const int ops = box ? 0 : (int)P[GAL_P_OPERATORS];
  if constexpr (GAL) {
  // spin%angularMomentum growth (halo_angular_momentum_interpolate.F90)
  f[GAL_V_SPIN_ANGULAR_MOMENTUM] += np.satellite ? 0.0 : np.spin_rate;
  // satelliteMassLoss (satellite/mass_loss/_class.F90:176-200): an isolated
  // halo's bound mass follows its basic mass; darkMatterHaloMassLossRate is
  // the default zero class for satellites.
  f[GAL_V_SATELLITE_BOUND_MASS] += ((ops & GAL_OPERATOR_SATELLITE_MASS_LOSS) && !np.satellite) ? np.m_rate : 0.0;

  const double R = P[GAL_P_RECYCLED_FRACTION], yld = P[GAL_P_METAL_YIELD];
  // ---- starFormationDisks (star_formation/disks.F90:200-325) -----------
  // dynamicalTimeTimescale (star_formation/timescales/dynamical_time.F90:164-188),
  // timescaleRate (star_formation/rates/disks/timescale.F90:109-130),
  // instantaneousRates (stellar_populations/properties/instantaneous.F90:133-182)
  const bool disk_ok = !(j_d < 0.0 || r_gal < 0.0 || mg_d < 0.0);
  const double x_sf = v_gal * D.inv_200, sqrt_x_sf = sqrt_v * D.inv_sqrt_200;
  double tau_d;
  {
    const double eff = P[GAL_P_SF_DISK_EFFICIENCY];
    const double pw = pow_param(x_sf, sqrt_x_sf, ln_v - D.ln_200, D.n2_sf_disk, P[GAL_P_SF_DISK_EXPONENT_VELOCITY]);
    tau_d = (v_gal <= 0.0 || eff == 0.0) ? 0.0 : fmax(tdyn_gal * pw * D.inv_eff_sf_disk, P[GAL_P_SF_DISK_TIMESCALE_MINIMUM]);
    if (box) tau_d = P[GAL_P_BOX_TIMESCALE];  // starFormationTimescale fixed
  }
  const double sfr_d = (disk_ok && tau_d > 0.0) ? mg_d * rcp(tau_d) : 0.0;
  {
    const double zfuel = zg_d * inv_mg_d;
    const double rs = (1.0 - R) * sfr_d;
    f[GAL_V_DISK_MASS_STELLAR] += rs;
    f[GAL_V_DISK_MASS_GAS] += -rs;
    f[GAL_V_DISK_ABUNDANCES_STELLAR] += zfuel * rs;
    f[GAL_V_DISK_ABUNDANCES_GAS] += -(zfuel * rs) + yld * sfr_d;
    f[GAL_V_DISK_MASS_STELLAR_FORMED] += sfr_d;
  }
  // ---- starFormationSpheroids (star_formation/spheroids.F90:152) --------
  const bool sph_ok = !(j_s < 0.0 || r_gal < 0.0 || mg_s < 0.0);
  double tau_s;
  {
    const double eff = P[GAL_P_SF_SPHEROID_EFFICIENCY];
    const double pw = pow_param(x_sf, sqrt_x_sf, ln_v - D.ln_200, D.n2_sf_sph, P[GAL_P_SF_SPHEROID_EXPONENT_VELOCITY]);
    tau_s = (v_gal <= 0.0 || eff == 0.0) ? 0.0 : fmax(tdyn_gal * pw * D.inv_eff_sf_sph, P[GAL_P_SF_SPHEROID_TIMESCALE_MINIMUM]);
    if (box) tau_s = P[GAL_P_BOX_TIMESCALE];  // starFormationTimescale fixed
  }
  const double sfr_s = (sph_ok && tau_s > 0.0) ? mg_s * rcp(tau_s) : 0.0;
  {
    const double zfuel = zg_s * inv_mg_s;
    const double rs = (1.0 - R) * sfr_s;
    f[GAL_V_SPHEROID_MASS_STELLAR] += rs;
    f[GAL_V_SPHEROID_MASS_GAS] += -rs;
    f[GAL_V_SPHEROID_ABUNDANCES_STELLAR] += zfuel * rs;
    f[GAL_V_SPHEROID_ABUNDANCES_GAS] += -(zfuel * rs) + yld * sfr_s;
    f[GAL_V_SPHEROID_MASS_STELLAR_FORMED] += sfr_s;
  }
  // ---- stellarFeedbackDisks / Spheroids (stellar_feedback/disks.F90:116-206,
  // spheroids.F90:116): powerLaw (outflows/power_law/_class.F90:118-164)
  // limited by rateLimit (outflows/rate_limit.F90:111-170); the outflow joins
  // the outflowed reservoir (hot_halo/standard/_class.F90:598-626,
  // fractionStripped = 0).
  const double inv_tdyn_fb = (v_gal <= 0.0 || r_gal <= 0.0) ? 1.0 : inv_tdyn_gal;
  double out_d, out_s;
  {
    const double w = D.v_fb_disk * H.inv_v;
    double o = pow_param(w, D.sqrt_v_fb_disk * inv_sqrt_v, D.ln_v_fb_disk - ln_v, D.n2_fb_disk,
                         P[GAL_P_FB_DISK_EXPONENT]) * sfr_d;
    o = v_gal <= 0.0 ? 0.0 : o;
    if (box) o = P[GAL_P_BOX_OUTFLOW_FRACTION] * sfr_d;  // stellarFeedbackOutflows fixed
    const double omax = fmax(mg_d * inv_tdyn_fb * D.inv_tmin_fb_disk, 0.0);
    o = o > omax ? omax : o;
    out_d = (disk_ok && o > 0.0) ? o : 0.0;
  }
  {
    const double w = D.v_fb_sph * H.inv_v;
    double o = pow_param(w, D.sqrt_v_fb_sph * inv_sqrt_v, D.ln_v_fb_sph - ln_v, D.n2_fb_sph,
                         P[GAL_P_FB_SPHEROID_EXPONENT]) * sfr_s;
    o = v_gal <= 0.0 ? 0.0 : o;
    if (box) o = P[GAL_P_BOX_OUTFLOW_FRACTION_SPHEROID] * sfr_s;  // stellarFeedbackOutflows fixed
    const double omax = fmax(mg_s * inv_tdyn_fb * D.inv_tmin_fb_sph, 0.0);
    o = o > omax ? omax : o;
    out_s = (sph_ok && o > 0.0) ? o : 0.0;
  }
  // Outflowing gas joins the outflowed reservoir (hot_halo/standard/_class.F90:598-720):
  // the stripped fraction is zero for centrals and for satellites within the
  // virial radius (outflow_stripping/standard.F90:133-173), and the angular
  // momentum is raised by 1/(1 - fractionLossAngularMomentum).
  {
    const double rj = j_d * (out_d * inv_mt_d);
    const double rz = (zg_d * inv_mg_d) * out_d;
    f[GAL_V_HOT_OUTFLOWED_MASS] += out_d;
    f[GAL_V_DISK_MASS_GAS] -= out_d;
    f[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] += rj * D.inv_one_minus_j_loss;
    f[GAL_V_DISK_ANGULAR_MOMENTUM] -= rj;
    f[GAL_V_HOT_OUTFLOWED_ABUNDANCES] += rz;
    f[GAL_V_DISK_ABUNDANCES_GAS] -= rz;
  }
  {
    const double rj = j_s * (out_s * inv_mt_s);
    const double rz = (zg_s * inv_mg_s) * out_s;
    f[GAL_V_HOT_OUTFLOWED_MASS] += out_s;
    f[GAL_V_SPHEROID_MASS_GAS] -= out_s;
    f[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] += rj * D.inv_one_minus_j_loss;
    f[GAL_V_SPHEROID_ANGULAR_MOMENTUM] -= rj;
    f[GAL_V_HOT_OUTFLOWED_ABUNDANCES] += rz;
    f[GAL_V_SPHEROID_ABUNDANCES_GAS] -= rz;
  }
  if (box) return;
  // ---- barInstability (bar_instability.F90:145-248) with efstathiou1982
  // (galactic_dynamics/bar_instability/Efstathiou1982.F90:153-258): unstable
  // when eps = 1.18 V / sqrt(G M_disk / r_disk) is below the threshold
  // interpolated in gas fraction; material moves to the spheroid on
  // t_dyn ((eps_c - eps_iso)/(eps_c - eps))^2.  A disk of zero mass gives a
  // NaN gas fraction and, as in the reference, no instability.
  {
    constexpr double kEpsIso = 0.6221297315, kBoost = 1.1800237580, kDimMax = 1.0e10;
    const double fgas = mg_d * rcp(mt_d);
    const double thr = P[GAL_P_BAR_THRESHOLD_STELLAR] * (1.0 - fgas) + P[GAL_P_BAR_THRESHOLD_GASEOUS] * fgas;
    const double v_self2 = kGravity * mt_d * rcp(r_gal);
    double est = fmax(kEpsIso, kBoost * v_gal * rsqrt(v_self2));
    est = (mt_d >= 0.0 && v_self2 > 0.0) ? est : HUGE_VAL;
    const bool unstable = (ops & GAL_OPERATOR_BAR_INSTABILITY) && j_d > 0.0 && r_gal > 0.0 && v_gal > 0.0 &&
                          mg_d >= 0.0 && est < thr;
    constexpr double kC = kSpeedLight / kKilo;
    const double t_dyn = v_gal < kC ? tdyn_gal : kMpcPerKmsToGyr * r_gal * (1.0 / kC);
    const double iso_rel = thr - kEpsIso, est_rel = thr - est;
    const double ratio = iso_rel * rcp(est_rel);
    const double dimless = iso_rel > kDimMax * est_rel ? kDimMax : ratio * ratio;
    const double inv_tau = rcp(fmax(t_dyn, P[GAL_P_BAR_TIMESCALE_MINIMUM]) * dimless);
    const double rg = unstable ? fmax(0.0, mg_d) * inv_tau : 0.0;
    const double rs = unstable ? fmax(0.0, ms_d) * inv_tau : 0.0;
    const double rj = unstable ? fmax(0.0, j_d) * inv_tau : 0.0;
    const double rzg = unstable ? fmax(0.0, zg_d) * inv_tau : 0.0;
    const double rzs = unstable ? fmax(0.0, y[GAL_V_DISK_ABUNDANCES_STELLAR]) * inv_tau : 0.0;
    f[GAL_V_DISK_MASS_GAS] -= rg;
    f[GAL_V_SPHEROID_MASS_GAS] += rg;
    f[GAL_V_DISK_MASS_STELLAR] -= rs;
    f[GAL_V_SPHEROID_MASS_STELLAR] += rs;
    f[GAL_V_DISK_ANGULAR_MOMENTUM] -= (1.0 - P[GAL_P_BAR_J_RETAINED_DISK]) * rj;
    f[GAL_V_SPHEROID_ANGULAR_MOMENTUM] += P[GAL_P_BAR_J_RETAINED_SPHEROID] * rj;
    f[GAL_V_DISK_ABUNDANCES_GAS] -= rzg;
    f[GAL_V_SPHEROID_ABUNDANCES_GAS] += rzg;
    f[GAL_V_DISK_ABUNDANCES_STELLAR] -= rzs;
    f[GAL_V_SPHEROID_ABUNDANCES_STELLAR] += rzs;
  }
  // ---- CGMAccretion (circumgalactic_medium/accretion.F90:517) with
  //      accretionHalo simple (accretion/halo/simple.F90:281-378) ---------
  {
    const bool acc = !np.satellite && H.mass > 0.0;
    const double failed = (H.a3 > D.a3_reion && H.v_vir < P[GAL_P_VELOCITY_SUPPRESSION_REIONIZATION]) ? 1.0 : 0.0;
    const double inv_mass = rcp(H.mass);
    const double growth = np.m_rate * inv_mass;
    const double m_un = y[GAL_V_HOT_UNACCRETED_MASS], z_un = y[GAL_V_HOT_UNACCRETED_ABUNDANCES];
    const double rate_acc = D.f_baryon * np.m_rate * (1.0 - failed) + m_un * growth * (1.0 - failed);
    const double rate_fail = D.f_baryon * np.m_rate * failed - m_un * growth * (1.0 - failed);
    const double rate_z = z_un * growth * (1.0 - failed);
    f[GAL_V_HOT_MASS] += acc ? rate_acc : 0.0;
    f[GAL_V_HOT_UNACCRETED_MASS] += acc ? rate_fail : 0.0;
    f[GAL_V_HOT_ABUNDANCES] += acc ? rate_z : 0.0;
    f[GAL_V_HOT_UNACCRETED_ABUNDANCES] -= acc ? rate_z : 0.0;
    f[GAL_V_HOT_ANGULAR_MOMENTUM] += acc ? rate_acc * (y[GAL_V_SPIN_ANGULAR_MOMENTUM] * inv_mass) : 0.0;
  }
  // ---- CGMOutflowReincorporation (outflow_reincorporation.F90:272-330)
  //      haloDynamicalTimeRate (halo_dynamical_time.F90:113-125) ----------
  {
    const double m_o = y[GAL_V_HOT_OUTFLOWED_MASS];
    const bool ok = m_o > 0.0;
    const double q = P[GAL_P_REINCORPORATION_MULTIPLIER] * H.inv_t_dyn;
    const double r = m_o * q;
    const double rr = ok ? r : 0.0;
    const double rj = ok ? y[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] * q : 0.0;
    const double rz = ok ? y[GAL_V_HOT_OUTFLOWED_ABUNDANCES] * q : 0.0;
    f[GAL_V_HOT_OUTFLOWED_MASS] -= rr;
    f[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] -= rj;
    f[GAL_V_HOT_OUTFLOWED_ABUNDANCES] -= rz;
    f[GAL_V_HOT_MASS] += rr;
    f[GAL_V_HOT_ANGULAR_MOMENTUM] += rj;
    f[GAL_V_HOT_ABUNDANCES] += rz;
  }
  // ---- CGMOuterRadiusRamPressureStripping
  //      (circumgalactic_medium/outer_radius/ram_pressure_stripping.F90:111-242)
  // Returning outflows push the outer radius out when it lies inside the
  // virial radius; a central's follows the virial radius; a satellite's
  // shrinks towards the font2008 stripping radius
  // (hot_halo/ram_pressure_stripping/Font2008.F90:115-267) on the
  // ram-pressure acceleration time (timescale/ram_pressure_acceleration.F90:79-134).
  // The stripping radius is the closed form of the binding condition
  // formFactor G M(<r) rho_hot(r) / r = F for the isothermal halo
  // M(<r) = M_vir r / r_vir: formFactor V_vir^2 rho_hot(r) = F.
  if (ops & GAL_OPERATOR_CGM_OUTER_RADIUS) {
    const double rv = H.r_vir;
    const double rho_ro = hot_ok ? rho_out : 0.0;
    const double m_return = y[GAL_V_HOT_OUTFLOWED_MASS] * P[GAL_P_REINCORPORATION_MULTIPLIER] * H.inv_t_dyn;
    const double rho_min = D.f_baryon * H.mass * (H.inv_r_vir * H.inv_r_vir * H.inv_r_vir) * (1.0 / (4.0 * kPi));
    const double ret_a = m_return * (1.0 / (4.0 * kPi)) * rcp(r_out * r_out * fmax(rho_ro, rho_min));
    const double ret_b = m_return * rcp(H.mass) * rv;
    double ret = (r_out > 0.0 && rho_ro > 0.0) ? ret_a : (m_return > 0.0 ? ret_b : 0.0);
    // the outer radius starts at the virial radius, an exact tie in the
    // reference; a last-place difference between the caller's value and the
    // recomputed virial radius is judged a tie (1e-12 relative)
    const bool at_rv = r_out >= rv * (1.0 - 1.0e-12);
    ret = at_rv ? 0.0 : ret;
    // stripping radius and timescale
    const double force = np.ram_force;
    const double bind = P[GAL_P_RAM_PRESSURE_FORM_FACTOR] * H.v_vir * H.v_vir;
    const double x_v = rv * inv_r_core, r_small = 1.0e-6 * rv, x_s = r_small * inv_r_core;
    const double rho_rv = (hot_ok && at_rv) ? rho0 * rcp(1.0 + x_v * x_v) : 0.0;
    const double rho_small = (hot_ok && r_small <= r_out) ? rho0 * rcp(1.0 + x_s * x_s) : 0.0;
    const double r_star = fmin(r_core * sqrt(fmax(bind * rho0 * rcp(force) - 1.0, 0.0)), r_out);
    const double r_rp = bind * rho_rv - force >= 0.0 ? rv : (bind * rho_small - force <= 0.0 ? 0.0 : r_star);
    const double sigma = r_out * rho_ro;
    const double t_floor = np.v_host > 0.0 ? r_out * 0.1 * rcp(np.v_host) : HUGE_VAL;
    double tau = kMpcPerKmsToGyr * fmax(sqrt(2.0 * r_out * sigma * rcp(force)), t_floor);
    tau = (r_out <= 0.0 || sigma <= 0.0) ? H.t_dyn : tau;
    tau = force <= 0.0 ? kDoubleMax : tau;
    const double dr = (r_rp < r_out && y[GAL_V_HOT_ANGULAR_MOMENTUM] > 0.0) ? (r_rp - r_out) / tau : 0.0;
    const bool strip = np.satellite && dr != 0.0 && m_hot > 0.0 && r_out > 1.0e-3 * rv;
    const double loss = strip ? 4.0 * kPi * rho_ro * r_out * r_out * dr : 0.0;
    const double lm = loss * inv_m_hot;
    f[GAL_V_HOT_OUTER_RADIUS] += ret + (np.satellite ? (strip ? dr : 0.0) : H.r_vir_rate);
    f[GAL_V_HOT_MASS] += loss;
    f[GAL_V_HOT_ANGULAR_MOMENTUM] += y[GAL_V_HOT_ANGULAR_MOMENTUM] * lm;
    f[GAL_V_HOT_ABUNDANCES] += y[GAL_V_HOT_ABUNDANCES] * lm;
    f[GAL_V_HOT_STRIPPED_MASS] -= loss;
    f[GAL_V_HOT_STRIPPED_ABUNDANCES] -= y[GAL_V_HOT_ABUNDANCES] * lm;
  }
  }  // galaxy half
  if constexpr (!HOT) return;
  // ---- cooling radius ------------------------------------------------------
  const double shape_r = small_xo ? uo * uo * (1.0 / 4.0 - uo * (1.0 / 6.0 - uo * (1.0 / 8.0 - uo * (1.0 / 10.0))))
                                  : 0.5 * uo - 0.5 * log1p(uo);
  const double r_mean = r_core * shape_r * inv_shape;
  // composition of metallicity-only abundances, fully ionized H and He
  constexpr double kNorm = 1.0 / (kHydrogenPrimordial + kHeliumPrimordial);
  const double X = (1.0 - zf) * kHydrogenPrimordial * kNorm;
  const double Yh = (1.0 - zf) * kHeliumPrimordial * kNorm;
  const double nh_n = X / kAtomicMassHydrogen, nhe_n = Yh / kAtomicMassHelium;
  const double nh_per_rho = X * D.nh_coeff;
  const double inv_nh_n = rcp(nh_n);
  const double nall_per_nh = (nh_n + nhe_n) * inv_nh_n + (1.0 + 2.0 * nhe_n * inv_nh_n);
  const double lambda_cool = cooling_function_ln(tab, cd, lt, lz);
  // simpleTime (cooling_time/simple.F90:128-181): t_cool = K / rho;
  // simpleRadius (cooling_radius/simple.F90:313-387) with t_available = t_dyn
  // (time_available/White-Frenk.F90:144-146, ageFactor = 0)
  const double t_avail = H.t_dyn;
  const double K = D.cool_k_pre * t_vir * nall_per_nh * rcp(lambda_cool * nh_per_rho);
  const double inv_k = rcp(K);
  double r_cool = fmin(r_core * sqrt(fmax(rho0 * t_avail * inv_k - 1.0, 0.0)), r_out);
  r_cool = (K - t_avail * rho0 > 0.0) ? 0.0 : r_cool;
  r_cool = (K - t_avail * rho_out < 0.0) ? r_out : r_cool;
  r_cool = (lambda_cool > 0.0 && hot_ok) ? r_cool : 0.0;
  // simpleRadiusGrowthRate (cooling_radius/simple.F90:229-309); the
  // time-available growth rate is 1 (White-Frenk.F90:161-168)
  const double xc = r_cool * inv_r_core;
  const double slope = 2.0 * xc * xc / (1.0 + xc * xc);
  double r_cool_rate = r_cool * H.inv_t_dyn * rcp(slope);
  r_cool_rate = (r_cool < r_out && r_cool > 0.0 && slope != 0.0) ? r_cool_rate : 0.0;

  // ---- blackHolesAccretion (black_holes/accretion.F90:113-170) ----------
  // rates: accretion_rates/standard.F90:236-440; Bondi-Hoyle-Lyttleton:
  // accretion/Bondi_Hoyle_Lyttleton.F90:34-78; thin disk:
  // accretion_disks/Shakura_Sunyaev.F90:70-179.
  const bool bh_ok = m_bh > 0.0;
  const double eff_thin = 1.0 - e_isco;
  const double edd_cap = D.m_edd_per_mass * m_bh * rcp(eff_thin);
  constexpr double kBhlUnit = kKilo * kGigaYear / kMegaParsec * 4.0 * kPi;
  double mdot_sph;
  {
    double racc = P[GAL_P_BH_TEMPERATURE_SPHEROID] > 0.0 ? kGravity * m_bh * D.inv_cs2_sph : HUGE_VAL;
    racc = fmax(racc, y[GAL_V_BH_RADIAL_POSITION]);
    const double rho_a = hernquist_density(mg_s, r_gal, racc);
    const double lj = fmin(rsqrt(D.jeans_coeff * rho_a), r_gal);
    const double rho_j = hernquist_density(mg_s, r_gal, lj);
    const double rho = lj > racc ? rho_j : rho_a;
    double m = fmax(D.bhl_sph_coeff * m_bh * m_bh * rho, 0.0);
    m = (m > 0.0 && eff_thin > 0.0) ? fmin(m, edd_cap) : m;
    mdot_sph = (bh_ok && rho_a > 1.0) ? m : 0.0;
  }
  double mdot_hot;
  {
    const double cs2 = D.cs2_per_t * t_vir;
    const double cs = sqrt(cs2);
    const double r2 = fmin(kGravity * m_bh * rcp(cs2), r_out);
    const double q = r_cool * H.inv_r_vir;
    // standardHotModeFraction (accretion_rates/standard.F90:442-466)
    const double xq = (q - 0.9) * 10.0;
    double fhot = xq * xq * (2.0 * xq - 3.0) + 1.0;
    fhot = q < 0.9 ? 1.0 : (q > 1.0 ? 0.0 : fhot);
    const double rho2 = fhot * (rho0 * rcp(1.0 + sq(r2 * inv_r_core)));
    double m = fmax(P[GAL_P_BH_ENHANCEMENT_HOT_HALO] * kBhlUnit * r2 * r2 * rho2 * cs, 0.0);
    m = fmin(m, fmax(m_hot * cs * rcp(r_out) * (1.0 / kMpcPerKmsToGyr), 0.0));
    m = (m > 0.0 && eff_thin > 0.0) ? fmin(m, edd_cap) : m;
    mdot_hot = (bh_ok && hot_ok && rho2 > 1.0) ? m : 0.0;
  }
  const double mdot = mdot_sph + mdot_hot;
  const bool acc_ok = bh_ok && mdot > 0.0;
  // Meier (2001) jet power (Shakura_Sunyaev.F90:92-140):
  // M^-0.3 Mdot^1.2 = (Mdot^4 / M)^0.3
  const double inv_m_bh = rcp(m_bh);
  const double lm = 0.3 * log(mdot * mdot * mdot * mdot * inv_m_bh);
  const double p_jet_k = exp(D.ln_jet_kerr + lm) * (1.0 + 1.1 * spin + 0.29 * spin * spin);
  const double p_jet_s = exp(D.ln_jet_schw + lm + 3.785 * spin);
  const double p_jet = acc_ok ? (spin > 0.8 ? p_jet_k : p_jet_s) : 0.0;
  {
    constexpr double kInvC2 = 1.0 / ((kSpeedLight / kKilo) * (kSpeedLight / kKilo));
    const double eff_jet = p_jet * rcp(mdot) * kInvC2;
    const double reduced = mdot * (1.0 - eff_thin - eff_jet);
    const double spin_up = (l_isco - 2.0 * spin * e_isco) * mdot * inv_m_bh;
    f[GAL_V_BH_MASS] += acc_ok ? reduced : 0.0;
    // spheroid%massGasSinkRate: gas, metals and angular momentum in proportion
    f[GAL_V_SPHEROID_MASS_GAS] -= acc_ok ? mdot_sph : 0.0;
    f[GAL_V_SPHEROID_ABUNDANCES_GAS] -= acc_ok ? mdot_sph * (zg_s * inv_mg_s) : 0.0;
    f[GAL_V_SPHEROID_ANGULAR_MOMENTUM] -= acc_ok ? mdot_sph * (j_s * inv_mt_s) : 0.0;
    // hotHalo%massSinkRate: mass, metals and angular momentum in proportion
    const bool hs = acc_ok && hot_ok;
    f[GAL_V_HOT_MASS] -= hs ? mdot_hot : 0.0;
    f[GAL_V_HOT_ABUNDANCES] -= hs ? mdot_hot * (y[GAL_V_HOT_ABUNDANCES] * inv_m_hot) : 0.0;
    f[GAL_V_HOT_ANGULAR_MOMENTUM] -= hs ? mdot_hot * (y[GAL_V_HOT_ANGULAR_MOMENTUM] * inv_m_hot) : 0.0;
    f[GAL_V_BH_SPIN] += acc_ok ? spin_up : 0.0;
  }
  // ---- blackHolesWinds (black_holes/winds.F90:102-133) with ciotti2009
  // (black_holes/winds/Ciotti2009.F90:113-193): the wind power enters the
  // spheroid gas, which drives an outflow efficiencyEnergeticOutflow P / V^2
  // into the hot halo (spheroid/standard/_class.F90:673-725).
  {
    const double eff = P[GAL_P_BH_WIND_EFFICIENCY] * (P[GAL_P_BH_WIND_SCALES_WITH_RADIATIVE] != 0.0 ? eff_thin : 1.0);
    const bool on = (ops & GAL_OPERATOR_BLACK_HOLE_WINDS) && acc_ok && eff > 0.0;
    const double x = D.wind_pressure_ratio * mg_s * rcp(r_gal * eff * mdot) - 0.5;
    double coupled = x <= 0.0 ? 0.0 : (x >= 1.0 ? 1.0 : x * x * (3.0 - 2.0 * x));
    coupled = (mg_s > 0.0 && r_gal > 0.0) ? coupled : 0.0;
    constexpr double kC2 = (kSpeedLight / kKilo) * (kSpeedLight / kKilo);
    const double power = eff * coupled * mdot * kC2;
    const bool drive = on && power > 0.0 && mg_s > 0.0 && mt_s > 0.0 && v_gal > 0.0;
    const double mo = drive ? P[GAL_P_SPHEROID_ENERGETIC_OUTFLOW] * power * H.inv_v * H.inv_v : 0.0;
    const double jo = mo * inv_mt_s * j_s;
    const double zo = mo * inv_mg_s * zg_s;
    f[GAL_V_SPHEROID_MASS_GAS] -= mo;
    f[GAL_V_SPHEROID_ANGULAR_MOMENTUM] -= jo;
    f[GAL_V_SPHEROID_ABUNDANCES_GAS] -= zo;
    f[GAL_V_HOT_OUTFLOWED_MASS] += mo;
    f[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] += jo * D.inv_one_minus_j_loss;
    f[GAL_V_HOT_OUTFLOWED_ABUNDANCES] += zo;
  }
  // ---- CGMCoolingHeating (circumgalactic_medium/cooling_heating.F90:216-381)
  //      whiteFrenk1991Rate (cooling_rate/White-Frenk.F90:132-185) ---------
  {
    const bool ok = hot_ok && y[GAL_V_HOT_ANGULAR_MOMENTUM] > 0.0;
    double cool = 4.0 * kPi * r_cool * r_cool * (rho0 * rcp(1.0 + xc * xc)) * r_cool_rate;
    cool = r_cool >= r_out ? m_hot * H.inv_t_dyn : cool;
    cool = H.v_vir > P[GAL_P_COOLING_VELOCITY_CUTOFF] ? 0.0 : cool;
    const double heat = P[GAL_P_EFFICIENCY_RADIO_MODE] * p_jet * sq(H.inv_v);
    const bool heating = ok && heat > cool;
    const bool cooling = ok && cool > heat;
    // excess heating expels gas from the node (excessHeatDrivesOutflow); a
    // satellite's joins its stripped reservoir (outflow_stripping/standard.F90:120-131)
    const bool expel = heating && P[GAL_P_EXCESS_HEAT_DRIVES_OUTFLOW] != 0.0;
    const double out = expel ? fmin(heat - cool, P[GAL_P_RATE_MAXIMUM_EXPULSION] * m_hot * H.inv_t_dyn) : 0.0;
    // net cooling onto the disk; constantRotation j(r) = (J/M) r / <r>
    const double c = cooling ? fmax(cool - heat, 0.0) : 0.0;
    const double j_hot = y[GAL_V_HOT_ANGULAR_MOMENTUM] * inv_m_hot;
    const double z_hot = y[GAL_V_HOT_ABUNDANCES] * inv_m_hot;
    const double oz = y[GAL_V_HOT_ABUNDANCES] * (out * inv_m_hot);
    const double oj = y[GAL_V_HOT_ANGULAR_MOMENTUM] * (out * inv_m_hot);
    const double cj = cooling ? c * (j_hot * r_cool * rcp(r_mean)) : 0.0;
    const double cz = c * z_hot;
    f[GAL_V_HOT_MASS] -= heating ? out : c;
    f[GAL_V_HOT_ABUNDANCES] -= heating ? oz : cz;
    f[GAL_V_HOT_ANGULAR_MOMENTUM] -= heating ? oj : cj;
    f[GAL_V_HOT_STRIPPED_MASS] += np.satellite ? out : 0.0;
    f[GAL_V_HOT_STRIPPED_ABUNDANCES] += np.satellite ? oz : 0.0;
    f[GAL_V_DISK_MASS_GAS] += c;
    f[GAL_V_DISK_ABUNDANCES_GAS] += cz;
    f[GAL_V_DISK_ANGULAR_MOMENTUM] += cj * (1.0 - P[GAL_P_FRACTION_LOSS_ANGULAR_MOMENTUM]);
  }
  // ---- coolingEnergyRadiated (cooling_energy_radiated.F90:200-260) -------
  {
    const double rho = m_hot * (3.0 / (4.0 * kPi)) * rcp(r_out * r_out * r_out);
    const double nh = rho * nh_per_rho;
    const double rcm = r_out * (kMegaParsec / kCenti);
    const double count = nh * nall_per_nh * (4.0 * kPi / 3.0) * rcm * rcm * rcm;
    f[GAL_V_HOT_ENERGY_RADIATED] += hot_ok ? lambda_cool * nh * count : 0.0;
  }
}

// The whole RHS: both halves, summed as f_galaxy + f_halo.
__device__ __forceinline__ void rhs(const double* __restrict__ P, const DerivedConstants& D,
                                    const double* __restrict__ tab, const CoolingDesc& cd, int mode,
                                    const NodePar& np, double t, const double (&y)[kNV], double (&f)[kNV]) {
  double fh[kNV];
  rhs_part<kRhsGalaxy>(P, D, tab, cd, mode, np, t, y, f);
  rhs_part<kRhsHalo>(P, D, tab, cd, mode, np, t, y, fh);
#pragma unroll
  for (int i = 0; i < kNV; ++i) f[i] += fh[i];
}

// Component scale-set rules at the start of an evolve (hot_halo/standard/_class.F90:804-840,
// disk/standard/_class.F90:735-828, spheroid/standard/_class.F90:804-860,
// black_hole/standard/_class.F90:195-240, cooling_energy_radiated.F90:150-185).
__device__ __forceinline__ void scales(const double* __restrict__ P, const DerivedConstants& D, const NodePar& np,
                                       double t, const double (&y)[kNV], double (&sc)[kNV]) {
  const Halo H = halo(P, D, np, t, y[GAL_V_SPIN_ANGULAR_MOMENTUM]);
  const double mass = H.mass, r_vir = H.r_vir, v_vir = H.v_vir;
  constexpr double mrel = 1.0e-3, rrel = 1.0e-1;
#pragma unroll
  for (int i = 0; i < kNV; ++i) sc[i] = 1.0;
  sc[GAL_V_HOT_MASS] = mass * mrel;
  sc[GAL_V_HOT_OUTFLOWED_MASS] = mass * mrel;
  sc[GAL_V_HOT_UNACCRETED_MASS] = mass * mrel;
  sc[GAL_V_HOT_ABUNDANCES] = mass * mrel;
  sc[GAL_V_HOT_UNACCRETED_ABUNDANCES] = mass * mrel;
  sc[GAL_V_HOT_OUTFLOWED_ABUNDANCES] = mass * mrel;
  sc[GAL_V_HOT_ANGULAR_MOMENTUM] = mass * r_vir * v_vir * mrel;
  sc[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] = mass * r_vir * v_vir * mrel;
  sc[GAL_V_HOT_OUTER_RADIUS] = r_vir * rrel;
  sc[GAL_V_HOT_STRIPPED_MASS] = mass * mrel;
  sc[GAL_V_HOT_STRIPPED_ABUNDANCES] = mass * mrel;
  sc[GAL_V_HOT_ENERGY_RADIATED] = D.unit_energy * mass * v_vir * v_vir * 1.0e-3;
  const double ang = fmax(fabs(y[GAL_V_DISK_ANGULAR_MOMENTUM]) + fabs(y[GAL_V_SPHEROID_ANGULAR_MOMENTUM]), 0.1);
  const double mcomp = fmax(fabs(y[GAL_V_DISK_MASS_GAS]) + fabs(y[GAL_V_SPHEROID_MASS_GAS]) +
                                fabs(y[GAL_V_DISK_MASS_STELLAR]) + fabs(y[GAL_V_SPHEROID_MASS_STELLAR]),
                            1.0);
  const double zmin = fmax(mcomp * P[GAL_P_TOLERANCE_RELATIVE_METALLICITY], 1.0);
  const double zsc = fmax(fabs(y[GAL_V_DISK_ABUNDANCES_GAS]) + fabs(y[GAL_V_DISK_ABUNDANCES_STELLAR]) +
                              fabs(y[GAL_V_SPHEROID_ABUNDANCES_GAS]) + fabs(y[GAL_V_SPHEROID_ABUNDANCES_STELLAR]),
                          zmin);
  sc[GAL_V_DISK_ANGULAR_MOMENTUM] = ang;
  sc[GAL_V_DISK_MASS_GAS] = mcomp;
  sc[GAL_V_DISK_MASS_STELLAR] = mcomp;
  sc[GAL_V_DISK_MASS_STELLAR_FORMED] = mcomp;
  sc[GAL_V_DISK_FRACTION_MASS_RETAINED] = 1.0e-4 * y[GAL_V_DISK_FRACTION_MASS_RETAINED];
  sc[GAL_V_DISK_ABUNDANCES_GAS] = zsc;
  sc[GAL_V_DISK_ABUNDANCES_STELLAR] = zsc;
  sc[GAL_V_SPHEROID_ANGULAR_MOMENTUM] = ang;
  sc[GAL_V_SPHEROID_MASS_GAS] = mcomp;
  sc[GAL_V_SPHEROID_MASS_STELLAR] = mcomp;
  sc[GAL_V_SPHEROID_MASS_STELLAR_FORMED] = mcomp;
  sc[GAL_V_SPHEROID_ABUNDANCES_GAS] = zsc;
  sc[GAL_V_SPHEROID_ABUNDANCES_STELLAR] = zsc;
  sc[GAL_V_BH_MASS] = fmax(fmax(1.0, 1.0e-4 * y[GAL_V_SPHEROID_MASS_STELLAR]), y[GAL_V_BH_MASS]);
  sc[GAL_V_BH_SPIN] = 1.0;
  sc[GAL_V_BH_RADIAL_POSITION] = fmax(1.0e-4 * H.r_gal, 1.0e-6);
  sc[GAL_V_SPIN_ANGULAR_MOMENTUM] = mass * r_vir * v_vir * mrel;
  sc[GAL_V_SATELLITE_BOUND_MASS] = mass * mrel;
}

// Post-step: standardPostStepProcessing (node_evolver/standard.F90:1160-1185)
// with the disk/spheroid (disk/standard/_class.F90:506-660) and hot-halo
// (hot_halo/standard/_class.F90:468-475) negative-value fixes.  The state
// change is a GSL_CONTINUE: no evolver reset.  The state is addressed as
// y[v*stride] so the same code serves a lane's LDS copy.
__device__ __forceinline__ void component_post_step(double* __restrict__ y, int stride, int i_ms, int i_zs,
                                                    int i_mg, int i_zg, int i_j, double rv) {
#define PY(v) y[(v) * stride]
  if (PY(i_mg) < 0.0) {
    const double m = PY(i_mg) + PY(i_ms);
    double j;
    if (m == 0.0) { j = 0.0; PY(i_ms) = 0.0; PY(i_zs) = 0.0; }
    else { j = PY(i_j) / m; if (j < 0.0) j = rv; }
    PY(i_mg) = 0.0; PY(i_zg) = 0.0; PY(i_j) = j * PY(i_ms);
  }
  if (PY(i_ms) < 0.0) {
    const double m = PY(i_mg) + PY(i_ms);
    double j;
    if (m == 0.0) { j = 0.0; PY(i_mg) = 0.0; PY(i_zg) = 0.0; }
    else { j = PY(i_j) / m; if (j < 0.0) j = rv; }
    PY(i_ms) = 0.0; PY(i_zs) = 0.0; PY(i_j) = j * PY(i_mg);
  }
  if (PY(i_j) < 0.0) PY(i_j) = 0.0;
#undef PY
}

// The rarely taken fix-up is kept out of line; its arguments are passed by
// value so that nothing of the caller's has to live in addressable memory.
struct HaloConsts {
  double om, ol, a3_norm, sinh_arg, rvir_coeff;
};

__device__ __noinline__ void post_step_fix(const double* __restrict__ P, HaloConsts hc, NodePar np, double t,
                                           double* __restrict__ y, int stride) {
  DerivedConstants D{};
  D.om = hc.om;
  D.ol = hc.ol;
  D.a3_norm = hc.a3_norm;
  D.sinh_arg = hc.sinh_arg;
  D.rvir_coeff = hc.rvir_coeff;
  const Halo H = halo(P, D, np, t, y[GAL_V_SPIN_ANGULAR_MOMENTUM * stride]);
  const double rv = H.r_gal * H.v_vir;
  if (P[GAL_P_POST_STEP_ZERO_NEGATIVE_MASSES] != 0.0) {
    component_post_step(y, stride, GAL_V_DISK_MASS_STELLAR, GAL_V_DISK_ABUNDANCES_STELLAR, GAL_V_DISK_MASS_GAS,
                        GAL_V_DISK_ABUNDANCES_GAS, GAL_V_DISK_ANGULAR_MOMENTUM, rv);
    component_post_step(y, stride, GAL_V_SPHEROID_MASS_STELLAR, GAL_V_SPHEROID_ABUNDANCES_STELLAR,
                        GAL_V_SPHEROID_MASS_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS, GAL_V_SPHEROID_ANGULAR_MOMENTUM, rv);
  }
  if (y[GAL_V_HOT_MASS * stride] < 0.0) y[GAL_V_HOT_MASS * stride] = 0.0;
  if (y[GAL_V_HOT_OUTER_RADIUS * stride] < 0.0) y[GAL_V_HOT_OUTER_RADIUS * stride] = 0.0;
}

__device__ __forceinline__ void post_step(const double* __restrict__ P, const DerivedConstants& D, const NodePar& np,
                                          int mode, double t, double* __restrict__ y, int stride) {
  if (mode == GAL_PHYSICS_TEST) return;
  const bool galaxy = P[GAL_P_POST_STEP_ZERO_NEGATIVE_MASSES] != 0.0;
  const bool neg = (galaxy && (y[GAL_V_DISK_MASS_GAS * stride] < 0.0 || y[GAL_V_DISK_MASS_STELLAR * stride] < 0.0 ||
                               y[GAL_V_DISK_ANGULAR_MOMENTUM * stride] < 0.0 ||
                               y[GAL_V_SPHEROID_MASS_GAS * stride] < 0.0 ||
                               y[GAL_V_SPHEROID_MASS_STELLAR * stride] < 0.0 ||
                               y[GAL_V_SPHEROID_ANGULAR_MOMENTUM * stride] < 0.0)) ||
                   y[GAL_V_HOT_MASS * stride] < 0.0 || y[GAL_V_HOT_OUTER_RADIUS * stride] < 0.0;
  if (neg) post_step_fix(P, HaloConsts{D.om, D.ol, D.a3_norm, D.sinh_arg, D.rvir_coeff}, np, t, y, stride);
}

}  // namespace gal
