/*
 * oracle.c -- CPU restatement of the reference's node-ODE evolution path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  The reference itself cannot be built here (it needs
 * gfortran >= 16, GSL 2.6, HDF5, FoX and a Python code generator), so this
 * file restates, in plain C and in the reference's own scalar one-node-at-a-
 * time form, what the reference computes:
 *
 *   integrator   GSL 2.6 gsl_odeiv2_step_rkck + gsl_odeiv2_evolve_apply
 *                (third party, not vendored; restated from the published
 *                Cash & Karp 1990 tableau and GSL's documented evolve
 *                semantics, called at source/external/gslODEInitVal2/driver2.c:192)
 *   controller   sc2_control_hadjust, source/external/gslODEInitVal2/cscal2.c:93-169
 *   driver       gsl_odeiv2_driver2_apply, driver2.c:148-250
 *   solve        odeSolverSolve, source/numerical/ODE_solver/solver.F90:492-691
 *   trial loop   standardEvolve, source/merger_trees/node_evolver/standard.F90:587-740
 *   RHS          standardODEs -> nodeOperatorMulti (standard.F90:831-946,
 *                nodes/operators/multi.F90:313-332) over the physics
 *                operators of parameters/quickTest.xml:220-289 (cited per
 *                function below)
 *   post-step    standardPostStepProcessing (standard.F90:1160-1185) with the
 *                component post-steps (disk/standard/_class.F90:473-660,
 *                hot_halo/standard/_class.F90:455-475)
 *
 * Parity status of each piece is listed in the header of tests/test_oracle.py.
 * Compiled with -ffp-contract=off so that products and sums round exactly as
 * written.
 */
#include "oracle.h"
#include "../include/gal_tree.h"
#include <math.h>
#include <string.h>
#include <float.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NV GAL_NVAR

/* ---------------------------------------------------------------------- */
/* Physical constants (SI unless stated).  Definitions follow
   source/numerical/constants/{astronomical,atomic,physical,units,prefixes}.F90;
   values are CODATA 2018 / IAU 2015 where the reference takes them from
   external tables.                                                          */
#define C_PI              3.14159265358979323846
#define C_KILO            1.0e3
#define C_HECTO           1.0e2
#define C_CENTI           1.0e-2
#define C_PARSEC          3.0856775814913673e16
#define C_MEGAPARSEC      (1.0e6*C_PARSEC)
#define C_YEAR            3.15581497635456e7       /* astronomical.F90 "year"   */
#define C_GIGAYEAR        (1.0e9*C_YEAR)
#define C_MASS_SOLAR      1.98847e30
#define C_G_SI            6.67430e-11
#define C_G_INTERNAL      (C_G_SI*C_MASS_SOLAR/(C_KILO*C_KILO)/C_MEGAPARSEC)
#define C_MPC_PER_KMS_GYR (C_MEGAPARSEC/C_KILO/C_GIGAYEAR) /* MpcPerKmPerSToGyr */
#define C_AMU             1.66053906660e-27
#define C_A_HYDROGEN      1.0078250322
#define C_A_HELIUM        4.002602
#define C_MASS_H_ATOM     (C_A_HYDROGEN*C_AMU)
#define C_BOLTZMANN       1.380649e-23
#define C_SPEED_LIGHT     2.99792458e8
#define C_THOMSON         6.6524587321e-29
#define C_ERGS            1.0e-7
#define C_X_PRIMORDIAL    0.7514
#define C_Y_PRIMORDIAL    0.2486
#define C_Z_SOLAR         0.0188
/* meanAtomicMassPrimordial (atomic.F90): 1/(2X/A_H + 3Y/A_He) */
#define C_MU_PRIMORDIAL   (1.0/(2.0*C_X_PRIMORDIAL/C_A_HYDROGEN+3.0*C_Y_PRIMORDIAL/C_A_HELIUM))

/* Status codes (see gal_ode.h for their reference origin). */
#define ST_SUCCESS   GAL_STATUS_SUCCESS
#define ST_FAILURE   GAL_STATUS_FAILURE
#define ST_CONTINUE  GAL_STATUS_CONTINUE
#define HADJ_DEC (-1)
#define HADJ_NIL   0
#define HADJ_INC   1

/* ---------------------------------------------------------------------- */
/* Per-node system (the module-level state of standard.F90:137-141 made
   explicit).                                                               */
typedef struct {
  const double *P;          /* GAL_NPARAM physics parameters */
  double np[GAL_NPARAM_NODE];
  const orc_cooling *ct;
  double scale[NV];
  int nonneg[NV];
  double eps_abs, eps_rel;
  double root_tol;
  int64_t n_rhs, n_attempt, n_accept;
} orc_sys;

/* ---------------------------------------------------------------------- */
/* Cosmology and halo scales.
   a(t): closed form of the flat matter+Lambda Friedmann equation
   (cosmologyFunctions matterLambda, source/cosmology/cosmology_functions/matter_lambda.F90);
   virial overdensity: Bryan & Norman (1998) fit to the spherical-collapse
   contrast (virialDensityContrast sphericalCollapseClsnlssMttrCsmlgclCnstnt);
   r_vir, V_vir, t_dyn, T_vir as darkMatterHaloScale virialDensityContrastDefinition
   (source/dark_matter_halos/scales/virial_density_contrast_definition.F90).   */
typedef struct {
  double a, z, mass, r_vir, v_vir, t_dyn, t_vir, lambda;
  double r_disk, v_disk, r_sph, v_sph;
  double r_vir_rate;   /* darkMatterHaloScale%radiusVirialGrowthRate */
} orc_halo;

static void halo_compute(const double *P, const double *np, double t,
                         const double *y, orc_halo *h)
{
  const double om = P[GAL_P_OMEGA_MATTER], ol = P[GAL_P_OMEGA_DARK_ENERGY];
  const double h0 = P[GAL_P_HUBBLE_CONSTANT];               /* km/s/Mpc */
  const double h0_gyr = h0 / C_MPC_PER_KMS_GYR;            /* 1/Gyr     */
  const int satellite = np[GAL_NP_IS_SATELLITE] != 0.0;
  /* basic%mass solved analytically: dark_matter_only_mass/interpolate.F90:215-235 */
  double rate = satellite ? 0.0 : np[GAL_NP_MASS_RATE];
  h->mass = np[GAL_NP_MASS_REFERENCE] + rate * (t - np[GAL_NP_TIME_REFERENCE]);
  /* The virial scales use the mean density at the time the node was last
     isolated (dark_matter_halos/scales/virial_density_contrast.F90:356-382,
     basic%timeLastIsolated): a satellite's reference time is that time. */
  const double t_iso = satellite ? np[GAL_NP_TIME_REFERENCE] : t;
  h->a = pow(om / ol, 1.0 / 3.0) * pow(sinh(1.5 * sqrt(ol) * h0_gyr * t_iso), 2.0 / 3.0);
  h->z = 1.0 / h->a - 1.0;
  double a3 = h->a * h->a * h->a;
  double rho_crit0 = 3.0 * h0 * h0 / (8.0 * C_PI * C_G_INTERNAL);  /* Msun/Mpc^3 */
  double rho_mean = om * rho_crit0 / a3;
  double om_a = (om / a3) / (om / a3 + ol);
  double x = om_a - 1.0;
  double delta_c = 18.0 * C_PI * C_PI + 82.0 * x - 39.0 * x * x;
  double delta_m = delta_c / om_a;
  h->r_vir = cbrt(3.0 * h->mass / (4.0 * C_PI * delta_m * rho_mean));
  h->v_vir = sqrt(C_G_INTERNAL * h->mass / h->r_vir);
  h->t_dyn = C_MPC_PER_KMS_GYR * h->r_vir / h->v_vir;
  /* darkMatterHaloScale%temperatureVirial: 0.5 mu m_u V^2 / k */
  h->t_vir = 0.5 * C_AMU * C_MU_PRIMORDIAL * (C_KILO * h->v_vir) * (C_KILO * h->v_vir) / C_BOLTZMANN;
  /* spin parameter = J / (M V r) (Dark_Matter_Halo_Angular_Momentum_Scale) */
  double jscale = h->mass * h->v_vir * h->r_vir;
  h->lambda = (jscale > 0.0) ? y[GAL_V_SPIN_ANGULAR_MOMENTUM] / jscale : 0.0;
  /* galacticStructureSolver fixed, radiusFixed=virial:
     galactic_structure/radius_solver/fixed.F90:318-325 */
  h->r_disk = P[GAL_P_RADIUS_FACTOR] * h->r_vir * h->lambda;
  h->v_disk = h->v_vir;
  h->r_sph = h->r_disk;
  h->v_sph = h->v_vir;
  /* radiusVirialGrowthRate (virial_density_contrast.F90:338-354):
     (r_vir/3) (Mdot/M - dln(rho_vir)/dt) with rho_vir = Delta_c rho_crit0 E^2;
     dln E^2/dt = -3 H Omega_m(a), dOmega_m(a)/dt = -3 H Omega_m(a) (1 - Omega_m(a)).
     Zero for satellites, whose scales are those of their last isolated time. */
  if (satellite) {
    h->r_vir_rate = 0.0;
  } else {
    double hubble = h0_gyr * sqrt(om / a3 + ol);
    double dx = -3.0 * hubble * om_a * (1.0 - om_a);
    double dln_delta = (82.0 - 78.0 * x) * dx / delta_c;
    double dln_e2 = -3.0 * hubble * om_a;
    h->r_vir_rate = h->r_vir / 3.0 * (rate / h->mass - (dln_delta + dln_e2));
  }
}

/* ---------------------------------------------------------------------- */
/* Gas composition for metallicity-only abundances. */
static void gas_fractions(double metal_fraction, double *x_mass, double *x_number,
                          double *ne_per_nh)
{
  double zf = metal_fraction;
  if (zf < 0.0) zf = 0.0;
  if (zf > 1.0) zf = 1.0;
  double X = (1.0 - zf) * C_X_PRIMORDIAL / (C_X_PRIMORDIAL + C_Y_PRIMORDIAL);
  double Y = (1.0 - zf) * C_Y_PRIMORDIAL / (C_X_PRIMORDIAL + C_Y_PRIMORDIAL);
  double nh = X / C_A_HYDROGEN, nhe = Y / C_A_HELIUM;
  *x_mass = X;
  *x_number = nh / (nh + nhe);
  /* fully ionized H and He (collisional ionization equilibrium above 1e5 K) */
  *ne_per_nh = 1.0 + 2.0 * nhe / nh;
}

/* cieFileInterpolatingFactors + cieFileInterpolate
   (source/cooling/cooling_function/CIE_file.F90:665-735) with "fix"
   extrapolation at all four edges (CIE_file.F90:256-296).  The grid is
   searched by bisection as the reference's interpolator%locate does.      */
static int locate(double x0, double dx, int n, double x)
{
  int lo = 0, hi = n - 1;
  if (x < x0) return 0;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (x0 + dx * mid <= x) lo = mid; else hi = mid;
  }
  if (lo > n - 2) lo = n - 2;
  if (lo < 0) lo = 0;
  return lo;
}

double orc_cooling_function(const orc_cooling *c, double temperature,
                            double metallicity_solar)
{
  double t_min = exp(c->ln_t_min), t_max = exp(c->ln_t_min + c->d_ln_t * (c->n_t - 1));
  double z_min = exp(c->ln_z_min), z_max = exp(c->ln_z_min + c->d_ln_z * (c->n_z - 1));
  double T = temperature, Z = metallicity_solar;
  if (T < t_min) T = t_min;
  if (T > t_max) T = t_max;
  if (Z < z_min) Z = z_min;
  if (Z > z_max) Z = z_max;
  double lt = log(T), lz = log(Z);
  int it = locate(c->ln_t_min, c->d_ln_t, c->n_t, lt);
  int iz = locate(c->ln_z_min, c->d_ln_z, c->n_z, lz);
  double t0 = c->ln_t_min + c->d_ln_t * it, t1 = c->ln_t_min + c->d_ln_t * (it + 1);
  double z0 = c->ln_z_min + c->d_ln_z * iz, z1 = c->ln_z_min + c->d_ln_z * (iz + 1);
  double ht = (lt - t0) / (t1 - t0);
  double hz = (lz - z0) / (z1 - z0);
  const double *tab = c->ln_lambda;
  int nt = c->n_t;
  double v = tab[iz * nt + it] * (1.0 - ht) * (1.0 - hz)
           + tab[(iz + 1) * nt + it] * (1.0 - ht) * hz
           + tab[iz * nt + it + 1] * ht * (1.0 - hz)
           + tab[(iz + 1) * nt + it + 1] * ht * hz;
  return exp(v);
}

static double number_density_hydrogen(double density, double x_mass)
{
  /* cooling_time/simple.F90:150-155 */
  double mpc_cm = C_HECTO * C_MEGAPARSEC;
  return density * x_mass * C_MASS_SOLAR / C_MASS_H_ATOM / (mpc_cm * mpc_cm * mpc_cm);
}

/* simpleTime (source/cooling/cooling_time/simple.F90:128-181) */
double orc_cooling_time(const double *P, const orc_cooling *c, double temperature,
                        double density, double metal_fraction)
{
  const double time_large = 1.0e10;
  double X, xn, ne;
  gas_fractions(metal_fraction, &X, &xn, &ne);
  double nh = number_density_hydrogen(density, X);
  double n_all = nh / xn + ne * nh;
  double lambda = orc_cooling_function(c, temperature, metal_fraction / C_Z_SOLAR) * nh * nh;
  if (lambda > 0.0) {
    double energy = P[GAL_P_DEGREES_OF_FREEDOM] / 2.0 * C_BOLTZMANN * temperature * n_all / C_ERGS;
    return energy / lambda / C_GIGAYEAR;
  }
  return time_large;
}

/* ---------------------------------------------------------------------- */
/* Hot halo: beta profile (beta=2/3) with core radius
   coreRadiusOverVirialRadius*r_vir truncated at outerRadius
   (hot_halo/mass_distribution/beta_profile.F90), virial (isothermal)
   temperature (hot_halo/temperature_profile/virial.F90).                   */
typedef struct {
  int valid;
  double r_core, r_outer, x_outer, rho0, metal_fraction, t_avail;
  double r_cool, r_cool_rate, r_mean;
} orc_hot;

static double beta_density(const orc_hot *hh, double r)
{
  if (r > hh->r_outer) return 0.0;
  double x = r / hh->r_core;
  return hh->rho0 / (1.0 + x * x);
}

typedef struct { const double *P; const orc_cooling *ct; const orc_hot *hh; double temperature; } root_ctx;

/* coolingRadiusRoot (cooling/cooling_radius/simple.F90:389-427) */
static double cooling_radius_root(const root_ctx *rc, double r)
{
  double rho = beta_density(rc->hh, r);
  return orc_cooling_time(rc->P, rc->ct, rc->temperature, rho, rc->hh->metal_fraction) - rc->hh->t_avail;
}

/* Brent's method (Brent 1973, "Algorithms for Minimization without
   Derivatives", ch. 4), the algorithm of GSL's gsl_root_fsolver_brent used
   by the reference's rootFinder.  Relative tolerance on the root.          */
static double brent_root(const root_ctx *rc, double a, double b, double fa, double fb, double tol_rel)
{
  double c = a, fc = fa, d = b - a, e = d;
  for (int iter = 0; iter < 200; iter++) {
    if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { c = a; fc = fa; d = b - a; e = d; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    double tol = 2.0 * DBL_EPSILON * fabs(b) + 0.5 * tol_rel * fabs(b);
    double m = 0.5 * (c - b);
    if (fb == 0.0 || fabs(m) <= tol) return b;
    if (fabs(e) < tol || fabs(fa) <= fabs(fb)) { d = m; e = m; }
    else {
      double p, q, r, s = fb / fa;
      if (a == c) { p = 2.0 * m * s; q = 1.0 - s; }
      else { q = fa / fc; r = fb / fc; p = s * (2.0 * m * q * (q - r) - (b - a) * (r - 1.0)); q = (q - 1.0) * (r - 1.0) * (s - 1.0); }
      if (p > 0.0) q = -q; else p = -p;
      if (2.0 * p < fmin(3.0 * m * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
      else { d = m; e = m; }
    }
    a = b; fa = fb;
    if (fabs(d) > tol) b += d; else b += (m > 0.0 ? tol : -tol);
    fb = cooling_radius_root(rc, b);
  }
  return b;
}

/* Shape integrals of the truncated beta=2/3 profile at x = r_outer / r_core:
   S(x) = x - atan(x) (mass: M = 4 pi rho0 r_c^3 S) and
   N(x) = (x^2 - ln(1 + x^2)) / 2 (mass-weighted mean radius <r> = r_c N / S).
   Below x = 0.01 both differences cancel to a few digits, so their Taylor
   series are summed instead (truncation below 1e-16 relative).             */
static double beta_shape_mass(double x)
{
  if (x < 0.01) {
    double u = x * x;
    return x * u * (1.0 / 3.0 - u * (1.0 / 5.0 - u * (1.0 / 7.0 - u * (1.0 / 9.0))));
  }
  return x - atan(x);
}
static double beta_shape_radius(double x)
{
  double u = x * x;
  if (x < 0.01) return u * u * (1.0 / 4.0 - u * (1.0 / 6.0 - u * (1.0 / 8.0 - u * (1.0 / 10.0))));
  return 0.5 * u - 0.5 * log(1.0 + u);
}

static void hot_compute(const orc_sys *s, const orc_halo *h, const double *y, orc_hot *hh)
{
  const double *P = s->P;
  double m = y[GAL_V_HOT_MASS], ro = y[GAL_V_HOT_OUTER_RADIUS];
  memset(hh, 0, sizeof(*hh));
  if (!(m > 0.0) || !(ro > 0.0)) return;
  hh->valid = 1;
  hh->r_core = P[GAL_P_CORE_RADIUS_OVER_VIRIAL] * h->r_vir;
  hh->r_outer = ro;
  hh->x_outer = ro / hh->r_core;
  double xo = hh->x_outer;
  hh->rho0 = m / (4.0 * C_PI * hh->r_core * hh->r_core * hh->r_core * beta_shape_mass(xo));
  hh->metal_fraction = y[GAL_V_HOT_ABUNDANCES] / m;
  /* mass-weighted mean radius of the truncated profile, used for the
     constantRotation specific angular momentum normalization
     (cooling/specific_angular_momentum/constant_rotation.F90) */
  hh->r_mean = hh->r_core * beta_shape_radius(xo) / beta_shape_mass(xo);
  /* timeAvailable, ageFactor=0: cooling/time_available/White-Frenk.F90:144-146 */
  hh->t_avail = h->t_dyn;
  /* simpleRadius (cooling/cooling_radius/simple.F90:313-387) */
  root_ctx rc = { P, s->ct, hh, h->t_vir };
  double root_outer = cooling_radius_root(&rc, ro);
  if (root_outer < 0.0) {
    hh->r_cool = ro;
  } else {
    double root_zero = cooling_radius_root(&rc, 0.0);
    if (root_zero > 0.0) hh->r_cool = 0.0;
    else hh->r_cool = brent_root(&rc, 0.0, ro, root_zero, root_outer, s->root_tol);
  }
  /* simpleRadiusGrowthRate (cooling_radius/simple.F90:229-309):
     slope = dln(rho)/dln(r) * dln(t_cool)/dln(rho) + dln(T)/dln(r)*(...);
     isothermal -> temperature term zero; CIE -> dln(t_cool)/dln(rho) = -1. */
  if (hh->r_cool >= ro || !(hh->r_cool > 0.0)) {
    hh->r_cool_rate = 0.0;
  } else {
    double x = hh->r_cool / hh->r_core;
    double dlnrho = -2.0 * x * x / (1.0 + x * x);
    double slope = dlnrho * (-1.0);
    /* timeAvailableIncreaseRate = 1 (White-Frenk.F90:161-168) */
    hh->r_cool_rate = (slope != 0.0) ? hh->r_cool / hh->t_avail * 1.0 / slope : 0.0;
  }
}

/* whiteFrenk1991Rate (cooling/cooling_rate/White-Frenk.F90:132-185) */
static double cooling_rate_wf(const orc_sys *s, const orc_halo *h, const orc_hot *hh, const double *y)
{
  if (h->v_vir > s->P[GAL_P_COOLING_VELOCITY_CUTOFF]) return 0.0;
  if (hh->r_cool >= hh->r_outer) return y[GAL_V_HOT_MASS] / h->t_dyn;
  double rho = beta_density(hh, hh->r_cool);
  return 4.0 * C_PI * hh->r_cool * hh->r_cool * rho * hh->r_cool_rate;
}

/* ---------------------------------------------------------------------- */
/* Star formation, stellar populations, feedback. */

/* dynamicalTimeTimescale (star_formation/timescales/dynamical_time.F90:121-188) */
static double sf_timescale(double velocity, double radius, double eff, double alpha, double tmin)
{
  if (velocity <= 0.0) return 0.0;
  if (eff == 0.0) return 0.0;
  double tdyn = C_MPC_PER_KMS_GYR * radius / velocity;
  double tau = tdyn * pow(velocity / 200.0, alpha) / eff;
  return (tau > tmin) ? tau : tmin;
}

/* Instantaneous-recycling rates (stellar_populations/properties/instantaneous.F90:133-182)
   applied to a component with indices (mstar, mformed, zstar, mgas, zgas). */
static void stellar_rates(const double *P, double sfr, double mgas, double zgas,
                          double *dydt, int i_mstar, int i_mformed, int i_zstar,
                          int i_mgas, int i_zgas)
{
  double R = P[GAL_P_RECYCLED_FRACTION], yld = P[GAL_P_METAL_YIELD];
  double zfuel = (mgas > 0.0) ? zgas / mgas : 0.0;   /* massToMassFraction */
  double rate_stellar = (1.0 - R) * sfr;
  double rate_fuel = -rate_stellar;
  double rate_z_stellar = zfuel * rate_stellar;
  double rate_z_fuel = -rate_z_stellar + yld * sfr;
  dydt[i_mstar] += rate_stellar;
  dydt[i_mgas] += rate_fuel;
  dydt[i_zstar] += rate_z_stellar;
  dydt[i_zgas] += rate_z_fuel;
  dydt[i_mformed] += sfr;
}

/* powerLawOutflowRate (stellar_feedback/outflows/power_law/_class.F90:118-164)
   wrapped by rateLimitOutflowRate (stellar_feedback/outflows/rate_limit.F90:111-170).
   rateEnergyInput/feedbackEnergyInputAtInfinityCanonical == SFR for the
   instantaneous population (instantaneous.F90:175). */
static double outflow_rate(double sfr, double velocity, double radius, double mgas,
                           double v0, double alpha, double tfrac, int fixed_mode, double fraction)
{
  double ejective;
  if (fixed_mode) ejective = fraction * sfr;           /* stellarFeedbackOutflows fixed */
  else if (velocity <= 0.0) ejective = 0.0;
  else ejective = pow(v0 / velocity, alpha) * sfr;
  double total = ejective;
  double tdyn = (velocity <= 0.0 || radius <= 0.0) ? 1.0 : C_MPC_PER_KMS_GYR * radius / velocity;
  double rmax = mgas / tdyn / tfrac;
  if (rmax < 0.0) rmax = 0.0;
  if (total > rmax) ejective = ejective * rmax / total;
  return ejective;
}

/* Outflowing gas joins the hot halo's outflowed reservoir
   (Node_Component_Hot_Halo_Standard_Outflowing_{Mass,Ang_Mom,Abundances}_Rate,
   hot_halo/standard/_class.F90:598-720).  The stripped fraction of
   hotHaloOutflowStripping standard, efficiency (1 - M(<r_outer)/M(<r_vir))
   (hot_halo/outflow_stripping/standard.F90:133-173), is zero for centrals and
   for satellites whose outer radius lies within the virial radius; the
   angular momentum is raised by 1/(1 - fractionLossAngularMomentum) so that
   the infall torque applied when the gas cools again returns it. */
static void outflowing_add(const double *P, double mass, double ang, double z, double *dydt)
{
  dydt[GAL_V_HOT_OUTFLOWED_MASS] += mass;
  dydt[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] += ang / (1.0 - P[GAL_P_FRACTION_LOSS_ANGULAR_MOMENTUM]);
  dydt[GAL_V_HOT_OUTFLOWED_ABUNDANCES] += z;
}

/* stellarFeedbackDisksDifferentialEvolution (nodes/operators/physics/stellar_feedback/disks.F90:116-206)
   and the spheroid twin. */
static void feedback_apply(const double *P, double outflow, double mgas, double mstar, double ang, double zgas,
                           double *dydt, int i_mgas, int i_zgas, int i_ang)
{
  if (!(outflow > 0.0)) return;
  double mtot = mgas + mstar;
  double rate_ang = (mtot > 0.0) ? ang * (outflow / mtot) : 0.0;
  double rate_z = (mgas > 0.0) ? (zgas / mgas) * outflow : 0.0;
  dydt[i_mgas] -= outflow;
  dydt[i_ang] -= rate_ang;
  dydt[i_zgas] -= rate_z;
  outflowing_add(P, outflow, rate_ang, rate_z, dydt);
}

/* barInstabilityDifferentialEvolution (nodes/operators/physics/bar_instability.F90:145-248)
   with galacticDynamicsBarInstabilityEfstathiou1982
   (galactic_dynamics/bar_instability/Efstathiou1982.F90:153-258): the disk is
   unstable when eps = 1.18 V / sqrt(G M_disk / r_disk) falls below the
   threshold interpolated in gas fraction; its material moves to the spheroid
   on t_dyn ((eps_c - eps_iso)/(eps_c - eps))^2.  Returns the timescale, or
   -1 when the disk is stable.                                               */
static double bar_timescale(const double *P, double mg, double ms, double ang, double radius, double velocity)
{
  const double eps_iso = 0.6221297315, boost = 1.1800237580, dimensionless_max = 1.0e10;
  if (ang <= 0.0) return -1.0;
  if (velocity <= 0.0 || radius <= 0.0) return -1.0;
  double mdisk = mg + ms;
  double fgas = mg / mdisk;
  double threshold = P[GAL_P_BAR_THRESHOLD_STELLAR] * (1.0 - fgas) + P[GAL_P_BAR_THRESHOLD_GASEOUS] * fgas;
  double estimator = DBL_MAX;                  /* huge(0d0): stable */
  if (mdisk >= 0.0) {
    double v_self = sqrt(C_G_INTERNAL * mdisk / radius);
    if (v_self > 0.0 && ilogb(boost * velocity) - ilogb(v_self) <= DBL_MAX_EXP) {
      double e = boost * velocity / v_self;
      estimator = (e > eps_iso) ? e : eps_iso;
    }
  }
  if (!(estimator < threshold)) return -1.0;
  double c_kms = C_SPEED_LIGHT / C_KILO;
  double t_dyn = C_MPC_PER_KMS_GYR * radius / ((velocity < c_kms) ? velocity : c_kms);
  double iso_rel = threshold - eps_iso, est_rel = threshold - estimator;
  double dimensionless = (iso_rel > dimensionless_max * est_rel) ? dimensionless_max : (iso_rel / est_rel) * (iso_rel / est_rel);
  double tmin = P[GAL_P_BAR_TIMESCALE_MINIMUM];
  return ((t_dyn > tmin) ? t_dyn : tmin) * dimensionless;
}

static void bar_instability(const double *P, const orc_halo *h, const double *y, double *dydt)
{
  double mg = y[GAL_V_DISK_MASS_GAS], ms = y[GAL_V_DISK_MASS_STELLAR], j = y[GAL_V_DISK_ANGULAR_MOMENTUM];
  if (j < 0.0 || h->r_disk < 0.0 || mg < 0.0) return;
  double tau = bar_timescale(P, mg, ms, j, h->r_disk, h->v_disk);
  if (tau < 0.0) return;
  double rg = fmax(0.0, mg) / tau;
  dydt[GAL_V_DISK_MASS_GAS] -= rg;
  dydt[GAL_V_SPHEROID_MASS_GAS] += rg;
  double rs = fmax(0.0, ms) / tau;
  dydt[GAL_V_DISK_MASS_STELLAR] -= rs;
  dydt[GAL_V_SPHEROID_MASS_STELLAR] += rs;
  double rj = fmax(0.0, j) / tau;
  dydt[GAL_V_DISK_ANGULAR_MOMENTUM] -= (1.0 - P[GAL_P_BAR_J_RETAINED_DISK]) * rj;
  dydt[GAL_V_SPHEROID_ANGULAR_MOMENTUM] += P[GAL_P_BAR_J_RETAINED_SPHEROID] * rj;
  double rzg = fmax(0.0, y[GAL_V_DISK_ABUNDANCES_GAS]) / tau;
  dydt[GAL_V_DISK_ABUNDANCES_GAS] -= rzg;
  dydt[GAL_V_SPHEROID_ABUNDANCES_GAS] += rzg;
  double rzs = fmax(0.0, y[GAL_V_DISK_ABUNDANCES_STELLAR]) / tau;
  dydt[GAL_V_DISK_ABUNDANCES_STELLAR] -= rzs;
  dydt[GAL_V_SPHEROID_ABUNDANCES_STELLAR] += rzs;
  /* the efstathiou1982 class applies no external torque (specific torque 0) */
}

/* ---------------------------------------------------------------------- */
/* Black holes: accretion_rates/standard.F90:236-440, accretion/Bondi_Hoyle_Lyttleton.F90:34-78,
   thermodynamics/ideal_gases.F90:34-69, black_holes/fundamentals.F90:123-260,
   accretion_disks/Shakura_Sunyaev.F90:70-179 (thin disk). */
static double sound_speed(double temperature)
{
  /* adiabatic, gamma = 5/3, primordial mean atomic mass; km/s */
  return sqrt(5.0 / 3.0 * C_BOLTZMANN * temperature / (C_MU_PRIMORDIAL * C_AMU)) / C_KILO;
}
static double bhl_radius(double mass, double temperature)
{
  if (!(temperature > 0.0)) return HUGE_VAL;
  double cs = sound_speed(temperature);
  return C_G_INTERNAL * mass / (cs * cs);
}
static double bhl_rate(double mass, double density, double temperature)
{
  double cs = sound_speed(temperature);
  double gm = C_G_INTERNAL * mass;
  return (C_KILO * C_GIGAYEAR / C_MEGAPARSEC) * 4.0 * C_PI * gm * gm * density / pow(cs * cs, 1.5);
}
static double bhl_rate_radius(double density, double temperature, double radius)
{
  double cs = sound_speed(temperature);
  return (C_KILO * C_GIGAYEAR / C_MEGAPARSEC) * 4.0 * C_PI * radius * radius * density * sqrt(cs * cs);
}
static double jeans_length(double temperature, double density)
{
  return sound_speed(temperature) / sqrt(C_G_INTERNAL) / sqrt(density);
}
static double eddington_rate(double mass)
{
  return 4.0 * C_PI * C_G_SI * mass * C_MASS_H_ATOM * C_GIGAYEAR / C_THOMSON / C_SPEED_LIGHT;
}
static double isco_radius(double a)
{
  double a1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
  double a2 = sqrt(3.0 * a * a + a1 * a1);
  /* A1 <= 3 for |a| <= 1; rounding can push it just above 3 for tiny spins */
  double d = 3.0 - a1;
  if (d < 0.0) d = 0.0;
  return 3.0 + a2 - sqrt(d * (3.0 + a1 + 2.0 * a2));
}
static double isco_energy(double a)
{
  double r = isco_radius(a), sr = sqrt(r);
  return (r * r - 2.0 * r + a * sr) / r / sqrt(r * r - 3.0 * r + 2.0 * a * sr);
}
static double isco_angular_momentum(double a)
{
  double r = isco_radius(a), sr = sqrt(r);
  return (r * r - 2.0 * a * sr + a * a) / sr / sqrt(r * r - 3.0 * r + 2.0 * a * sr);
}
static double clamp_spin(double a)
{
  if (a < 0.0) return 0.0;
  if (a > 0.998) return 0.998;
  return a;
}
static double efficiency_radiative(double a, double mdot)
{
  return (mdot > 0.0) ? 1.0 - isco_energy(clamp_spin(a)) : 0.0;
}
static double power_jet(double mbh, double a_in, double mdot)
{
  /* Meier (2001) eqs. 4-5 (Shakura_Sunyaev.F90:92-140) */
  if (mdot <= 0.0) return 0.0;
  double a = clamp_spin(a_in);
  const double unit = C_ERGS * C_GIGAYEAR / C_MASS_SOLAR / (C_KILO * C_KILO);
  double mdim = mdot / eddington_rate(mbh);
  double mbh9 = mbh / 1.0e9;
  if (!(mbh9 > 0.0) || !(mdim > 0.0)) return 0.0;
  if (a > 0.8)
    return pow(10.0, 42.7) * unit * pow(mbh9, 0.9) * pow(mdim, 1.2) * (1.0 + 1.1 * a + 0.29 * a * a);
  return pow(10.0, 41.7) * unit * pow(mbh9, 0.9) * pow(mdim, 1.2) * exp(3.785 * a);
}
static double spin_up_rate(double mbh, double a_in, double mdot)
{
  if (mdot == 0.0 || !(mbh > 0.0)) return 0.0;
  double a = clamp_spin(a_in);
  return (isco_angular_momentum(a) - 2.0 * a * isco_energy(a)) * mdot / mbh;
}
static double hernquist_density(double mass, double scale, double r)
{
  if (!(mass > 0.0) || !(scale > 0.0) || !(r > 0.0)) return 0.0;
  double rs = r + scale;
  return mass * scale / (2.0 * C_PI * r * rs * rs * rs);
}
/* standardHotModeFraction (accretion_rates/standard.F90:442-466) */
static double hot_mode_fraction(double rc_over_rvir)
{
  if (rc_over_rvir < 0.9) return 1.0;
  if (rc_over_rvir > 1.0) return 0.0;
  double x = (rc_over_rvir - 0.9) / (1.0 - 0.9);
  return x * x * (2.0 * x - 3.0) + 1.0;
}

typedef struct { double mdot_sph, mdot_hot, mdot, p_jet, eff_rad; } orc_bh;

/* blackHoleWindCiotti2009%power (black_holes/winds/Ciotti2009.F90:113-193):
   the wind carries efficiencyWind (times the radiative efficiency) of the
   accreted rest-mass energy, coupled to the spheroid gas by the smoothstep of
   P_ISM / P_wind - 1/2.  [Msun (km/s)^2 / Gyr]                              */
static double wind_power(const double *P, const orc_bh *b, double mgas_sph, double radius_sph)
{
  const double velocity_wind = 1.0e4, temperature_ism = 1.0e4;   /* km/s, K */
  double eff = P[GAL_P_BH_WIND_EFFICIENCY];
  if (P[GAL_P_BH_WIND_SCALES_WITH_RADIATIVE] != 0.0) eff *= b->eff_rad;
  if (b->mdot <= 0.0 || eff <= 0.0) return 0.0;
  double coupled = 0.0;
  if (mgas_sph > 0.0 && radius_sph > 0.0) {
    double r = radius_sph * C_MEGAPARSEC;
    double p_wind = eff * b->mdot * C_MASS_SOLAR / C_GIGAYEAR * C_SPEED_LIGHT * C_SPEED_LIGHT / 4.0 / C_PI
                    / (velocity_wind * C_KILO) / (r * r);
    double p_ism = 3.0 / 4.0 / C_PI * mgas_sph * C_MASS_SOLAR / (r * r * r) / C_MASS_H_ATOM * 3.0 / 2.0
                   * C_BOLTZMANN * temperature_ism;
    double x = p_ism / p_wind - 0.5;
    if (x <= 0.0) coupled = 0.0;
    else if (x >= 1.0) coupled = 1.0;
    else coupled = 3.0 * x * x - 2.0 * x * x * x;
  }
  return eff * coupled * b->mdot * (C_SPEED_LIGHT / C_KILO) * (C_SPEED_LIGHT / C_KILO);
}

static void bh_accretion(const orc_sys *s, const orc_halo *h, const orc_hot *hh, const double *y, orc_bh *b)
{
  const double *P = s->P;
  memset(b, 0, sizeof(*b));
  double mbh = y[GAL_V_BH_MASS], a = y[GAL_V_BH_SPIN];
  if (!(mbh > 0.0)) return;
  /* spheroid mode */
  double tsph = P[GAL_P_BH_TEMPERATURE_SPHEROID];
  double racc = bhl_radius(mbh, tsph);
  if (y[GAL_V_BH_RADIAL_POSITION] > racc) racc = y[GAL_V_BH_RADIAL_POSITION];
  double mgas_s = y[GAL_V_SPHEROID_MASS_GAS];
  double rho = hernquist_density(mgas_s, h->r_sph, racc);
  if (rho > 1.0) {
    double lj = jeans_length(tsph, rho);
    if (lj > h->r_sph) lj = h->r_sph;
    if (lj > racc) rho = hernquist_density(mgas_s, h->r_sph, lj);
    double m = P[GAL_P_BH_ENHANCEMENT_SPHEROID] * bhl_rate(mbh, rho, tsph);
    if (m < 0.0) m = 0.0;
    double e = efficiency_radiative(a, m);
    if (e > 0.0) { double cap = eddington_rate(mbh) / e; if (cap < m) m = cap; }
    b->mdot_sph = m;
  }
  /* hot-halo mode (hot mode only) */
  if (hh->valid) {
    double thot = h->t_vir;
    double r2 = bhl_radius(mbh, thot);
    if (r2 > hh->r_outer) r2 = hh->r_outer;
    double fhot = hot_mode_fraction(hh->r_cool / h->r_vir);
    double rho2 = fhot * beta_density(hh, r2);
    if (rho2 > 1.0) {
      double m = P[GAL_P_BH_ENHANCEMENT_HOT_HALO] * bhl_rate_radius(rho2, thot, r2);
      if (m < 0.0) m = 0.0;
      double mmax = y[GAL_V_HOT_MASS] / (hh->r_outer / sound_speed(thot) * C_MPC_PER_KMS_GYR);
      if (mmax < 0.0) mmax = 0.0;
      if (m > mmax) m = mmax;
      double e = efficiency_radiative(a, m);
      if (e > 0.0) { double cap = eddington_rate(mbh) / e; if (cap < m) m = cap; }
      b->mdot_hot = m;
    }
  }
  b->mdot = b->mdot_sph + b->mdot_hot;
  if (b->mdot > 0.0) {
    b->eff_rad = efficiency_radiative(a, b->mdot);
    b->p_jet = power_jet(mbh, a, b->mdot);
  }
}

/* ---------------------------------------------------------------------- */
/* CGMOuterRadiusRamPressureStripping
   (circumgalactic_medium/outer_radius/ram_pressure_stripping.F90:111-242). */
static double hot_density(const orc_hot *hh, double r)
{
  return hh->valid ? beta_density(hh, r) : 0.0;
}

/* The outer radius starts at the virial radius (hot_halo/standard.F90 sets
   outerRadius = radiusVirial), so whether it lies inside the virial radius is
   an exact tie in the reference, where both come from one function.  Here the
   initial value is written by the caller and the virial radius recomputed in
   the RHS; a last-place difference between the two must not switch the
   returning-outflow term on, so the tie is judged to 1e-12 relative. */
static int reaches_virial(double r_outer, double r_vir)
{
  return r_outer >= r_vir * (1.0 - 1.0e-12);
}

/* hotHaloRamPressureStrippingFont2008%radiusStripped
   (hot_halo/ram_pressure_stripping/Font2008.F90:115-267): the radius where
   the gravitational binding force per area, formFactor G M(<r) rho_hot(r) / r,
   equals the ram pressure.  With the total enclosed mass of the isothermal
   halo, M(<r) = M_vir r / r_vir (the profile behind the virial temperature of
   the hot gas), the binding force is formFactor V_vir^2 rho_hot(r), and the
   beta profile inverts it in closed form: r = r_c sqrt(formFactor V_vir^2
   rho_0 / F - 1), capped at the outer radius where the truncated profile
   ends.  The reference's bracketing decisions are kept: the virial radius if
   the binding force there exceeds F, zero if it does not at 1e-6 r_vir.     */
static double ram_pressure_radius(const double *P, const orc_halo *h, const orc_hot *hh, double force)
{
  const double rv = h->r_vir, ff = P[GAL_P_RAM_PRESSURE_FORM_FACTOR];
  const double bind = ff * h->v_vir * h->v_vir;
  const double x_v = rv / hh->r_core;
  const double rho_rv = (hh->valid && reaches_virial(hh->r_outer, rv)) ? hh->rho0 / (1.0 + x_v * x_v) : 0.0;
  if (bind * rho_rv - force >= 0.0) return rv;
  if (bind * hot_density(hh, 1.0e-6 * rv) - force <= 0.0) return 0.0;
  double r = hh->r_core * sqrt(bind * hh->rho0 / force - 1.0);
  if (r > hh->r_outer) r = hh->r_outer;
  return r;
}

/* hotHaloRamPressureTimescaleRamPressureAcceleration
   (hot_halo/ram_pressure_stripping/timescale/ram_pressure_acceleration.F90:79-134) */
static double ram_pressure_timescale(const orc_halo *h, const orc_hot *hh, double r_outer, double force,
                                     double v_host)
{
  if (force <= 0.0) return DBL_MAX;
  double sigma = r_outer * hot_density(hh, r_outer);
  if (r_outer <= 0.0 || sigma <= 0.0) return h->t_dyn;
  double a = sqrt(2.0 * r_outer * sigma / force), b = r_outer / 10.0 / v_host;
  return C_MPC_PER_KMS_GYR * ((a > b) ? a : b);
}

static void cgm_outer_radius(const orc_sys *s, const orc_halo *h, const orc_hot *hh, const double *y,
                             double *dydt)
{
  const double *P = s->P;
  const double ro = y[GAL_V_HOT_OUTER_RADIUS], rv = h->r_vir;
  /* haloDynamicalTimeRate (hot_halo/outflow_reincorporation/halo_dynamical_time.F90:113-128) */
  const double m_return = y[GAL_V_HOT_OUTFLOWED_MASS] * P[GAL_P_REINCORPORATION_MULTIPLIER] / h->t_dyn;
  if (!reaches_virial(ro, rv)) {
    double rho = hot_density(hh, ro);
    if (ro > 0.0 && rho > 0.0) {
      double fb = P[GAL_P_OMEGA_BARYON] / P[GAL_P_OMEGA_MATTER];
      double rho_min = fb * h->mass / (rv * rv * rv) / 4.0 / C_PI;
      dydt[GAL_V_HOT_OUTER_RADIUS] += m_return / 4.0 / C_PI / (ro * ro) / ((rho > rho_min) ? rho : rho_min);
    } else if (m_return > 0.0) {
      dydt[GAL_V_HOT_OUTER_RADIUS] += m_return / h->mass * rv;
    }
  }
  if (s->np[GAL_NP_IS_SATELLITE] != 0.0) {
    const double force = s->np[GAL_NP_RAM_PRESSURE_FORCE];
    double r_rp = ram_pressure_radius(P, h, hh, force);
    double dr = 0.0;
    if (r_rp < ro && y[GAL_V_HOT_ANGULAR_MOMENTUM] > 0.0)
      dr = (r_rp - ro) / ram_pressure_timescale(h, hh, ro, force, s->np[GAL_NP_HOST_VELOCITY_VIRIAL]);
    if (dr != 0.0 && y[GAL_V_HOT_MASS] > 0.0 && ro > 1.0e-3 * rv) {
      double m = y[GAL_V_HOT_MASS];
      double loss = 4.0 * C_PI * hot_density(hh, ro) * ro * ro * dr;
      dydt[GAL_V_HOT_OUTER_RADIUS] += dr;
      dydt[GAL_V_HOT_MASS] += loss;
      dydt[GAL_V_HOT_ANGULAR_MOMENTUM] += y[GAL_V_HOT_ANGULAR_MOMENTUM] * (loss / m);
      dydt[GAL_V_HOT_ABUNDANCES] += y[GAL_V_HOT_ABUNDANCES] * (loss / m);
      /* hotHaloOutflowStripping standard: satellites are stripped */
      dydt[GAL_V_HOT_STRIPPED_MASS] -= loss;
      dydt[GAL_V_HOT_STRIPPED_ABUNDANCES] += y[GAL_V_HOT_ABUNDANCES] * (-loss / m);
    }
  } else {
    dydt[GAL_V_HOT_OUTER_RADIUS] += h->r_vir_rate;
  }
}

/* ---------------------------------------------------------------------- */
/* The RHS: nodeOperatorMulti%differentialEvolution over the operators of
   parameters/quickTest.xml:220-289 in list order. */
static int rhs_eval(orc_sys *s, double t, const double *y, double *dydt)
{
  const double *P = s->P;
  s->n_rhs++;
  for (int i = 0; i < NV; i++) dydt[i] = 0.0;
  int mode = (int)P[GAL_P_PHYSICS_MODE];
  if (mode == GAL_PHYSICS_TEST) {
    /* source/tests/ODE_solver/functions.F90:37-80 */
    dydt[0] = sin(t);
    dydt[1] = y[2];
    dydt[2] = -1.0 * y[1];
    return ST_SUCCESS;
  }
  orc_halo h;
  halo_compute(P, s->np, t, y, &h);
  /* spin%angularMomentum growth (halo_angular_momentum_interpolate.F90) */
  dydt[GAL_V_SPIN_ANGULAR_MOMENTUM] += (s->np[GAL_NP_IS_SATELLITE] != 0.0) ? 0.0 : s->np[GAL_NP_SPIN_RATE];
  const int ops = (mode == GAL_PHYSICS_STANDARD) ? (int)P[GAL_P_OPERATORS] : 0;
  /* satelliteMassLoss (nodes/operators/physics/satellite/mass_loss/_class.F90:176-200):
     the bound mass of an isolated halo follows its basic mass; a satellite's
     loses mass at darkMatterHaloMassLossRate, whose default (zero,
     dark_matter_halos/mass_loss_rates/_class.F90:39) quickTest.xml keeps. */
  if (ops & GAL_OPERATOR_SATELLITE_MASS_LOSS)
    dydt[GAL_V_SATELLITE_BOUND_MASS] += (s->np[GAL_NP_IS_SATELLITE] != 0.0) ? 0.0 : s->np[GAL_NP_MASS_RATE];

  /* starFormationDisks: nodes/operators/physics/star_formation/disks.F90:200-325 */
  double mg_d = y[GAL_V_DISK_MASS_GAS], ms_d = y[GAL_V_DISK_MASS_STELLAR];
  double j_d = y[GAL_V_DISK_ANGULAR_MOMENTUM], zg_d = y[GAL_V_DISK_ABUNDANCES_GAS];
  int disk_ok = !(j_d < 0.0 || h.r_disk < 0.0 || mg_d < 0.0);
  double sfr_d = 0.0;
  if (disk_ok) {
    double tau = (mode == GAL_PHYSICS_BOX) ? P[GAL_P_BOX_TIMESCALE]
               : sf_timescale(h.v_disk, h.r_disk, P[GAL_P_SF_DISK_EFFICIENCY],
                              P[GAL_P_SF_DISK_EXPONENT_VELOCITY], P[GAL_P_SF_DISK_TIMESCALE_MINIMUM]);
    sfr_d = (tau > 0.0) ? mg_d / tau : 0.0;   /* timescaleRate, star_formation/rates/disks/timescale.F90:109-130 */
    stellar_rates(P, sfr_d, mg_d, zg_d, dydt, GAL_V_DISK_MASS_STELLAR, GAL_V_DISK_MASS_STELLAR_FORMED,
                  GAL_V_DISK_ABUNDANCES_STELLAR, GAL_V_DISK_MASS_GAS, GAL_V_DISK_ABUNDANCES_GAS);
  }
  /* starFormationSpheroids: nodes/operators/physics/star_formation/spheroids.F90:152 */
  double mg_s = y[GAL_V_SPHEROID_MASS_GAS], ms_s = y[GAL_V_SPHEROID_MASS_STELLAR];
  double j_s = y[GAL_V_SPHEROID_ANGULAR_MOMENTUM], zg_s = y[GAL_V_SPHEROID_ABUNDANCES_GAS];
  int sph_ok = !(j_s < 0.0 || h.r_sph < 0.0 || mg_s < 0.0);
  double sfr_s = 0.0;
  if (sph_ok) {
    double tau = (mode == GAL_PHYSICS_BOX) ? P[GAL_P_BOX_TIMESCALE]
               : sf_timescale(h.v_sph, h.r_sph, P[GAL_P_SF_SPHEROID_EFFICIENCY],
                              P[GAL_P_SF_SPHEROID_EXPONENT_VELOCITY], P[GAL_P_SF_SPHEROID_TIMESCALE_MINIMUM]);
    sfr_s = (tau > 0.0) ? mg_s / tau : 0.0;
    stellar_rates(P, sfr_s, mg_s, zg_s, dydt, GAL_V_SPHEROID_MASS_STELLAR, GAL_V_SPHEROID_MASS_STELLAR_FORMED,
                  GAL_V_SPHEROID_ABUNDANCES_STELLAR, GAL_V_SPHEROID_MASS_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS);
  }
  /* stellarFeedbackDisks: nodes/operators/physics/stellar_feedback/disks.F90:116-206 */
  if (disk_ok) {
    double out = outflow_rate(sfr_d, h.v_disk, h.r_disk, mg_d, P[GAL_P_FB_DISK_VELOCITY], P[GAL_P_FB_DISK_EXPONENT],
                              P[GAL_P_FB_DISK_TIMESCALE_MINIMUM], mode == GAL_PHYSICS_BOX, P[GAL_P_BOX_OUTFLOW_FRACTION]);
    feedback_apply(P, out, mg_d, ms_d, j_d, zg_d, dydt, GAL_V_DISK_MASS_GAS, GAL_V_DISK_ABUNDANCES_GAS, GAL_V_DISK_ANGULAR_MOMENTUM);
  }
  /* stellarFeedbackSpheroids: nodes/operators/physics/stellar_feedback/spheroids.F90:116 */
  if (sph_ok) {
    double out = outflow_rate(sfr_s, h.v_sph, h.r_sph, mg_s, P[GAL_P_FB_SPHEROID_VELOCITY], P[GAL_P_FB_SPHEROID_EXPONENT],
                              P[GAL_P_FB_SPHEROID_TIMESCALE_MINIMUM], mode == GAL_PHYSICS_BOX,
                              P[GAL_P_BOX_OUTFLOW_FRACTION_SPHEROID]);
    feedback_apply(P, out, mg_s, ms_s, j_s, zg_s, dydt, GAL_V_SPHEROID_MASS_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS, GAL_V_SPHEROID_ANGULAR_MOMENTUM);
  }
  if (mode == GAL_PHYSICS_BOX) return ST_SUCCESS;
  if (ops & GAL_OPERATOR_BAR_INSTABILITY) bar_instability(P, &h, y, dydt);
  orc_hot hh;
  hot_compute(s, &h, y, &hh);
  /* blackHolesAccretion: nodes/operators/physics/black_holes/accretion.F90:113-170 */
  orc_bh b;
  bh_accretion(s, &h, &hh, y, &b);
  if (b.mdot > 0.0) {
    double c_kms = C_SPEED_LIGHT / C_KILO;
    double eff_jet = b.p_jet / b.mdot / (c_kms * c_kms);
    double reduced = b.mdot * (1.0 - b.eff_rad - eff_jet);
    if (!(y[GAL_V_BH_MASS] <= 0.0 && reduced < 0.0)) {
      double spin_up = spin_up_rate(y[GAL_V_BH_MASS], y[GAL_V_BH_SPIN], b.mdot);
      dydt[GAL_V_BH_MASS] += reduced;
      /* spheroid%massGasSinkRate: gas, metals and angular momentum in proportion */
      if (b.mdot_sph != 0.0) {
        dydt[GAL_V_SPHEROID_MASS_GAS] -= b.mdot_sph;
        if (mg_s > 0.0) dydt[GAL_V_SPHEROID_ABUNDANCES_GAS] -= b.mdot_sph * (zg_s / mg_s);
        if (mg_s + ms_s > 0.0) dydt[GAL_V_SPHEROID_ANGULAR_MOMENTUM] -= b.mdot_sph * (j_s / (mg_s + ms_s));
      }
      /* hotHalo%massSinkRate: mass, metals and angular momentum in proportion */
      if (b.mdot_hot != 0.0 && y[GAL_V_HOT_MASS] > 0.0) {
        double mh = y[GAL_V_HOT_MASS];
        dydt[GAL_V_HOT_MASS] -= b.mdot_hot;
        dydt[GAL_V_HOT_ABUNDANCES] -= b.mdot_hot * (y[GAL_V_HOT_ABUNDANCES] / mh);
        dydt[GAL_V_HOT_ANGULAR_MOMENTUM] -= b.mdot_hot * (y[GAL_V_HOT_ANGULAR_MOMENTUM] / mh);
      }
      dydt[GAL_V_BH_SPIN] += spin_up;
    }
  }
  /* blackHolesWinds (nodes/operators/physics/black_holes/winds.F90:102-133):
     the wind power enters the spheroid gas as energy, which drives an outflow
     efficiencyEnergeticOutflow P / V_spheroid^2 into the hot halo
     (Node_Component_Spheroid_Standard_Energy_Gas_Input_Rate,
     spheroid/standard/_class.F90:673-725). */
  if (ops & GAL_OPERATOR_BLACK_HOLE_WINDS) {
    double power = wind_power(P, &b, mg_s, h.r_sph);
    if (power > 0.0 && mg_s > 0.0 && mg_s + ms_s > 0.0 && h.v_sph > 0.0) {
      double mo = P[GAL_P_SPHEROID_ENERGETIC_OUTFLOW] * power / (h.v_sph * h.v_sph);
      double jo = (mo / (mg_s + ms_s)) * j_s;
      double zo = (mo / mg_s) * zg_s;
      dydt[GAL_V_SPHEROID_MASS_GAS] -= mo;
      dydt[GAL_V_SPHEROID_ANGULAR_MOMENTUM] -= jo;
      dydt[GAL_V_SPHEROID_ABUNDANCES_GAS] -= zo;
      outflowing_add(P, mo, jo, zo, dydt);
    }
  }
  /* CGMAccretion (nodes/operators/physics/circumgalactic_medium/accretion.F90:517)
     with accretionHalo simple (accretion/halo/simple.F90:281-378). */
  if (s->np[GAL_NP_IS_SATELLITE] == 0.0 && h.mass > 0.0) {
    double fb = P[GAL_P_OMEGA_BARYON] / P[GAL_P_OMEGA_MATTER];
    double mrate = s->np[GAL_NP_MASS_RATE];
    double failed = (h.z < P[GAL_P_REDSHIFT_REIONIZATION] && h.v_vir < P[GAL_P_VELOCITY_SUPPRESSION_REIONIZATION]) ? 1.0 : 0.0;
    double growth = mrate / h.mass;
    double m_un = y[GAL_V_HOT_UNACCRETED_MASS], z_un = y[GAL_V_HOT_UNACCRETED_ABUNDANCES];
    double rate_acc = fb * mrate * (1.0 - failed) + m_un * growth * (1.0 - failed);
    double rate_fail = fb * mrate * failed - m_un * growth * (1.0 - failed);
    double rate_z = z_un * growth * (1.0 - failed);
    double j_halo = y[GAL_V_SPIN_ANGULAR_MOMENTUM] / h.mass;
    dydt[GAL_V_HOT_MASS] += rate_acc;
    dydt[GAL_V_HOT_UNACCRETED_MASS] += rate_fail;
    dydt[GAL_V_HOT_ABUNDANCES] += rate_z;
    dydt[GAL_V_HOT_UNACCRETED_ABUNDANCES] -= rate_z;
    dydt[GAL_V_HOT_ANGULAR_MOMENTUM] += rate_acc * j_halo;
  }
  /* CGMOutflowReincorporation (circumgalactic_medium/outflow_reincorporation.F90:272-330)
     with haloDynamicalTimeRate (hot_halo/outflow_reincorporation/halo_dynamical_time.F90:113-125) */
  {
    double m_out = y[GAL_V_HOT_OUTFLOWED_MASS];
    if (m_out > 0.0) {
      double r = m_out * P[GAL_P_REINCORPORATION_MULTIPLIER] / h.t_dyn;
      double rj = y[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] * (r / m_out);
      double rz = y[GAL_V_HOT_OUTFLOWED_ABUNDANCES] * (r / m_out);
      dydt[GAL_V_HOT_OUTFLOWED_MASS] -= r;
      dydt[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] -= rj;
      dydt[GAL_V_HOT_OUTFLOWED_ABUNDANCES] -= rz;
      dydt[GAL_V_HOT_MASS] += r;
      dydt[GAL_V_HOT_ANGULAR_MOMENTUM] += rj;
      dydt[GAL_V_HOT_ABUNDANCES] += rz;
    }
  }
  /* CGMCoolingHeating (circumgalactic_medium/cooling_heating.F90:216-381) */
  if (hh.valid && y[GAL_V_HOT_ANGULAR_MOMENTUM] > 0.0) {
    double mh = y[GAL_V_HOT_MASS];
    double cool = cooling_rate_wf(s, &h, &hh, y);
    double heat = P[GAL_P_EFFICIENCY_RADIO_MODE] * b.p_jet / (h.v_vir * h.v_vir);
    if (heat > cool) {
      if (P[GAL_P_EXCESS_HEAT_DRIVES_OUTFLOW] != 0.0) {
        double out = heat - cool;
        double cap = P[GAL_P_RATE_MAXIMUM_EXPULSION] * mh / h.t_dyn;
        if (cap < out) out = cap;
        double rz = y[GAL_V_HOT_ABUNDANCES] * (out / mh);
        double rj = y[GAL_V_HOT_ANGULAR_MOMENTUM] * (out / mh);
        dydt[GAL_V_HOT_MASS] -= out;
        dydt[GAL_V_HOT_ABUNDANCES] -= rz;
        dydt[GAL_V_HOT_ANGULAR_MOMENTUM] -= rj;
        /* the expelled gas leaves the node; a satellite's joins its stripped
           reservoir (hotHaloOutflowStripping standard never strips centrals,
           outflow_stripping/standard.F90:120-131) */
        if (s->np[GAL_NP_IS_SATELLITE] != 0.0) {
          dydt[GAL_V_HOT_STRIPPED_MASS] += out;
          dydt[GAL_V_HOT_STRIPPED_ABUNDANCES] += rz;
        }
      }
    } else if (cool > heat) {
      cool = cool - heat;
      if (cool < 0.0) cool = 0.0;
      /* constantRotation: j(r) = (J/M) r / <r> */
      double jspec = (y[GAL_V_HOT_ANGULAR_MOMENTUM] / mh) * hh.r_cool / hh.r_mean;
      double rj = cool * jspec;
      double rz = cool * y[GAL_V_HOT_ABUNDANCES] / mh;
      dydt[GAL_V_HOT_MASS] -= cool;
      dydt[GAL_V_HOT_ANGULAR_MOMENTUM] -= rj;
      dydt[GAL_V_HOT_ABUNDANCES] -= rz;
      dydt[GAL_V_DISK_MASS_GAS] += cool;
      dydt[GAL_V_DISK_ABUNDANCES_GAS] += rz;
      dydt[GAL_V_DISK_ANGULAR_MOMENTUM] += rj * (1.0 - P[GAL_P_FRACTION_LOSS_ANGULAR_MOMENTUM]);
    }
  }
  if (ops & GAL_OPERATOR_CGM_OUTER_RADIUS) cgm_outer_radius(s, &h, &hh, y, dydt);
  /* coolingEnergyRadiated (nodes/operators/physics/cooling_energy_radiated.F90:200-260) */
  if (hh.valid) {
    double ro = y[GAL_V_HOT_OUTER_RADIUS];
    double rho = y[GAL_V_HOT_MASS] / (4.0 / 3.0 * C_PI * ro * ro * ro);
    double X, xn, ne;
    gas_fractions(hh.metal_fraction, &X, &xn, &ne);
    double nh = number_density_hydrogen(rho, X);
    double n_all = nh / xn + ne * nh;
    double lam = orc_cooling_function(s->ct, h.t_vir, hh.metal_fraction / C_Z_SOLAR) * nh * nh;
    double rcm = ro * C_MEGAPARSEC / C_CENTI;
    double count = n_all * 4.0 * C_PI / 3.0 * rcm * rcm * rcm;
    dydt[GAL_V_HOT_ENERGY_RADIATED] += lam / nh * count;
  }
  return ST_SUCCESS;
}

/* ---------------------------------------------------------------------- */
/* Component scale-set rules evaluated at the start of each evolve:
   hot_halo/standard/_class.F90:804-840, disk/standard/_class.F90:735-828,
   spheroid/standard/_class.F90:804-860, black_hole/standard/_class.F90:195-240,
   cooling_energy_radiated.F90:150-185. */
static void scales_compute(const double *P, const double *np, double t, const double *y, double *sc)
{
  orc_halo h;
  halo_compute(P, np, t, y, &h);
  const double mrel = 1.0e-3, rrel = 1.0e-1;
  double mv = h.mass, rv = h.r_vir, vv = h.v_vir;
  for (int i = 0; i < NV; i++) sc[i] = 1.0;
  sc[GAL_V_HOT_MASS] = mv * mrel;
  sc[GAL_V_HOT_OUTFLOWED_MASS] = mv * mrel;
  sc[GAL_V_HOT_UNACCRETED_MASS] = mv * mrel;
  sc[GAL_V_HOT_ABUNDANCES] = mv * mrel;
  sc[GAL_V_HOT_UNACCRETED_ABUNDANCES] = mv * mrel;
  sc[GAL_V_HOT_OUTFLOWED_ABUNDANCES] = mv * mrel;
  sc[GAL_V_HOT_ANGULAR_MOMENTUM] = mv * rv * vv * mrel;
  sc[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] = mv * rv * vv * mrel;
  sc[GAL_V_HOT_OUTER_RADIUS] = rv * rrel;
  sc[GAL_V_HOT_STRIPPED_MASS] = mv * mrel;
  sc[GAL_V_HOT_STRIPPED_ABUNDANCES] = mv * mrel;
  double unit_energy = C_MASS_SOLAR * C_KILO * C_KILO / C_ERGS / C_GIGAYEAR;
  sc[GAL_V_HOT_ENERGY_RADIATED] = unit_energy * mv * vv * vv * 1.0e-3;
  double ang = fabs(y[GAL_V_DISK_ANGULAR_MOMENTUM]) + fabs(y[GAL_V_SPHEROID_ANGULAR_MOMENTUM]);
  if (ang < 0.1) ang = 0.1;
  double mass = fabs(y[GAL_V_DISK_MASS_GAS]) + fabs(y[GAL_V_SPHEROID_MASS_GAS])
              + fabs(y[GAL_V_DISK_MASS_STELLAR]) + fabs(y[GAL_V_SPHEROID_MASS_STELLAR]);
  if (mass < 1.0) mass = 1.0;
  double zmin = mass * P[GAL_P_TOLERANCE_RELATIVE_METALLICITY];
  if (zmin < 1.0) zmin = 1.0;
  double zsc = fabs(y[GAL_V_DISK_ABUNDANCES_GAS]) + fabs(y[GAL_V_DISK_ABUNDANCES_STELLAR])
             + fabs(y[GAL_V_SPHEROID_ABUNDANCES_GAS]) + fabs(y[GAL_V_SPHEROID_ABUNDANCES_STELLAR]);
  if (zsc < zmin) zsc = zmin;
  sc[GAL_V_DISK_ANGULAR_MOMENTUM] = ang;
  sc[GAL_V_DISK_MASS_GAS] = mass;
  sc[GAL_V_DISK_MASS_STELLAR] = mass;
  sc[GAL_V_DISK_MASS_STELLAR_FORMED] = mass;
  sc[GAL_V_DISK_FRACTION_MASS_RETAINED] = 1.0e-4 * y[GAL_V_DISK_FRACTION_MASS_RETAINED];
  sc[GAL_V_DISK_ABUNDANCES_GAS] = zsc;
  sc[GAL_V_DISK_ABUNDANCES_STELLAR] = zsc;
  sc[GAL_V_SPHEROID_ANGULAR_MOMENTUM] = ang;
  sc[GAL_V_SPHEROID_MASS_GAS] = mass;
  sc[GAL_V_SPHEROID_MASS_STELLAR] = mass;
  sc[GAL_V_SPHEROID_MASS_STELLAR_FORMED] = mass;
  sc[GAL_V_SPHEROID_ABUNDANCES_GAS] = zsc;
  sc[GAL_V_SPHEROID_ABUNDANCES_STELLAR] = zsc;
  double bm = 1.0;
  if (1.0e-4 * y[GAL_V_SPHEROID_MASS_STELLAR] > bm) bm = 1.0e-4 * y[GAL_V_SPHEROID_MASS_STELLAR];
  if (y[GAL_V_BH_MASS] > bm) bm = y[GAL_V_BH_MASS];
  sc[GAL_V_BH_MASS] = bm;
  sc[GAL_V_BH_SPIN] = 1.0;
  double rp = 1.0e-4 * h.r_sph;
  sc[GAL_V_BH_RADIAL_POSITION] = (rp > 1.0e-6) ? rp : 1.0e-6;
  sc[GAL_V_SPIN_ANGULAR_MOMENTUM] = mv * rv * vv * mrel;
  sc[GAL_V_SATELLITE_BOUND_MASS] = mv * mrel;
}

/* isNonNegative attributes of the component properties (_class.F90 XML). */
static void nonneg_flags(int *nn)
{
  for (int i = 0; i < NV; i++) nn[i] = 1;
  nn[GAL_V_HOT_ANGULAR_MOMENTUM] = 0;
  nn[GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM] = 0;
  nn[GAL_V_HOT_ENERGY_RADIATED] = 0;
  nn[GAL_V_BH_RADIAL_POSITION] = 0;
  nn[GAL_V_SPIN_ANGULAR_MOMENTUM] = 0;
}

/* ---------------------------------------------------------------------- */
/* Post-step: standardPostStepProcessing (standard.F90:1160-1185) calling the
   component post-steps.  Returns ST_SUCCESS or ST_CONTINUE (state changed,
   no evolver reset). */
static void component_post_step(double *y, int i_mstar, int i_zstar, int i_mgas, int i_zgas, int i_ang,
                                double radius, double velocity, double spin_j, double halo_mass, int *status)
{
  /* disk/standard/_class.F90:506-660 (spheroid/standard/_class.F90 is the same) */
  if (y[i_mgas] < 0.0) {
    double m = y[i_mgas] + y[i_mstar], j;
    if (m == 0.0) { j = 0.0; y[i_mstar] = 0.0; y[i_zstar] = 0.0; }
    else { j = y[i_ang] / m; if (j < 0.0) j = radius * velocity; }
    y[i_mgas] = 0.0; y[i_zgas] = 0.0; y[i_ang] = j * y[i_mstar];
    if (*status == ST_SUCCESS) *status = ST_CONTINUE;
  }
  if (y[i_mstar] < 0.0) {
    double m = y[i_mgas] + y[i_mstar], j;
    if (m == 0.0) { j = 0.0; y[i_mgas] = 0.0; y[i_zgas] = 0.0; }
    else { j = y[i_ang] / m; if (j < 0.0) j = radius * velocity; }
    y[i_mstar] = 0.0; y[i_zstar] = 0.0; y[i_ang] = j * y[i_mgas];
    if (*status == ST_SUCCESS) *status = ST_CONTINUE;
  }
  if (y[i_ang] < 0.0) {
    /* both branches of disk/standard/_class.F90:641-660 end with zero J
       (the reference aborts in the large-|J| branch) */
    (void)spin_j; (void)halo_mass;
    y[i_ang] = 0.0;
    if (*status == ST_SUCCESS) *status = ST_CONTINUE;
  }
}

static int post_step(orc_sys *s, double t, double *y)
{
  int status = ST_SUCCESS;
  int mode = (int)s->P[GAL_P_PHYSICS_MODE];
  if (mode == GAL_PHYSICS_TEST) return status;
  orc_halo h;
  halo_compute(s->P, s->np, t, y, &h);
  /* postStepZeroNegativeMasses (disk/standard/_class.F90, spheroid/standard/_class.F90) */
  if (s->P[GAL_P_POST_STEP_ZERO_NEGATIVE_MASSES] != 0.0) {
  component_post_step(y, GAL_V_DISK_MASS_STELLAR, GAL_V_DISK_ABUNDANCES_STELLAR, GAL_V_DISK_MASS_GAS,
                      GAL_V_DISK_ABUNDANCES_GAS, GAL_V_DISK_ANGULAR_MOMENTUM, h.r_disk, h.v_disk,
                      y[GAL_V_SPIN_ANGULAR_MOMENTUM], h.mass, &status);
  component_post_step(y, GAL_V_SPHEROID_MASS_STELLAR, GAL_V_SPHEROID_ABUNDANCES_STELLAR, GAL_V_SPHEROID_MASS_GAS,
                      GAL_V_SPHEROID_ABUNDANCES_GAS, GAL_V_SPHEROID_ANGULAR_MOMENTUM, h.r_sph, h.v_sph,
                      y[GAL_V_SPIN_ANGULAR_MOMENTUM], h.mass, &status);
  }
  /* hot_halo/standard/_class.F90:468-475 */
  if (y[GAL_V_HOT_MASS] < 0.0) { y[GAL_V_HOT_MASS] = 0.0; if (status == ST_SUCCESS) status = ST_CONTINUE; }
  if (y[GAL_V_HOT_OUTER_RADIUS] < 0.0) { y[GAL_V_HOT_OUTER_RADIUS] = 0.0; if (status == ST_SUCCESS) status = ST_CONTINUE; }
  return status;
}

/* ---------------------------------------------------------------------- */
/* Controller: sc2_control_hadjust, cscal2.c:93-169. */
int orc_hadjust(int dim, int ord, double eps_abs, double eps_rel, double a_y, double a_dydt,
                const double *y, const double *yerr, const double *yp, const double *scale_abs,
                const int *is_non_negative, double *h)
{
  const double S = 0.9;
  const double h_old = *h;
  double rmax = DBL_MIN;
  int forbidden = 0;
  for (int i = 0; i < dim; i++) {
    const double D0 = eps_rel * (a_y * fabs(y[i]) + a_dydt * fabs(h_old * yp[i])) + eps_abs * scale_abs[i];
    const double r = fabs(yerr[i]) / fabs(D0);
    rmax = (r > rmax) ? r : rmax;          /* GSL_MAX_DBL */
    if (is_non_negative[i] && y[i] < 0.0) forbidden = 1;
  }
  if (rmax > 1.1) {
    double r = S / pow(rmax, 1.0 / ord);
    if (r < 0.2) r = 0.2;
    *h = r * h_old;
    return HADJ_DEC;
  } else if (forbidden == 1) {
    *h = 0.5 * h_old;
    return HADJ_DEC;
  } else if (rmax < 0.5) {
    double r = S / pow(rmax, 1.0 / (ord + 1.0));
    if (r > 4.9) r = 4.9;
    if (r < 1.0) r = 1.0;
    *h = r * h_old;
    return HADJ_INC;
  }
  return HADJ_NIL;
}

/* ---------------------------------------------------------------------- */
/* Cash-Karp (4,5) stepper, Cash & Karp (1990) ACM TOMS 16:201 Table 5,
   with GSL rkck semantics: k1 from dydt_in, dydt_out evaluated at the new
   point, error = 5th-order minus embedded 4th-order estimate; order 4.    */
static const double ck_a2 = 1.0 / 5.0, ck_a3 = 3.0 / 10.0, ck_a4 = 3.0 / 5.0, ck_a5 = 1.0, ck_a6 = 7.0 / 8.0;
static const double ck_b21 = 1.0 / 5.0;
static const double ck_b31 = 3.0 / 40.0, ck_b32 = 9.0 / 40.0;
static const double ck_b41 = 3.0 / 10.0, ck_b42 = -9.0 / 10.0, ck_b43 = 6.0 / 5.0;
static const double ck_b51 = -11.0 / 54.0, ck_b52 = 5.0 / 2.0, ck_b53 = -70.0 / 27.0, ck_b54 = 35.0 / 27.0;
static const double ck_b61 = 1631.0 / 55296.0, ck_b62 = 175.0 / 512.0, ck_b63 = 575.0 / 13824.0,
                    ck_b64 = 44275.0 / 110592.0, ck_b65 = 253.0 / 4096.0;
static const double ck_c1 = 37.0 / 378.0, ck_c3 = 250.0 / 621.0, ck_c4 = 125.0 / 594.0, ck_c6 = 512.0 / 1771.0;
#define CK_EC1 (37.0 / 378.0 - 2825.0 / 27648.0)
#define CK_EC3 (250.0 / 621.0 - 18575.0 / 48384.0)
#define CK_EC4 (125.0 / 594.0 - 13525.0 / 55296.0)
#define CK_EC5 (-277.0 / 14336.0)
#define CK_EC6 (512.0 / 1771.0 - 1.0 / 4.0)
#define RKCK_ORDER 4

typedef int (*rhs_fn)(void *ctx, double t, const double *y, double *dydt);

typedef struct {
  int dim;
  rhs_fn f;
  void *ctx;
  int (*post)(void *ctx, double t, double *y);
  /* evolve object */
  int count;
  double y0[NV], yerr[NV], dydt_in[NV], dydt_out[NV];
  double last_step;
  /* driver */
  double h;
  /* control */
  double eps_abs, eps_rel;
  const double *scale;
  const int *nonneg;
  int64_t n_attempt, n_accept;
} orc_driver;

static int rkck_apply(orc_driver *d, double t, double h, double *y, double *yerr,
                      const double *dydt_in, double *dydt_out)
{
  const int n = d->dim;
  double k1[NV], k2[NV], k3[NV], k4[NV], k5[NV], k6[NV], ytmp[NV] = {0}, y0[NV];
  int st;
  memcpy(y0, y, sizeof(double) * n);
  memcpy(k1, dydt_in, sizeof(double) * n);
  for (int i = 0; i < n; i++) ytmp[i] = y[i] + h * (ck_b21 * k1[i]);
  if ((st = d->f(d->ctx, t + ck_a2 * h, ytmp, k2))) return st;
  for (int i = 0; i < n; i++) ytmp[i] = y[i] + h * (ck_b31 * k1[i] + ck_b32 * k2[i]);
  if ((st = d->f(d->ctx, t + ck_a3 * h, ytmp, k3))) return st;
  for (int i = 0; i < n; i++) ytmp[i] = y[i] + h * (ck_b41 * k1[i] + ck_b42 * k2[i] + ck_b43 * k3[i]);
  if ((st = d->f(d->ctx, t + ck_a4 * h, ytmp, k4))) return st;
  for (int i = 0; i < n; i++) ytmp[i] = y[i] + h * (ck_b51 * k1[i] + ck_b52 * k2[i] + ck_b53 * k3[i] + ck_b54 * k4[i]);
  if ((st = d->f(d->ctx, t + ck_a5 * h, ytmp, k5))) return st;
  for (int i = 0; i < n; i++) ytmp[i] = y[i] + h * (ck_b61 * k1[i] + ck_b62 * k2[i] + ck_b63 * k3[i] + ck_b64 * k4[i] + ck_b65 * k5[i]);
  if ((st = d->f(d->ctx, t + ck_a6 * h, ytmp, k6))) return st;
  for (int i = 0; i < n; i++) y[i] = y[i] + h * (ck_c1 * k1[i] + ck_c3 * k3[i] + ck_c4 * k4[i] + ck_c6 * k6[i]);
  if ((st = d->f(d->ctx, t + h, y, dydt_out))) { memcpy(y, y0, sizeof(double) * n); return st; }
  for (int i = 0; i < n; i++)
    yerr[i] = h * (CK_EC1 * k1[i] + CK_EC3 * k3[i] + CK_EC4 * k4[i] + CK_EC5 * k5[i] + CK_EC6 * k6[i]);
  return ST_SUCCESS;
}

/* gsl_odeiv2_evolve_apply (GSL 2.6 ode-initval2/evolve.c), restated. */
static int evolve_apply(orc_driver *d, double *t, double t1, double *h, double *y)
{
  const double t0 = *t;
  double h0 = *h;
  int final_step = 0;
  const double dt = t1 - t0;
  const int n = d->dim;
  if ((dt < 0.0 && h0 > 0.0) || (dt > 0.0 && h0 < 0.0)) return GAL_STATUS_FAILURE;  /* GSL_EINVAL in GSL */
  memcpy(d->y0, y, sizeof(double) * n);
  if (d->count == 0) {
    int st = d->f(d->ctx, t0, y, d->dydt_in);
    if (st) return st;
  } else {
    memcpy(d->dydt_in, d->dydt_out, sizeof(double) * n);
  }
try_step:
  if ((dt >= 0.0 && h0 > dt) || (dt < 0.0 && h0 < dt)) { h0 = dt; final_step = 1; }
  else final_step = 0;
  d->n_attempt++;
  {
    int st = rkck_apply(d, t0, h0, y, d->yerr, d->dydt_in, d->dydt_out);
    if (st == GAL_STATUS_INTERRUPT) return st;
    if (st != ST_SUCCESS) {
      /* stepper failure: halve the step (evolve.c) */
      double h_old = h0;
      h0 *= 0.5;
      volatile double t_curr = *t, t_next = *t + h0;
      if (fabs(h0) < fabs(h_old) && t_next != t_curr) {
        memcpy(y, d->y0, sizeof(double) * n);
        goto try_step;
      }
      *h = h0; *t = t0;
      return st;
    }
  }
  d->count++;
  d->last_step = h0;
  if (final_step) *t = t1; else *t = t0 + h0;
  {
    double h_old = h0;
    int hs = orc_hadjust(n, RKCK_ORDER, d->eps_abs, d->eps_rel, 1.0, 0.0, y, d->yerr, d->dydt_out,
                         d->scale, d->nonneg, &h0);
    if (hs == HADJ_DEC) {
      volatile double t_curr = *t, t_next = *t + h0;
      if (fabs(h0) < fabs(h_old) && t_next != t_curr) {
        memcpy(y, d->y0, sizeof(double) * n);
        goto try_step;
      }
      /* step size can not be reduced further: restore and fail */
      memcpy(y, d->y0, sizeof(double) * n);
      *t = t0;
      return GAL_STATUS_FAILURE;
    }
  }
  if (final_step == 0) *h = h0;
  d->n_accept++;
  return ST_SUCCESS;
}

/* gsl_odeiv2_driver2_apply (driver2.c:148-250) with hmin=0, hmax=DBL_MAX,
   nmax=0 (driver2.c:93-97) and no step analyzer / latent integrator. */
static int driver2_apply(orc_driver *d, double *t, double t1, double *y)
{
  int sign = (d->h > 0.0) ? 1 : -1;
  if (sign * (t1 - *t) < 0.0) return GAL_STATUS_FAILURE;
  while (sign * (t1 - *t) > 0.0) {
    int s = evolve_apply(d, t, t1, &d->h, y);
    if (s != ST_SUCCESS) return s;
    if (d->post) {
      int ps = d->post(d->ctx, *t, y);
      if (ps != ST_SUCCESS && ps != ST_CONTINUE) d->count = 0;   /* gsl_odeiv2_evolve_reset */
    }
  }
  return ST_SUCCESS;
}

/* odeSolverSolve (solver.F90:492-691) without interrupts or latent variables. */
static int ode_solve(orc_driver *d, double *x0, double x1, double *y, double *xstep)
{
  double xs = x1 - *x0;
  if (xstep && *xstep > 0.0) xs = (*xstep < xs) ? *xstep : xs;
  double x = *x0;
  int forward = x1 > *x0;
  d->count = 0;                      /* GSL_ODEIV2_Driver_Reset */
  if (xs != 0.0) d->h = xs;          /* GSL_ODEIV2_Driver_Reset_hStart */
  int status = ST_SUCCESS;
  while ((forward && x < x1) || (!forward && x > x1)) {
    status = driver2_apply(d, &x, x1, y);
    if (status == ST_SUCCESS) {
      if (xstep) *xstep = d->h;
    } else {
      *x0 = x;
      return status;
    }
  }
  *x0 = x;
  return status;
}

/* ---------------------------------------------------------------------- */
static int sys_rhs(void *ctx, double t, const double *y, double *dydt)
{
  return rhs_eval((orc_sys *)ctx, t, y, dydt);
}
static int sys_post(void *ctx, double t, double *y)
{
  return post_step((orc_sys *)ctx, t, y);
}

/* standardEvolve trial loop (node_evolver/standard.F90:587-740). */
static int node_evolve(orc_sys *s, double t_start, double t_end, double *y, double *h_node,
                       int flags, int64_t *cnt)
{
  const int trial_max = 8;
  double y_saved[NV];
  memcpy(y_saved, y, sizeof(y_saved));
  double step_size = -1.0;
  int status = ST_SUCCESS;
  orc_driver d;
  memset(&d, 0, sizeof(d));
  d.dim = NV; d.f = sys_rhs; d.ctx = s; d.post = sys_post;
  d.eps_abs = s->eps_abs; d.eps_rel = s->eps_rel; d.scale = s->scale; d.nonneg = s->nonneg;
  d.h = 1.0;                         /* hStart default (solver.F90:368) */
  if (t_start != t_end) {
    int trial = 0;
    status = ST_FAILURE;
    while (trial < trial_max && !(status == ST_SUCCESS || status == GAL_STATUS_INTERRUPT)) {
      if (flags & GAL_FLAG_REUSE_STEP_SIZE) step_size = *h_node / pow(2.0, trial);
      else step_size = -1.0;
      double t = t_start;
      status = ode_solve(&d, &t, t_end, y, &step_size);
      if ((flags & GAL_FLAG_ENFORCE_NON_NEGATIVITY) && !(status == ST_SUCCESS || status == GAL_STATUS_INTERRUPT)) {
        int any = 0;
        for (int i = 0; i < NV; i++) if (y[i] < 0.0 && s->nonneg[i]) any = 1;
        if (any) {
          for (int i = 0; i < NV; i++) if (y[i] < 0.0 && s->nonneg[i]) y[i] = 0.0;
          step_size = *h_node / pow(2.0, trial);
          status = ST_SUCCESS;
        }
      }
      if (!(status == ST_SUCCESS || status == GAL_STATUS_INTERRUPT)) {
        trial++;
        memcpy(y, y_saved, sizeof(y_saved));
      }
    }
    if (!(status == ST_SUCCESS || status == GAL_STATUS_INTERRUPT)) status = GAL_STATUS_UNDERFLOW;
  }
  if (status == ST_SUCCESS) *h_node = step_size;
  if (cnt) { cnt[0] = d.n_attempt; cnt[1] = d.n_accept; cnt[2] = s->n_rhs; }
  return status;
}

/* ---------------------------------------------------------------------- */
int orc_evolve_batch(int n, int ld, double *y, const double *np, const double *t_start,
                     const double *t_end, double *h_step, double *scale, int *status,
                     int64_t *counters, const double *params, const orc_cooling *table,
                     double tol_abs, double tol_rel, int flags, double root_tol_rel, int nthreads)
{
  if (n < 0 || ld < n) return GAL_E_INVALID;
  int nbad = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads) reduction(+:nbad)
#endif
  for (int i = 0; i < n; i++) {
    orc_sys s;
    memset(&s, 0, sizeof(s));
    s.P = params; s.ct = table; s.eps_abs = tol_abs; s.eps_rel = tol_rel; s.root_tol = root_tol_rel;
    double yr[NV];
    for (int v = 0; v < NV; v++) yr[v] = y[(size_t)v * ld + i];
    for (int p = 0; p < GAL_NPARAM_NODE; p++) s.np[p] = np[(size_t)p * ld + i];
    if (flags & GAL_FLAG_SCALES_PROVIDED) {
      for (int v = 0; v < NV; v++) s.scale[v] = scale[(size_t)v * ld + i];
    } else {
      if ((int)params[GAL_P_PHYSICS_MODE] == GAL_PHYSICS_TEST) for (int v = 0; v < NV; v++) s.scale[v] = 1.0;
      else scales_compute(params, s.np, t_start[i], yr, s.scale);
      if (scale) for (int v = 0; v < NV; v++) scale[(size_t)v * ld + i] = s.scale[v];
    }
    if (flags & GAL_FLAG_ENFORCE_NON_NEGATIVITY) nonneg_flags(s.nonneg);
    else for (int v = 0; v < NV; v++) s.nonneg[v] = 0;
    double hn = h_step[i];
    int64_t cnt[3];
    int st = node_evolve(&s, t_start[i], t_end[i], yr, &hn, flags, cnt);
    for (int v = 0; v < NV; v++) y[(size_t)v * ld + i] = yr[v];
    h_step[i] = hn;
    status[i] = st;
    if (counters) { counters[i] = cnt[0]; counters[(size_t)n + i] = cnt[1]; counters[2 * (size_t)n + i] = cnt[2]; }
    if (st != ST_SUCCESS) nbad++;
  }
  return nbad;
}

/* ---------------------------------------------------------------------- */
/* Tree schedules (the sequential restatement of oracle/trees.py
   run_sequential + merge_galaxy_simple, which follow the walk of
   merger_trees/evolver/standard.F90:291-1035 and the mass movements of
   satellites/merging/mass_movements/simple.F90:208-290). */
static void merge_galaxy_simple_c(double *y, int ld, int sat, int host, double major_ratio)
{
#define Y(v, s) y[(size_t)(v) * ld + (s)]
  const double ms = Y(GAL_V_DISK_MASS_GAS, sat) + Y(GAL_V_DISK_MASS_STELLAR, sat) +
                    Y(GAL_V_SPHEROID_MASS_GAS, sat) + Y(GAL_V_SPHEROID_MASS_STELLAR, sat);
  const double mh = Y(GAL_V_DISK_MASS_GAS, host) + Y(GAL_V_DISK_MASS_STELLAR, host) +
                    Y(GAL_V_SPHEROID_MASS_GAS, host) + Y(GAL_V_SPHEROID_MASS_STELLAR, host);
  const int major = ms > 0.0 && mh > 0.0 && fmin(ms, mh) >= major_ratio * fmax(ms, mh);
  static const int to_sph[6][2] = {
      {GAL_V_DISK_MASS_GAS, GAL_V_SPHEROID_MASS_GAS}, {GAL_V_DISK_ABUNDANCES_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS},
      {GAL_V_DISK_MASS_STELLAR, GAL_V_SPHEROID_MASS_STELLAR},
      {GAL_V_DISK_ABUNDANCES_STELLAR, GAL_V_SPHEROID_ABUNDANCES_STELLAR},
      {GAL_V_DISK_MASS_STELLAR_FORMED, GAL_V_SPHEROID_MASS_STELLAR_FORMED},
      {GAL_V_DISK_ANGULAR_MOMENTUM, GAL_V_SPHEROID_ANGULAR_MOMENTUM}};
  static const int same[16] = {
      GAL_V_SPHEROID_MASS_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS, GAL_V_SPHEROID_MASS_STELLAR,
      GAL_V_SPHEROID_ABUNDANCES_STELLAR, GAL_V_SPHEROID_MASS_STELLAR_FORMED, GAL_V_SPHEROID_ANGULAR_MOMENTUM,
      GAL_V_HOT_MASS, GAL_V_HOT_ABUNDANCES, GAL_V_HOT_ANGULAR_MOMENTUM, GAL_V_HOT_OUTFLOWED_MASS,
      GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM, GAL_V_HOT_OUTFLOWED_ABUNDANCES, GAL_V_HOT_UNACCRETED_MASS,
      GAL_V_HOT_UNACCRETED_ABUNDANCES, GAL_V_HOT_ENERGY_RADIATED, GAL_V_BH_MASS};
  for (int k = 0; k < 6; k++) {
    Y(to_sph[k][1], host) += Y(to_sph[k][0], sat);
    Y(to_sph[k][0], sat) = 0.0;
  }
  for (int k = 0; k < 16; k++) {
    Y(same[k], host) += Y(same[k], sat);
    Y(same[k], sat) = 0.0;
  }
  /* stripped gas joins the host's outflowed reservoir (mass and metals;
     hot_halo/standard/_class.F90:532-597 moves it after each evolve) */
  Y(GAL_V_HOT_OUTFLOWED_MASS, host) += Y(GAL_V_HOT_STRIPPED_MASS, sat);
  Y(GAL_V_HOT_STRIPPED_MASS, sat) = 0.0;
  Y(GAL_V_HOT_OUTFLOWED_ABUNDANCES, host) += Y(GAL_V_HOT_STRIPPED_ABUNDANCES, sat);
  Y(GAL_V_HOT_STRIPPED_ABUNDANCES, sat) = 0.0;
  if (major) {
    for (int k = 0; k < 6; k++) {
      Y(to_sph[k][1], host) += Y(to_sph[k][0], host);
      Y(to_sph[k][0], host) = 0.0;
    }
  }
#undef Y
}

int orc_tree_run(int64_t n_ops, const int *kind, const int *a, const int *b, const double *t0, const double *t1,
                 const double *node_par, int ld, const int *tree_of_node, int n_trees, double *y, double *h,
                 int *status, const double *params, const orc_cooling *table, double tol_abs, double tol_rel,
                 int flags, double root_tol_rel, double major_ratio, int nthreads)
{
  if (n_ops < 0 || ld < 1 || n_trees < 1) return GAL_E_INVALID;
  for (int64_t k = 0; k < n_ops; k++) {
    if (a[k] < 0 || a[k] >= ld || b[k] >= ld) return GAL_E_INVALID;
    const int t = tree_of_node[a[k]];
    if (t < 0 || t >= n_trees) return GAL_E_INVALID;
  }
  /* operations of each tree, in schedule order */
  int64_t *start = calloc((size_t)n_trees + 1, sizeof(int64_t));
  int64_t *list = malloc(sizeof(int64_t) * (size_t)(n_ops > 0 ? n_ops : 1));
  int64_t *fill = malloc(sizeof(int64_t) * (size_t)n_trees);
  if (!start || !list || !fill) { free(start); free(list); free(fill); return GAL_E_NOMEM; }
  for (int64_t k = 0; k < n_ops; k++) start[tree_of_node[a[k]] + 1]++;
  for (int t = 0; t < n_trees; t++) start[t + 1] += start[t];
  for (int t = 0; t < n_trees; t++) fill[t] = start[t];
  for (int64_t k = 0; k < n_ops; k++) list[fill[tree_of_node[a[k]]]++] = k;
  int nbad = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+:nbad)
#endif
  for (int t = 0; t < n_trees; t++) {
    for (int64_t j = start[t]; j < start[t + 1]; j++) {
      const int64_t k = list[j];
      const int na = a[k];
      if (kind[k] == GAL_OP_EVOLVE) {
        orc_sys s;
        memset(&s, 0, sizeof(s));
        s.P = params; s.ct = table; s.eps_abs = tol_abs; s.eps_rel = tol_rel; s.root_tol = root_tol_rel;
        double yr[NV];
        for (int v = 0; v < NV; v++) yr[v] = y[(size_t)v * ld + na];
        for (int p = 0; p < GAL_NPARAM_NODE; p++) s.np[p] = node_par[(size_t)GAL_NPARAM_NODE * k + p];
        scales_compute(params, s.np, t0[k], yr, s.scale);
        if (flags & GAL_FLAG_ENFORCE_NON_NEGATIVITY) nonneg_flags(s.nonneg);
        double hn = h[na];
        const int st = node_evolve(&s, t0[k], t1[k], yr, &hn, flags, NULL);
        for (int v = 0; v < NV; v++) y[(size_t)v * ld + na] = yr[v];
        h[na] = hn;
        status[na] = st;
        if (st != ST_SUCCESS) nbad++;
      } else if (kind[k] == GAL_OP_PROMOTE) {
        for (int v = 0; v < NV; v++) y[(size_t)v * ld + b[k]] = y[(size_t)v * ld + na];
        h[b[k]] = h[na];
      } else if (kind[k] == GAL_OP_MERGE_GALAXY) {
        merge_galaxy_simple_c(y, ld, na, b[k], major_ratio);
      }
    }
  }
  free(start); free(list); free(fill);
  return nbad;
}

int orc_rhs_batch(int n, int ld, const double *t, const double *y, const double *np, double *dydt,
                  const double *params, const orc_cooling *table, double root_tol_rel)
{
  int nbad = 0;
  for (int i = 0; i < n; i++) {
    orc_sys s;
    memset(&s, 0, sizeof(s));
    s.P = params; s.ct = table; s.root_tol = root_tol_rel;
    double yr[NV], dr[NV];
    for (int v = 0; v < NV; v++) yr[v] = y[(size_t)v * ld + i];
    for (int p = 0; p < GAL_NPARAM_NODE; p++) s.np[p] = np[(size_t)p * ld + i];
    if (rhs_eval(&s, t[i], yr, dr) != ST_SUCCESS) nbad++;
    for (int v = 0; v < NV; v++) dydt[(size_t)v * ld + i] = dr[v];
  }
  return nbad;
}

void orc_scales_batch(int n, int ld, const double *t, const double *y, const double *np, double *scale,
                      const double *params)
{
  for (int i = 0; i < n; i++) {
    double yr[NV], sc[NV], npr[GAL_NPARAM_NODE];
    for (int v = 0; v < NV; v++) yr[v] = y[(size_t)v * ld + i];
    for (int p = 0; p < GAL_NPARAM_NODE; p++) npr[p] = np[(size_t)p * ld + i];
    scales_compute(params, npr, t[i], yr, sc);
    for (int v = 0; v < NV; v++) scale[(size_t)v * ld + i] = sc[v];
  }
}

void orc_halo_properties(const double *params, const orc_cooling *table, const double *np_row, double t,
                         const double *y_row, double root_tol_rel, double out[8])
{
  orc_sys s;
  memset(&s, 0, sizeof(s));
  s.P = params; s.ct = table; s.root_tol = root_tol_rel;
  memcpy(s.np, np_row, sizeof(s.np));
  orc_halo h;
  orc_hot hh;
  halo_compute(params, s.np, t, y_row, &h);
  hot_compute(&s, &h, y_row, &hh);
  out[0] = h.r_vir; out[1] = h.v_vir; out[2] = h.t_dyn; out[3] = h.t_vir; out[4] = h.a;
  out[5] = hh.r_cool; out[6] = hh.r_cool_rate; out[7] = hh.t_avail;
}

/* Test systems of source/tests/ODE_solver/functions.F90 through the same
   solve / driver / evolve / rkck / controller path, scale = 1. */
typedef struct { int which; int64_t n_rhs; } test_ctx;
static int test_rhs(void *ctx, double t, const double *y, double *dydt)
{
  test_ctx *c = (test_ctx *)ctx;
  c->n_rhs++;
  if (c->which == 1) dydt[0] = sin(t);
  else if (c->which == 3) dydt[0] = y[0];
  else { dydt[0] = y[1]; dydt[1] = -1.0 * y[0]; }
  return ST_SUCCESS;
}

int orc_test_rkck_step(int which, double t, double h, double *y, double *yerr)
{
  test_ctx c = { which, 0 };
  orc_driver d;
  memset(&d, 0, sizeof(d));
  d.dim = (which == 2) ? 2 : 1;
  d.f = test_rhs; d.ctx = &c;
  double dydt_in[NV], dydt_out[NV];
  int st = test_rhs(&c, t, y, dydt_in);
  if (st) return st;
  return rkck_apply(&d, t, h, y, yerr, dydt_in, dydt_out);
}

int orc_test_ode_solve(int which, double x0, double x1, double *y, double tol_abs, double tol_rel,
                       double *h_out, int64_t counters[3])
{
  static const double ones[NV] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                  1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static const int zeros[NV] = {0};
  test_ctx c = { which, 0 };
  orc_driver d;
  memset(&d, 0, sizeof(d));
  d.dim = (which == 1) ? 1 : 2;
  d.f = test_rhs; d.ctx = &c; d.post = NULL;
  d.eps_abs = tol_abs; d.eps_rel = tol_rel; d.scale = ones; d.nonneg = zeros;
  d.h = 1.0;
  double x = x0;
  int st = ode_solve(&d, &x, x1, y, NULL);
  if (h_out) *h_out = d.h;
  if (counters) { counters[0] = d.n_attempt; counters[1] = d.n_accept; counters[2] = c.n_rhs; }
  return st;
}
