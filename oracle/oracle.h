/*
 * oracle.h -- CPU restatement of the reference's node-ODE path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the MI355X engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates; see oracle.c.
 * Array conventions follow include/gal_ode.h (structure-of-arrays with a
 * leading dimension) so results can be compared element for element.
 */
#ifndef GAL_ORACLE_H
#define GAL_ORACLE_H
#include <stdint.h>
#include "../include/gal_ode.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const double *ln_lambda;   /* [n_z * n_t] */
  int n_t, n_z;
  double ln_t_min, d_ln_t, ln_z_min, d_ln_z;
} orc_cooling;

/* Evolve n nodes (SoA, leading dimension ld).  Same argument meaning as
   gal_ode_batch_evolve_d.  root_tol_rel: relative tolerance of the
   cooling-radius root finder (the reference uses 1e-6,
   cooling/cooling_radius/simple.F90:133).  nthreads: OpenMP threads.     */
int orc_evolve_batch(int n, int ld, double *y, const double *node_par,
                     const double *t_start, const double *t_end,
                     double *h_step, double *scale, int *status,
                     int64_t *counters, const double *params,
                     const orc_cooling *table, double tol_abs,
                     double tol_rel, int flags, double root_tol_rel,
                     int nthreads);

/* RHS of n nodes at times t (SoA). Returns number of non-success statuses. */
int orc_rhs_batch(int n, int ld, const double *t, const double *y,
                  const double *node_par, double *dydt, const double *params,
                  const orc_cooling *table, double root_tol_rel);

/* Default component scales (scale-set rules) for n nodes at t (SoA). */
void orc_scales_batch(int n, int ld, const double *t, const double *y,
                      const double *node_par, double *scale,
                      const double *params);

/* The scaled controller on one case (cscal2.c:93-169). Returns -1/0/+1. */
int orc_hadjust(int dim, int ord, double eps_abs, double eps_rel, double a_y,
                double a_dydt, const double *y, const double *yerr,
                const double *yp, const double *scale,
                const int *is_non_negative, double *h);

/* Cooling-function lookup Lambda(T, Z/Zsun) [erg cm^3 s^-1]. */
double orc_cooling_function(const orc_cooling *table, double temperature,
                            double metallicity_solar);

/* Cooling time [Gyr] of hot gas (simpleTime, cooling_time/simple.F90:128-181). */
double orc_cooling_time(const double *params, const orc_cooling *table,
                        double temperature, double density,
                        double metal_fraction);

/* Halo quantities at time t for one node (virial radius [Mpc], virial
   velocity [km/s], dynamical time [Gyr], virial temperature [K], scale
   factor) and hot-halo cooling radius [Mpc] and its growth rate [Mpc/Gyr]. */
void orc_halo_properties(const double *params, const orc_cooling *table,
                         const double *node_par_row, double t,
                         const double *y_row, double root_tol_rel,
                         double out[8]);

/* Apply a tree schedule on the CPU: operations one at a time in schedule
   order within each tree (trees are independent and run in parallel over
   nthreads).  kind/a/b/t0/t1/node_par[5*op] as gal_schedule_op of
   include/gal_tree.h; tree_of_node[ld] names each node's tree (0..n_trees-1).
   Evolve: orc_evolve of node a over [t0, t1] with step size h[a]; promote:
   y[:, b] = y[:, a], h[b] = h[a]; galaxy merger: the simple mass movements
   (satellite a into host b, spheroid destinations, major if the smaller
   galaxy is at least major_ratio of the larger); becoming a satellite moves
   no mass.  status[a] is the last evolve's status.  Returns the number of
   non-success evolves.                                                      */
int orc_tree_run(int64_t n_ops, const int *kind, const int *a, const int *b,
                 const double *t0, const double *t1, const double *node_par,
                 int ld, const int *tree_of_node, int n_trees, double *y,
                 double *h, int *status, const double *params,
                 const orc_cooling *table, double tol_abs, double tol_rel,
                 int flags, double root_tol_rel, double major_ratio,
                 int nthreads);

/* Generic fixed-dimension Cash-Karp solve of a built-in test system used to
   pin the integrator against closed forms (source/tests/ODE_solver.F90:59-104).
   which: 1 = y'=sin(x) ; 2 = y''=-y (y1'=y2, y2'=-y1).                   */
int orc_test_ode_solve(int which, double x0, double x1, double *y,
                       double tol_abs, double tol_rel, double *h_out,
                       int64_t counters[3]);

/* One Cash-Karp step (no step-size control) of a built-in test system from
   (t, y): y is replaced by the fifth-order solution, yerr receives the
   embedded error estimate.  which: 1, 2 as above; 3 = y'=y.              */
int orc_test_rkck_step(int which, double t, double h, double *y, double *yerr);

#ifdef __cplusplus
}
#endif
#endif
