/*
 * gal_tree.h -- merger-tree forests and the order in which their nodes evolve.
 *
 * Host side of the tree path:
 *  - a forest builder that follows the Parkinson, Cole & Helly (2008)
 *    branching rates with the Cole et al. (2000) binary-split construction
 *    (reference: merger_trees/branching_probability/Parkinson_Cole_Helly.F90:265-1406,
 *    merger_trees/construct/builder/Cole2000/_class.F90:360-915), or a forest
 *    given as arrays (the fully-specified tree format,
 *    merger_trees/construct/fully_specified.F90:269-450);
 *  - the reference's tree walk (mergerTreeEvolverStandard%evolve,
 *    merger_trees/evolver/standard.F90:291-635, with nodeIsEvolvable :723-760,
 *    timeEvolveTo :762-1035, the all-nodes walker
 *    merger_trees/walkers/all_nodes/_class.F90:101-222, promotion and halo
 *    merging node_evolver/standard.F90:1241-1356 and
 *    node_merger/single_level_hierarchy.F90:65-115, satellite mergers
 *    evolve/timesteps/satellite.F90:210-278).  Node evolution intervals and
 *    events depend only on halo times and masses, so the walk is run once on
 *    the host and recorded as a schedule: the exact sequence of evolution
 *    intervals, promotions, halo mergers and galaxy mergers of the reference;
 *  - the decomposition of that schedule into waves: an operation's wave is
 *    one more than the latest wave of any earlier operation on the same
 *    node, so running wave after wave reproduces the sequential order
 *    operation for operation.
 * gal_tree_evolve_d (tree_exec.hip) runs a schedule on the GPU.
 */
#ifndef GAL_TREE_H
#define GAL_TREE_H

#include <stdint.h>
#include "gal_ode.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gal_forest gal_forest;
typedef struct gal_schedule gal_schedule;

/* Cosmology and power spectrum of the builder (flat LCDM; Eisenstein & Hu
   1998 zero-baryon-wiggle transfer function; sigma_8 normalisation). */
typedef struct gal_cosmology {
  double omega_matter, omega_dark_energy, omega_baryon, hubble_constant; /* H0 in km/s/Mpc */
  double sigma_8, index_spectral, temperature_cmb;
} gal_cosmology;

/* Builder parameters (Cole2000 defaults: mergeProbability 0.1, accretionLimit
   0.1; PCH defaults G0 0.57, gamma1 0.38, gamma2 -0.01, epsilon1 0.1). */
typedef struct gal_pch_params {
  double mass_resolution;     /* Msun */
  double merge_probability;   /* maximum split probability per step */
  double accretion_limit;     /* sub-resolution accretion fraction that inserts a node */
  double g0, gamma1, gamma2;  /* PCH modification G(s1/s2, d2/s2) */
  double epsilon1;            /* step bound: dw <= epsilon1 sqrt(2 (S(M/2) - S(M))) */
  double time_earliest;       /* Gyr: branches stop here */
} gal_pch_params;

/* Parameters of the tree walk (mergerTreeEvolverStandard and the timesteps). */
typedef struct gal_walk_params {
  double time_end;                       /* Gyr: evolve to this time (normally the roots' time) */
  double timestep_relative, timestep_absolute;          /* simple timestep: min(rel/H, abs) */
  double timestep_host_relative, timestep_host_absolute;  /* satellites vs their host */
  double satellite_offset_absolute, satellite_offset_relative;  /* satellite timestep offsets */
  double merging_multiplier;             /* Jiang (2008) timescale multiplier */
  double orbit_radius_ratio, orbit_circularity;  /* orbit used for the merging time */
  double spin_parameter;                 /* spin of every halo: J = lambda M V_vir r_vir */
} gal_walk_params;

/* Operations of a schedule. */
#define GAL_OP_EVOLVE          0  /* evolve node a from t0 to t1                         */
#define GAL_OP_PROMOTE         1  /* node a's galaxy moves into its parent b             */
#define GAL_OP_BECOME_SATELLITE 2 /* node a becomes a satellite of b (halo merger)        */
#define GAL_OP_MERGE_GALAXY    3  /* satellite a merges into host b (mass movements)     */

typedef struct gal_schedule_op {
  int32_t kind, a, b, wave;
  double t0, t1;
  double node_par[GAL_NPARAM_NODE];      /* EVOLVE: halo of node a over [t0, t1] */
} gal_schedule_op;

/* ---- forests ------------------------------------------------------------- */
int gal_forest_build_pch(const gal_cosmology* cosmo, const gal_pch_params* pch, int n_trees,
                         const double* root_mass, double time_root, uint64_t seed, gal_forest** out);
/* Forest from arrays: indices are global (0..n_nodes-1), -1 for none;
   tree t holds nodes tree_offset[t] .. tree_offset[t+1]-1, its root first. */
int gal_forest_from_arrays(int64_t n_nodes, int n_trees, const int64_t* tree_offset, const int32_t* parent,
                           const int32_t* first_child, const int32_t* sibling, const double* mass,
                           const double* time, gal_forest** out);
int gal_forest_size(const gal_forest* f, int64_t* n_nodes, int* n_trees);
int gal_forest_arrays(const gal_forest* f, int64_t* tree_offset, int32_t* parent, int32_t* first_child,
                      int32_t* sibling, double* mass, double* time);
int gal_forest_destroy(gal_forest* f);

/* sigma(M) and delta_c(t) of the builder's cosmology (for tests). */
int gal_cosmology_sigma(const gal_cosmology* cosmo, int n, const double* mass, double* sigma);
int gal_cosmology_delta_critical(const gal_cosmology* cosmo, int n, const double* time, double* delta);

/* ---- schedules ----------------------------------------------------------- */
int gal_schedule_build(const gal_forest* f, const gal_cosmology* cosmo, const gal_walk_params* walk,
                       gal_schedule** out);
int gal_schedule_size(const gal_schedule* s, int64_t* n_ops, int* n_waves);
int gal_schedule_ops(const gal_schedule* s, gal_schedule_op* ops);
/* Node-Gyr the schedule evolves: the lengths t1 - t0 of its EVOLVE
   operations summed in schedule order, without copying the operations out.
   No counterpart in the reference (a workload statistic). */
int gal_schedule_node_gyr(const gal_schedule* s, double* node_gyr);
int gal_schedule_destroy(gal_schedule* s);

/* Initial galaxies of the leaf nodes (state y[v*ld + i], SoA, for every node
   of the forest; non-leaves are zero): hot gas f_b M with metallicity
   metal_fraction, a black-hole seed, spin angular momentum from the walk's
   spin parameter. */
int gal_forest_initial_state(const gal_forest* f, const gal_cosmology* cosmo, const gal_walk_params* walk,
                             double metal_fraction, double black_hole_seed, int ld, double* y);

/* ---- execution on the GPU ------------------------------------------------ */
/* Run a schedule (the work of mergerTreeEvolverStandard%evolve,
   merger_trees/evolver/standard.F90:291-1035, for every tree of the forest)
   as an operation graph in one persistent launch of the integrator: each
   operation waits only for the previous operations on the nodes it touches;
   idle lanes claim ready operations, evolve intervals with gal_ode.h
   semantics, apply events inline and release what follows.  y, scale: SoA
   [GAL_NVAR*ld] over the forest's nodes (device); status: [ld] (device), the
   status of each node's latest evolve.  Merger mass movements follow
   mergerMassMovementsSimple (satellites/merging/mass_movements/simple.F90:
   208-290) with major_ratio the major-merger mass ratio.  The handle's
   wall-time limit, if set, covers the whole forest.  Returns GAL_E_HIP if the
   graph stops making progress (no operation becomes ready for a minute). */
int gal_tree_evolve_d(gal_ode_batch* h, const gal_schedule* s, int ld, double* y, double* scale, int* status,
                      double major_ratio, void* stream);

/* The same schedule wave by wave: each wave's EVOLVE operations as one
   batched evolve of the listed nodes, then its event operations; a wave
   waits for the slowest interval of the wave before it.  The result equals
   gal_tree_evolve_d's bit for bit (the handle's wall-time limit, if set,
   applies to each wave's launch). */
int gal_tree_evolve_waves_d(gal_ode_batch* h, const gal_schedule* s, int ld, double* y, double* scale,
                            int* status, double major_ratio, void* stream);

/* Apply node events to host arrays (copied to the device and back): each
   event is GAL_OP_PROMOTE (node's galaxy moves into its parent target:
   mergerTreeNodeEvolverStandard%promote, node_evolver/standard.F90:1241-1356)
   or GAL_OP_MERGE_GALAXY (satellite node merges into host target:
   mergerTreeNodeEvolverStandard%merge through mergerMassMovementsSimple,
   satellites/merging/mass_movements/simple.F90:208-290, the satellite's
   stripped gas joining the host's outflowed reservoir).  y: SoA
   [GAL_NVAR*n_nodes]; h_step [n_nodes] (a promotion hands the node's step
   size to its parent).  Events run in parallel, so every node may appear in
   at most one of them (GAL_E_INVALID otherwise). */
int gal_node_events_h(int n_events, const int* kind, const int* node, const int* target, int n_nodes, double* y,
                      double* h_step, double major_ratio);

/* The same with the learned surrogate in front of the integrator
   (include/gal_surrogate.h): in every wave the evolve operations of halos
   that are not satellites go through gal_surrogate_evolve_d (network, gate,
   integrator for the rejected), those of satellites -- outside the
   network's training distribution -- straight to the integrator.
   *n_accepted (host, may be NULL) counts the intervals the surrogate took.
   Needs gal_surrogate.h's types. */
struct gal_mlp;
struct gal_surrogate_gate;

/* gal_tree_evolve_d that also records every evolve operation, in execution
   order: the node's state before (rec_y0, device) and after (rec_y1, device)
   the interval, SoA [GAL_NVAR*capacity]; its times (rec_t0, rec_t1, host
   [capacity]) and node parameters (rec_node_par, host
   [GAL_NPARAM_NODE*capacity], SoA).  *n_recorded (host) = number of evolve
   operations; capacity must hold them all.  The surrogate's training pairs
   come from these trajectories. */
int gal_tree_evolve_record_d(gal_ode_batch* h, const gal_schedule* s, int ld, double* y, double* scale,
                             int* status, double major_ratio, double* rec_y0, double* rec_y1, double* rec_t0,
                             double* rec_t1, double* rec_node_par, int64_t capacity, int64_t* n_recorded,
                             void* stream);
int gal_tree_evolve_surrogate_d(gal_ode_batch* h, struct gal_mlp* m, const float* norm,
                                const struct gal_surrogate_gate* gate, const gal_schedule* s, int ld, double* y,
                                double* scale, int* status, double major_ratio, int64_t* n_accepted, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GAL_TREE_H */
