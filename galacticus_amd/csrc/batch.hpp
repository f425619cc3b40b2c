// batch.hpp -- the gal_ode_batch handle shared by the translation units of
// the library (the evolver in gal_ode.hip, the learned surrogate in
// surrogate.hip).  Internal: not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "../../include/gal_ode.h"
#include "physics.hpp"
#include "events.hpp"

struct gal_ode_batch {
  int n_max = 0;
  double tol_abs = 0.0, tol_rel = 0.0;
  int flags = 0;
  int device = 0;
  int n_cus = 0;
  int nodes_per_block = 64;
  int blocks_per_cu = 1;
  int rhs_waves = 2;          // waves per group of 64 nodes (2: the RHS split over two SIMDs)
  double wall_limit_s = 0.0;  // per-launch wall-time limit (0: none)
  double wall_clock_hz = 0.0; // frequency of the device's realtime counter
  double graph_stall_s = 60.0;  // operation-graph executor: give up after this long without progress
  double params[GAL_NPARAM] = {};
  gal::DerivedConstants D{};
  gal::CoolingDesc cd{};
  double* d_params = nullptr;
  double* d_table = nullptr;
  unsigned int* d_work = nullptr;
  double* d_ywork = nullptr;  // lane-private state/scales for two workgroups per CU
  int ywork_lanes = 0;
  double* d_kwork = nullptr;  // lane-private stage derivatives for four workgroups per CU
  int kwork_lanes = 0;
  int* d_order = nullptr;
  // claim order planned by gal_ode_batch_plan_d in d_order, for the next
  // gal_ode_batch_evolve_planned_d of the same y, n and ld
  const double* plan_y = nullptr;
  int plan_n = 0, plan_ld = 0;
  int* d_index = nullptr;
  float* d_key = nullptr;
  float* d_key_sorted = nullptr;
  void* d_sort_temp = nullptr;
  size_t sort_temp_bytes = 0;
  bool have_physics = false, have_table = false;
  // staging for the host-pointer entry point
  double *d_y = nullptr, *d_np = nullptr, *d_t0 = nullptr, *d_t1 = nullptr, *d_h = nullptr, *d_scale = nullptr;
  int* d_status = nullptr;
  long long* d_counters = nullptr;
};

#define HIP_TRY(x)                                                              \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "gal_ode: %s failed: %s\n", #x, hipGetErrorString(e_));   \
      return GAL_E_HIP;                                                         \
    }                                                                           \
  } while (0)


// Evolve the nodes listed in subset[0..n_subset) (device indices into the
// batch of n) in place; nodes not listed are not touched.  The claim order
// is the listed order, or the cost order of the listed nodes when the handle
// has GAL_FLAG_COST_ORDER.
int gal_ode_evolve_subset(gal_ode_batch* h, int n, int ld, double* y, const double* node_par, const double* t_start,
                          const double* t_end, double* h_step, double* scale, int* status, int64_t* counters,
                          const int* subset, int n_subset, hipStream_t s,
                          const int* n_subset_dev = nullptr);

// Run a forest's operation graph (events.hpp) in one persistent launch:
// lanes claim ready operations, evolve intervals in place on the slot-indexed
// state, apply events inline and release the operations that follow.  Every
// operation of one slot runs in graph order, so the state equals that of the
// per-wave executor bit for bit.  *graph_error (device) is set if lanes gave
// up waiting on a graph that stopped making progress.
int gal_ode_evolve_graph(gal_ode_batch* h, int ld, double* y, double* h_step, double* scale, int* status,
                         const gal::OpGraph& g, int* graph_error, hipStream_t s);
