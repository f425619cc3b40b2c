// events.hpp -- the tree events that move state between slots, as device
// functions shared by the per-wave event kernel and the integrator's
// operation-graph mode.  Internal: not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

#include "../../include/gal_ode.h"
#include "../../include/gal_tree.h"
#include "physics.hpp"

namespace gal {

// Apply one event to the state y (SoA, leading dimension ld) and the
// per-slot step sizes:
//  GAL_OP_PROMOTE: the galaxy in slot a, and its step size, move into the
//    parent's slot b (node_evolver/standard.F90:1241-1327);
//  GAL_OP_MERGE_GALAXY: satellite a merges into host b with the simple mass
//    movements (satellites/merging/mass_movements/simple.F90:208-290):
//    satellite gas and stars to the host spheroid, and in a major merger the
//    host disk joins them; the satellite's hot-halo reservoirs and black hole
//    join the host's.
__device__ inline void apply_event(int kind, size_t a, size_t b, int ld, double* __restrict__ y,
                                   double* __restrict__ h_step, double major_ratio) {
  auto Y = [&](int v, size_t s) -> double& { return y[(size_t)v * ld + s]; };
  if (kind == GAL_OP_PROMOTE) {
#pragma unroll
    for (int v = 0; v < kNV; ++v) Y(v, b) = Y(v, a);
    h_step[b] = h_step[a];
    return;
  }
  if (kind != GAL_OP_MERGE_GALAXY) return;
  const double m_sat = Y(GAL_V_DISK_MASS_GAS, a) + Y(GAL_V_DISK_MASS_STELLAR, a) + Y(GAL_V_SPHEROID_MASS_GAS, a) +
                       Y(GAL_V_SPHEROID_MASS_STELLAR, a);
  const double m_host = Y(GAL_V_DISK_MASS_GAS, b) + Y(GAL_V_DISK_MASS_STELLAR, b) + Y(GAL_V_SPHEROID_MASS_GAS, b) +
                        Y(GAL_V_SPHEROID_MASS_STELLAR, b);
  const bool major = m_sat > 0.0 && m_host > 0.0 && fmin(m_sat, m_host) >= major_ratio * fmax(m_sat, m_host);
  struct Move { int from, to; };
  // satellite disk and spheroid -> host spheroid
  const Move sat_moves[] = {
      {GAL_V_DISK_MASS_GAS, GAL_V_SPHEROID_MASS_GAS},
      {GAL_V_DISK_ABUNDANCES_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS},
      {GAL_V_DISK_MASS_STELLAR, GAL_V_SPHEROID_MASS_STELLAR},
      {GAL_V_DISK_ABUNDANCES_STELLAR, GAL_V_SPHEROID_ABUNDANCES_STELLAR},
      {GAL_V_DISK_MASS_STELLAR_FORMED, GAL_V_SPHEROID_MASS_STELLAR_FORMED},
      {GAL_V_DISK_ANGULAR_MOMENTUM, GAL_V_SPHEROID_ANGULAR_MOMENTUM},
      {GAL_V_SPHEROID_MASS_GAS, GAL_V_SPHEROID_MASS_GAS},
      {GAL_V_SPHEROID_ABUNDANCES_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS},
      {GAL_V_SPHEROID_MASS_STELLAR, GAL_V_SPHEROID_MASS_STELLAR},
      {GAL_V_SPHEROID_ABUNDANCES_STELLAR, GAL_V_SPHEROID_ABUNDANCES_STELLAR},
      {GAL_V_SPHEROID_MASS_STELLAR_FORMED, GAL_V_SPHEROID_MASS_STELLAR_FORMED},
      {GAL_V_SPHEROID_ANGULAR_MOMENTUM, GAL_V_SPHEROID_ANGULAR_MOMENTUM},
      // the satellite's circumgalactic reservoirs and black hole
      {GAL_V_HOT_MASS, GAL_V_HOT_MASS},
      {GAL_V_HOT_ABUNDANCES, GAL_V_HOT_ABUNDANCES},
      {GAL_V_HOT_ANGULAR_MOMENTUM, GAL_V_HOT_ANGULAR_MOMENTUM},
      {GAL_V_HOT_OUTFLOWED_MASS, GAL_V_HOT_OUTFLOWED_MASS},
      {GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM, GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM},
      {GAL_V_HOT_OUTFLOWED_ABUNDANCES, GAL_V_HOT_OUTFLOWED_ABUNDANCES},
      {GAL_V_HOT_UNACCRETED_MASS, GAL_V_HOT_UNACCRETED_MASS},
      {GAL_V_HOT_UNACCRETED_ABUNDANCES, GAL_V_HOT_UNACCRETED_ABUNDANCES},
      {GAL_V_HOT_ENERGY_RADIATED, GAL_V_HOT_ENERGY_RADIATED},
      {GAL_V_BH_MASS, GAL_V_BH_MASS},
  };
  for (const Move& m : sat_moves) {
    Y(m.to, b) += Y(m.from, a);
    Y(m.from, a) = 0.0;
  }
  // the satellite's stripped gas joins the host's outflowed reservoir, as
  // hot_halo/standard/_class.F90:532-597 moves it after each satellite
  // evolve (here at the merger: mass and metals; its angular momentum,
  // M_stripped J_host/M_host there, is not carried)
  Y(GAL_V_HOT_OUTFLOWED_MASS, b) += Y(GAL_V_HOT_STRIPPED_MASS, a);
  Y(GAL_V_HOT_STRIPPED_MASS, a) = 0.0;
  Y(GAL_V_HOT_OUTFLOWED_ABUNDANCES, b) += Y(GAL_V_HOT_STRIPPED_ABUNDANCES, a);
  Y(GAL_V_HOT_STRIPPED_ABUNDANCES, a) = 0.0;
  if (major) {
    const Move host_moves[] = {
        {GAL_V_DISK_MASS_GAS, GAL_V_SPHEROID_MASS_GAS},
        {GAL_V_DISK_ABUNDANCES_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS},
        {GAL_V_DISK_MASS_STELLAR, GAL_V_SPHEROID_MASS_STELLAR},
        {GAL_V_DISK_ABUNDANCES_STELLAR, GAL_V_SPHEROID_ABUNDANCES_STELLAR},
        {GAL_V_DISK_MASS_STELLAR_FORMED, GAL_V_SPHEROID_MASS_STELLAR_FORMED},
        {GAL_V_DISK_ANGULAR_MOMENTUM, GAL_V_SPHEROID_ANGULAR_MOMENTUM},
    };
    for (const Move& m : host_moves) {
      Y(m.to, b) += Y(m.from, b);
      Y(m.from, b) = 0.0;
    }
  }
}

// A forest's operations as a dependency graph for one persistent launch of
// the integrator.  ops is the device copy of the schedule's operations,
// indexed as in the schedule; the graph holds n of them (becoming a
// satellite changes no state and is left out).  Operation k touches slot
// ops[k].a (and ops[k].b for events); it waits on the previous operation on
// each slot it touches, and next_a[k], next_b[k] are the operations that
// follow it on those slots (-1: none).  npred[k] counts the links still
// unfinished; an operation whose count reaches zero is appended to ready[]
// (tail counts the appends), from which idle lanes claim entries in order.
// Every operation of the graph is appended exactly once, so ready[] has n
// entries; entries not yet written hold -1.
struct OpGraph {
  int n = 0;                               // operations in the graph
  const gal_schedule_op* ops = nullptr;    // the schedule's operations (device)
  const int* next_a = nullptr;
  const int* next_b = nullptr;
  int* npred = nullptr;
  int* ready = nullptr;
  unsigned int* tail = nullptr;
  double major_ratio = 0.25;
};

__device__ __forceinline__ int graph_ready_entry(const OpGraph& g, int j) {
  return __hip_atomic_load(&g.ready[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Operation k has finished and its writes are issued: make them visible at
// device scope, then count it off each successor; the lane that counts a
// successor to zero appends it to the ready list.
__device__ inline void graph_release(const OpGraph& g, int k) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  const int nx[2] = {g.next_a[k], g.next_b[k]};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int m = nx[i];
    if (m < 0) continue;
    const int left = __hip_atomic_fetch_add(&g.npred[m], -1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (left == 1) {
      const unsigned int pos = __hip_atomic_fetch_add(g.tail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (pos < (unsigned int)g.n) __hip_atomic_store(&g.ready[pos], m, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace gal
