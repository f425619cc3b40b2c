// tree_exec.hip -- runs a tree schedule (include/gal_tree.h) on the GPU.
//
// Every node of the forest owns one slot of the structure-of-arrays state.
// A wave is a set of operations on distinct slots whose earlier operations
// all sit in earlier waves, so executing waves in order gives every slot the
// same sequence of operations as the reference's sequential walk.  Per wave:
//   1. the wave's EVOLVE intervals are scattered into the per-slot start/end
//      time and halo-parameter arrays;
//   2. the persistent integrator evolves exactly the listed slots
//      (gal_ode_evolve_subset: lanes claim the listed nodes, costliest first);
//   3. one thread per event applies promotions (the galaxy, and its step
//      size, move into the parent's slot: node_evolver/standard.F90:1241-1327)
//      and galaxy mergers (mergerMassMovementsSimple,
//      satellites/merging/mass_movements/simple.F90:208-290: satellite gas and
//      stars to the host spheroid; in a major merger the host disk joins
//      them; the satellite's hot-halo reservoirs and black hole join the host's).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/gal_ode.h"
#include "../../include/gal_tree.h"
#include "physics.hpp"
#include "batch.hpp"

using namespace gal;

namespace {

__global__ void fill_kernel(int n, double value, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = value;
}

__global__ void scatter_kernel(int n, const int* __restrict__ slot, const double* __restrict__ t0,
                               const double* __restrict__ t1, const double* __restrict__ np, int ld,
                               double* __restrict__ slot_t0, double* __restrict__ slot_t1,
                               double* __restrict__ slot_np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = slot[i];
  slot_t0[s] = t0[i];
  slot_t1[s] = t1[i];
#pragma unroll
  for (int k = 0; k < GAL_NPARAM_NODE; ++k) slot_np[(size_t)k * ld + s] = np[(size_t)GAL_NPARAM_NODE * i + k];
}

__global__ void event_kernel(int n, const int* __restrict__ kind, const int* __restrict__ from,
                             const int* __restrict__ to, int ld, double* __restrict__ y, double* __restrict__ h_step,
                             double major_ratio) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t a = (size_t)from[i], b = (size_t)to[i];
  auto Y = [&](int v, size_t s) -> double& { return y[(size_t)v * ld + s]; };
  if (kind[i] == GAL_OP_PROMOTE) {
#pragma unroll
    for (int v = 0; v < kNV; ++v) Y(v, b) = Y(v, a);
    h_step[b] = h_step[a];
    return;
  }
  // GAL_OP_MERGE_GALAXY: satellite a into host b
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

template <typename T>
T* upload(const std::vector<T>& v, hipStream_t s, bool& ok) {
  if (v.empty() || !ok) return nullptr;
  T* d = nullptr;
  if (hipMalloc(&d, sizeof(T) * v.size()) != hipSuccess) { ok = false; return nullptr; }
  if (hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s) != hipSuccess) ok = false;
  return d;
}

}  // namespace

extern "C" int gal_tree_evolve_d(gal_ode_batch* h, const gal_schedule* sched, int ld, double* y, double* scale,
                                 int* status, double major_ratio, void* stream) {
  if (!h || !sched || ld < 1 || ld > h->n_max || !y || !status) return GAL_E_INVALID;
  if ((h->flags & GAL_FLAG_SCALES_PROVIDED) && !scale) return GAL_E_INVALID;
  int64_t n_ops = 0;
  int n_waves = 0;
  if (gal_schedule_size(sched, &n_ops, &n_waves) != GAL_OK) return GAL_E_INVALID;
  if (n_ops == 0) return GAL_OK;
  std::vector<gal_schedule_op> ops((size_t)n_ops);
  gal_schedule_ops(sched, ops.data());
  for (const gal_schedule_op& op : ops)
    if (op.a < 0 || op.a >= ld || op.b >= ld) return GAL_E_INVALID;
  // Bucket the operations by wave (stable: within a wave the order of the
  // schedule is kept, although operations of one wave touch distinct slots).
  std::vector<int64_t> ev_count(n_waves + 1, 0), evt_count(n_waves + 1, 0);
  for (const gal_schedule_op& op : ops) {
    if (op.kind == GAL_OP_EVOLVE) ev_count[op.wave + 1]++;
    else if (op.kind == GAL_OP_PROMOTE || op.kind == GAL_OP_MERGE_GALAXY) evt_count[op.wave + 1]++;
  }
  for (int w = 0; w < n_waves; ++w) { ev_count[w + 1] += ev_count[w]; evt_count[w + 1] += evt_count[w]; }
  const size_t n_ev = (size_t)ev_count[n_waves], n_evt = (size_t)evt_count[n_waves];
  std::vector<int> ev_slot(n_ev), evt_kind(n_evt), evt_a(n_evt), evt_b(n_evt);
  std::vector<double> ev_t0(n_ev), ev_t1(n_ev), ev_np(GAL_NPARAM_NODE * n_ev);
  {
    std::vector<int64_t> pe(ev_count.begin(), ev_count.end() - 1), pt(evt_count.begin(), evt_count.end() - 1);
    for (const gal_schedule_op& op : ops) {
      if (op.kind == GAL_OP_EVOLVE) {
        const int64_t k = pe[op.wave]++;
        ev_slot[k] = op.a;
        ev_t0[k] = op.t0;
        ev_t1[k] = op.t1;
        for (int j = 0; j < GAL_NPARAM_NODE; ++j) ev_np[GAL_NPARAM_NODE * k + j] = op.node_par[j];
      } else if (op.kind == GAL_OP_PROMOTE || op.kind == GAL_OP_MERGE_GALAXY) {
        const int64_t k = pt[op.wave]++;
        evt_kind[k] = op.kind;
        evt_a[k] = op.a;
        evt_b[k] = op.b;
      }
    }
  }
  hipStream_t s = (hipStream_t)stream;
  bool ok = true;
  int* d_ev_slot = upload(ev_slot, s, ok);
  double* d_ev_t0 = upload(ev_t0, s, ok);
  double* d_ev_t1 = upload(ev_t1, s, ok);
  double* d_ev_np = upload(ev_np, s, ok);
  int* d_evt_kind = upload(evt_kind, s, ok);
  int* d_evt_a = upload(evt_a, s, ok);
  int* d_evt_b = upload(evt_b, s, ok);
  double *d_t0 = nullptr, *d_t1 = nullptr, *d_np = nullptr, *d_h = nullptr;
  ok = ok && hipMalloc(&d_t0, sizeof(double) * ld) == hipSuccess && hipMalloc(&d_t1, sizeof(double) * ld) == hipSuccess &&
       hipMalloc(&d_np, sizeof(double) * GAL_NPARAM_NODE * (size_t)ld) == hipSuccess &&
       hipMalloc(&d_h, sizeof(double) * ld) == hipSuccess;
  int rc = ok ? GAL_OK : GAL_E_NOMEM;
  if (rc == GAL_OK) {
    hipLaunchKernelGGL(fill_kernel, dim3((ld + 255) / 256), dim3(256), 0, s, ld, -1.0, d_h);
    hipLaunchKernelGGL(fill_kernel, dim3((ld + 255) / 256), dim3(256), 0, s, ld, 0.0, d_t0);
    hipLaunchKernelGGL(fill_kernel, dim3((ld + 255) / 256), dim3(256), 0, s, ld, 0.0, d_t1);
    hipLaunchKernelGGL(fill_kernel, dim3((GAL_NPARAM_NODE * ld + 255) / 256), dim3(256), 0, s, GAL_NPARAM_NODE * ld,
                       0.0, d_np);
    if (hipGetLastError() != hipSuccess) rc = GAL_E_HIP;
  }
  for (int w = 0; w < n_waves && rc == GAL_OK; ++w) {
    const int64_t e0 = ev_count[w], ne = ev_count[w + 1] - ev_count[w];
    if (ne > 0) {
      hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, (int)ne, d_ev_slot + e0,
                         (const double*)d_ev_t0 + e0, (const double*)d_ev_t1 + e0,
                         (const double*)d_ev_np + GAL_NPARAM_NODE * e0, ld, d_t0, d_t1, d_np);
      rc = gal_ode_evolve_subset(h, ld, ld, y, d_np, d_t0, d_t1, d_h, scale, status, nullptr, d_ev_slot + e0, (int)ne,
                                 s);
      if (rc != GAL_OK) break;
    }
    const int64_t v0 = evt_count[w], nv = evt_count[w + 1] - evt_count[w];
    if (nv > 0) {
      hipLaunchKernelGGL(event_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, (int)nv,
                         (const int*)d_evt_kind + v0, (const int*)d_evt_a + v0, (const int*)d_evt_b + v0, ld, y, d_h,
                         major_ratio);
      if (hipGetLastError() != hipSuccess) rc = GAL_E_HIP;
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess && rc == GAL_OK) rc = GAL_E_HIP;
  hipFree(d_ev_slot); hipFree(d_ev_t0); hipFree(d_ev_t1); hipFree(d_ev_np);
  hipFree(d_evt_kind); hipFree(d_evt_a); hipFree(d_evt_b);
  hipFree(d_t0); hipFree(d_t1); hipFree(d_np); hipFree(d_h);
  return rc;
}
