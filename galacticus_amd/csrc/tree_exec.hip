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
#include <cstdio>
#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/gal_ode.h"
#include "../../include/gal_tree.h"
#include "../../include/gal_surrogate.h"
#include "physics.hpp"
#include "batch.hpp"
#include "events.hpp"
#include "schedule.hpp"

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
  apply_event(kind[i], (size_t)from[i], (size_t)to[i], ld, y, h_step, major_ratio);
}

// Compact copies of a wave's non-satellite evolve operations for the
// surrogate: state, step size, times and node parameters (SoA, leading
// dimension ldc).
__global__ void gather_kernel(int n, const int* __restrict__ slot, int ld, const double* __restrict__ y,
                              const double* __restrict__ h, const double* __restrict__ t0,
                              const double* __restrict__ t1, const double* __restrict__ np, int ldc,
                              double* __restrict__ yc, double* __restrict__ hc, double* __restrict__ t0c,
                              double* __restrict__ t1c, double* __restrict__ npc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t s = (size_t)slot[k];
#pragma unroll
  for (int v = 0; v < kNV; ++v) yc[(size_t)v * ldc + k] = y[(size_t)v * ld + s];
  hc[k] = h[s];
  t0c[k] = t0[k];
  t1c[k] = t1[k];
#pragma unroll
  for (int j = 0; j < GAL_NPARAM_NODE; ++j) npc[(size_t)j * ldc + k] = np[(size_t)GAL_NPARAM_NODE * k + j];
}

__global__ void scatter_back_kernel(int n, const int* __restrict__ slot, int ldc, const double* __restrict__ yc,
                                    const double* __restrict__ hc, const int* __restrict__ stc,
                                    const double* __restrict__ scc, const int* __restrict__ acc, int ld,
                                    double* __restrict__ y, double* __restrict__ h, int* __restrict__ status,
                                    double* __restrict__ scale, unsigned long long* __restrict__ n_accepted) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t s = (size_t)slot[k];
#pragma unroll
  for (int v = 0; v < kNV; ++v) y[(size_t)v * ld + s] = yc[(size_t)v * ldc + k];
  if (scale) {
#pragma unroll
    for (int v = 0; v < kNV; ++v) scale[(size_t)v * ld + s] = scc[(size_t)v * ldc + k];
  }
  h[s] = hc[k];
  status[s] = stc[k];
  if (acc[k]) atomicAdd(n_accepted, 1ull);
}

// Copy the states of a wave's evolve operations into a record (SoA, leading
// dimension ldr, column r0 + k for operation k of the wave).
__global__ void record_kernel(int n, const int* __restrict__ slot, int ld, const double* __restrict__ y, int ldr,
                              int64_t r0, double* __restrict__ rec) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t s = (size_t)slot[k];
#pragma unroll
  for (int v = 0; v < kNV; ++v) rec[(size_t)v * ldr + (size_t)(r0 + k)] = y[(size_t)v * ld + s];
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

struct SurrogateRoute {
  gal_mlp* m;
  const float* norm;
  const gal_surrogate_gate* gate;
  int64_t* n_accepted;   // host out
};

struct Recording {
  double* y0;            // device [GAL_NVAR * cap]: state before each evolve operation
  double* y1;            // device [GAL_NVAR * cap]: state after it
  double* t0;            // host [cap]
  double* t1;            // host [cap]
  double* node_par;      // host [GAL_NPARAM_NODE * cap], SoA with leading dimension cap
  int64_t cap;
  int64_t* n;            // host out: number of evolve operations recorded
};

static int tree_evolve(gal_ode_batch* h, const gal_schedule* sched, int ld, double* y, double* scale, int* status,
                       double major_ratio, void* stream, const SurrogateRoute* sr, const Recording* rec = nullptr) {
  if (!h || !sched || ld < 1 || ld > h->n_max || !y || !status) return GAL_E_INVALID;
  if ((h->flags & GAL_FLAG_SCALES_PROVIDED) && !scale) return GAL_E_INVALID;
  if (sr && (!sr->m || !sr->norm || !sr->gate)) return GAL_E_INVALID;
  if (sr && sr->n_accepted) *sr->n_accepted = 0;
  int64_t n_ops = 0;
  int n_waves = 0;
  if (gal_schedule_size(sched, &n_ops, &n_waves) != GAL_OK) return GAL_E_INVALID;
  if (n_ops == 0) return GAL_OK;
  std::vector<gal_schedule_op> ops((size_t)n_ops);
  gal_schedule_ops(sched, ops.data());
  for (const gal_schedule_op& op : ops)
    if (op.a < 0 || op.a >= ld || op.b >= ld) return GAL_E_INVALID;
  // Per wave: evolve operations (with the surrogate, those of halos that are
  // not satellites first) and event operations.  Operations of one wave
  // touch distinct nodes, so their order within the wave is free.
  std::vector<int64_t> ev_count(n_waves + 1, 0), evt_count(n_waves + 1, 0), ns_count(n_waves, 0);
  for (const gal_schedule_op& op : ops) {
    if (op.kind == GAL_OP_EVOLVE) {
      ev_count[op.wave + 1]++;
      if (sr && op.node_par[GAL_NP_IS_SATELLITE] == 0.0) ns_count[op.wave]++;
    } else if (op.kind == GAL_OP_PROMOTE || op.kind == GAL_OP_MERGE_GALAXY) {
      evt_count[op.wave + 1]++;
    }
  }
  for (int w = 0; w < n_waves; ++w) { ev_count[w + 1] += ev_count[w]; evt_count[w + 1] += evt_count[w]; }
  const size_t n_ev = (size_t)ev_count[n_waves], n_evt = (size_t)evt_count[n_waves];
  if (rec && (!rec->y0 || !rec->y1 || !rec->t0 || !rec->t1 || !rec->node_par || (int64_t)n_ev > rec->cap))
    return GAL_E_INVALID;
  std::vector<int> ev_slot(n_ev), evt_kind(n_evt), evt_a(n_evt), evt_b(n_evt);
  std::vector<double> ev_t0(n_ev), ev_t1(n_ev), ev_np(GAL_NPARAM_NODE * n_ev);
  int64_t cap = 0;
  {
    std::vector<int64_t> pe(ev_count.begin(), ev_count.end() - 1), pt(evt_count.begin(), evt_count.end() - 1);
    std::vector<int64_t> ps(n_waves);
    for (int w = 0; w < n_waves; ++w) {
      ps[w] = ev_count[w];                  // satellites follow the first ns_count[w] entries
      pe[w] = ev_count[w] + ns_count[w];
      cap = std::max(cap, ns_count[w]);
    }
    for (const gal_schedule_op& op : ops) {
      if (op.kind == GAL_OP_EVOLVE) {
        const bool front = sr && op.node_par[GAL_NP_IS_SATELLITE] == 0.0;
        const int64_t k = front ? ps[op.wave]++ : pe[op.wave]++;
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
  if (rec) {
    for (size_t k = 0; k < n_ev; ++k) {
      rec->t0[k] = ev_t0[k];
      rec->t1[k] = ev_t1[k];
      for (int j = 0; j < GAL_NPARAM_NODE; ++j) rec->node_par[(size_t)j * rec->cap + k] = ev_np[GAL_NPARAM_NODE * k + j];
    }
    if (rec->n) *rec->n = (int64_t)n_ev;
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
  // compact buffers of the surrogate route
  double *c_y = nullptr, *c_h = nullptr, *c_t0 = nullptr, *c_t1 = nullptr, *c_np = nullptr, *c_sc = nullptr;
  int *c_st = nullptr, *c_acc = nullptr;
  unsigned long long* d_nacc = nullptr;
  if (sr && cap > 0) {
    const size_t C = (size_t)cap;
    ok = ok && hipMalloc(&c_y, sizeof(double) * kNV * C) == hipSuccess &&
         hipMalloc(&c_sc, sizeof(double) * kNV * C) == hipSuccess && hipMalloc(&c_h, sizeof(double) * C) == hipSuccess &&
         hipMalloc(&c_t0, sizeof(double) * C) == hipSuccess && hipMalloc(&c_t1, sizeof(double) * C) == hipSuccess &&
         hipMalloc(&c_np, sizeof(double) * GAL_NPARAM_NODE * C) == hipSuccess &&
         hipMalloc(&c_st, sizeof(int) * C) == hipSuccess && hipMalloc(&c_acc, sizeof(int) * C) == hipSuccess &&
         hipMalloc(&d_nacc, sizeof(unsigned long long)) == hipSuccess &&
         hipMemsetAsync(d_nacc, 0, sizeof(unsigned long long), s) == hipSuccess;
  }
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
    const int64_t ns = sr ? ns_count[w] : 0;
    if (rec && ne > 0)
      hipLaunchKernelGGL(record_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, (int)ne,
                         (const int*)d_ev_slot + e0, ld, (const double*)y, (int)rec->cap, e0, rec->y0);
    if (ns > 0) {
      const unsigned blocks = (unsigned)((ns + 255) / 256);
      hipLaunchKernelGGL(gather_kernel, dim3(blocks), dim3(256), 0, s, (int)ns, (const int*)d_ev_slot + e0, ld,
                         (const double*)y, (const double*)d_h, (const double*)d_ev_t0 + e0,
                         (const double*)d_ev_t1 + e0, (const double*)d_ev_np + GAL_NPARAM_NODE * e0, (int)cap, c_y,
                         c_h, c_t0, c_t1, c_np);
      rc = gal_surrogate_evolve_d(h, sr->m, sr->norm, sr->gate, (int)ns, (int)cap, c_y, c_np, c_t0, c_t1, c_h, c_sc,
                                  c_st, c_acc, nullptr, s);
      if (rc != GAL_OK) break;
      hipLaunchKernelGGL(scatter_back_kernel, dim3(blocks), dim3(256), 0, s, (int)ns, (const int*)d_ev_slot + e0,
                         (int)cap, (const double*)c_y, (const double*)c_h, (const int*)c_st, (const double*)c_sc,
                         (const int*)c_acc, ld, y, d_h, status, scale, d_nacc);
      if (hipGetLastError() != hipSuccess) { rc = GAL_E_HIP; break; }
    }
    if (ne - ns > 0) {
      const int64_t r0 = e0 + ns, nr = ne - ns;
      hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, (int)nr, d_ev_slot + r0,
                         (const double*)d_ev_t0 + r0, (const double*)d_ev_t1 + r0,
                         (const double*)d_ev_np + GAL_NPARAM_NODE * r0, ld, d_t0, d_t1, d_np);
      rc = gal_ode_evolve_subset(h, ld, ld, y, d_np, d_t0, d_t1, d_h, scale, status, nullptr, d_ev_slot + r0, (int)nr,
                                 s);
      if (rc != GAL_OK) break;
    }
    if (rec && ne > 0)
      hipLaunchKernelGGL(record_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, (int)ne,
                         (const int*)d_ev_slot + e0, ld, (const double*)y, (int)rec->cap, e0, rec->y1);
    const int64_t v0 = evt_count[w], nv = evt_count[w + 1] - evt_count[w];
    if (nv > 0) {
      hipLaunchKernelGGL(event_kernel, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, (int)nv,
                         (const int*)d_evt_kind + v0, (const int*)d_evt_a + v0, (const int*)d_evt_b + v0, ld, y, d_h,
                         major_ratio);
      if (hipGetLastError() != hipSuccess) rc = GAL_E_HIP;
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess && rc == GAL_OK) rc = GAL_E_HIP;
  if (rc == GAL_OK && sr && sr->n_accepted && d_nacc) {
    unsigned long long na = 0;
    if (hipMemcpy(&na, d_nacc, sizeof(na), hipMemcpyDeviceToHost) != hipSuccess) rc = GAL_E_HIP;
    *sr->n_accepted = (int64_t)na;
  }
  hipFree(d_ev_slot); hipFree(d_ev_t0); hipFree(d_ev_t1); hipFree(d_ev_np);
  hipFree(d_evt_kind); hipFree(d_evt_a); hipFree(d_evt_b);
  hipFree(d_t0); hipFree(d_t1); hipFree(d_np); hipFree(d_h);
  hipFree(c_y); hipFree(c_sc); hipFree(c_h); hipFree(c_t0); hipFree(c_t1); hipFree(c_np); hipFree(c_st);
  hipFree(c_acc); hipFree(d_nacc);
  return rc;
}

namespace {

// ---- the operation graph --------------------------------------------------
// A schedule's operation graph on one device, built on the first call on
// that device and kept with the schedule: the operations themselves, the
// successor links and the pristine predecessor counts and ready list, plus
// the state of a run (counts, ready list, counters, per-slot step sizes),
// which every call resets on the device.  Calls that share a schedule on one
// device therefore must not overlap.
struct DeviceGraph {
  int device = -1;
  int max_slot = -1;  // highest slot an operation names
  int n_all = 0;      // operations of the schedule
  int n = 0;          // operations in the graph
  unsigned int r0 = 0;  // operations ready at the start
  gal_schedule_op* ops = nullptr;
  int *next_a = nullptr, *next_b = nullptr, *npred0 = nullptr, *ready0 = nullptr;
  int *npred = nullptr, *ready = nullptr, *err = nullptr;
  unsigned int* tail = nullptr;
  double* h = nullptr;
  int h_cap = 0;
  ~DeviceGraph() {
    int cur = -1;
    const bool switch_device = hipGetDevice(&cur) == hipSuccess && cur != device && device >= 0;
    if (switch_device) (void)hipSetDevice(device);
    (void)hipFree(ops); (void)hipFree(next_a); (void)hipFree(next_b); (void)hipFree(npred0);
    (void)hipFree(ready0); (void)hipFree(npred); (void)hipFree(ready); (void)hipFree(err); (void)hipFree(tail);
    (void)hipFree(h);
    if (switch_device) (void)hipSetDevice(cur);
  }
};

void delete_device_graph(void* p) { delete static_cast<DeviceGraph*>(p); }

// Links and initial ready list over the schedule's operations (host, once).
int build_device_graph(const gal_schedule& sc, int device, hipStream_t s, DeviceGraph& g) {
  const std::vector<gal_schedule_op>& ops = sc.ops;
  if (ops.size() > (size_t)INT32_MAX) return GAL_E_INVALID;
  const int n_all = (int)ops.size();
  int max_slot = -1;
  for (const gal_schedule_op& op : ops) {
    if (op.a < 0) return GAL_E_INVALID;
    if (op.kind == GAL_OP_PROMOTE || op.kind == GAL_OP_MERGE_GALAXY) {
      if (op.b < 0 || op.b == op.a) return GAL_E_INVALID;
    } else if (op.kind != GAL_OP_EVOLVE && op.kind != GAL_OP_BECOME_SATELLITE) {
      return GAL_E_INVALID;
    }
    max_slot = std::max(max_slot, std::max(op.a, op.b));
  }
  std::vector<int> next_a(n_all, -1), next_b(n_all, -1), npred(n_all, 0), ready;
  std::vector<int> last((size_t)max_slot + 1, -1);   // latest operation on each slot
  int n = 0;
  for (int k = 0; k < n_all; ++k) {
    const gal_schedule_op& op = ops[(size_t)k];
    // becoming a satellite changes no state and orders only its own node:
    // the operations either side of it on that node follow each other
    if (op.kind == GAL_OP_BECOME_SATELLITE) continue;
    ++n;
    const int touched[2] = {op.a, op.kind == GAL_OP_EVOLVE ? -1 : op.b};
    for (int slot : touched) {
      if (slot < 0) continue;
      const int p = last[(size_t)slot];
      if (p >= 0) {
        (ops[(size_t)p].a == slot ? next_a[(size_t)p] : next_b[(size_t)p]) = k;
        ++npred[(size_t)k];
      }
      last[(size_t)slot] = k;
    }
  }
  ready.assign((size_t)std::max(n, 1), -1);
  unsigned int r0 = 0;
  for (int k = 0; k < n_all; ++k)
    if (ops[(size_t)k].kind != GAL_OP_BECOME_SATELLITE && npred[(size_t)k] == 0) ready[r0++] = k;
  g.device = device;
  g.max_slot = max_slot;
  g.n_all = n_all;
  g.n = n;
  g.r0 = r0;
  bool ok = true;
  g.ops = upload(ops, s, ok);
  g.next_a = upload(next_a, s, ok);
  g.next_b = upload(next_b, s, ok);
  g.npred0 = upload(npred, s, ok);
  g.ready0 = upload(ready, s, ok);
  ok = ok && hipMalloc(&g.npred, sizeof(int) * (size_t)n_all) == hipSuccess &&
       hipMalloc(&g.ready, sizeof(int) * ready.size()) == hipSuccess &&
       hipMalloc(&g.err, sizeof(int)) == hipSuccess && hipMalloc(&g.tail, sizeof(unsigned int)) == hipSuccess;
  // the host vectors are released on return: the copies must have landed
  if (hipStreamSynchronize(s) != hipSuccess) return GAL_E_HIP;
  return ok ? GAL_OK : GAL_E_NOMEM;
}

// The schedule's graph on the current device, built on first use.
int device_graph(gal_schedule* sc, hipStream_t s, DeviceGraph** out) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return GAL_E_HIP;
  std::lock_guard<std::mutex> lock(sc->attach_mutex);
  for (auto& at : sc->attachments)
    if (at.first == device) {
      *out = static_cast<DeviceGraph*>(at.second.get());
      return GAL_OK;
    }
  auto g = std::make_unique<DeviceGraph>();
  const int rc = build_device_graph(*sc, device, s, *g);
  if (rc != GAL_OK) return rc;
  *out = g.get();
  sc->attachments.emplace_back(device, gal_schedule::Attachment(g.release(), delete_device_graph));
  return GAL_OK;
}

// Reset a run: predecessor counts, ready list, counters and step sizes.
__global__ void graph_reset_kernel(int n_all, const int* __restrict__ npred0, int* __restrict__ npred, int n,
                                   const int* __restrict__ ready0, int* __restrict__ ready, unsigned int r0,
                                   unsigned int* __restrict__ tail, int* __restrict__ err, int ld,
                                   double* __restrict__ h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_all) npred[i] = npred0[i];
  if (i < n) ready[i] = ready0[i];
  if (i < ld) h[i] = -1.0;
  if (i == 0) {
    *tail = r0;
    *err = 0;
  }
}

}  // namespace

// The schedule as an operation graph (events.hpp) run by one persistent
// launch of the integrator: no wave waits for the slowest interval of the
// one before it, only for the operations on its own slots.
static int tree_evolve_graph(gal_ode_batch* h, const gal_schedule* sched_c, int ld, double* y, double* scale,
                             int* status, double major_ratio, void* stream) {
  if (!h || !sched_c || ld < 1 || ld > h->n_max || !y || !status) return GAL_E_INVALID;
  if ((h->flags & GAL_FLAG_SCALES_PROVIDED) && !scale) return GAL_E_INVALID;
  gal_schedule* sched = const_cast<gal_schedule*>(sched_c);   // the graph is a cache kept with it
  if (sched->ops.empty()) return GAL_OK;
  hipStream_t s = (hipStream_t)stream;
  DeviceGraph* dg = nullptr;
  int rc = device_graph(sched, s, &dg);
  if (rc != GAL_OK) return rc;
  if (dg->max_slot >= ld) return GAL_E_INVALID;
  if (dg->n == 0) return GAL_OK;
  if (dg->h_cap < ld) {
    (void)hipFree(dg->h);
    dg->h = nullptr;
    dg->h_cap = 0;
    if (hipMalloc(&dg->h, sizeof(double) * (size_t)ld) != hipSuccess) return GAL_E_NOMEM;
    dg->h_cap = ld;
  }
  const int span = std::max(std::max(dg->n_all, dg->n), ld);
  hipLaunchKernelGGL(graph_reset_kernel, dim3((unsigned)((span + 255) / 256)), dim3(256), 0, s, dg->n_all,
                     (const int*)dg->npred0, dg->npred, dg->n, (const int*)dg->ready0, dg->ready, dg->r0, dg->tail,
                     dg->err, ld, dg->h);
  if (hipGetLastError() != hipSuccess) return GAL_E_HIP;
  gal::OpGraph g;
  g.n = dg->n;
  g.ops = dg->ops;
  g.next_a = dg->next_a;
  g.next_b = dg->next_b;
  g.npred = dg->npred;
  g.ready = dg->ready;
  g.tail = dg->tail;
  g.major_ratio = major_ratio;
  rc = gal_ode_evolve_graph(h, ld, y, dg->h, scale, status, g, dg->err, s);
  if (hipStreamSynchronize(s) != hipSuccess && rc == GAL_OK) rc = GAL_E_HIP;
  if (rc == GAL_OK) {
    int err = 0;
    unsigned int appended = 0;
    if (hipMemcpy(&err, dg->err, sizeof(err), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&appended, dg->tail, sizeof(appended), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = GAL_E_HIP;
    } else if (err != 0 || appended != (unsigned int)dg->n) {
      fprintf(stderr, "gal_tree_evolve_d: the operation graph stopped with %u of %d operations ready\n", appended,
              dg->n);
      rc = GAL_E_HIP;
    }
  }
  return rc;
}

extern "C" int gal_node_events_h(int n_events, const int* kind, const int* node, const int* target, int n_nodes,
                                 double* y, double* h_step, double major_ratio) {
  if (n_events < 0 || n_nodes < 1 || (n_events > 0 && (!kind || !node || !target || !y || !h_step)))
    return GAL_E_INVALID;
  if (n_events == 0) return GAL_OK;
  // events run in parallel on the device: each must name valid, distinct nodes
  std::vector<char> used((size_t)n_nodes, 0);
  for (int i = 0; i < n_events; ++i) {
    if (kind[i] != GAL_OP_PROMOTE && kind[i] != GAL_OP_MERGE_GALAXY) return GAL_E_INVALID;
    const int a = node[i], b = target[i];
    if (a < 0 || a >= n_nodes || b < 0 || b >= n_nodes || a == b || used[a] || used[b]) return GAL_E_INVALID;
    used[a] = used[b] = 1;
  }
  int *d_kind = nullptr, *d_a = nullptr, *d_b = nullptr;
  double *d_y = nullptr, *d_h = nullptr;
  const size_t ny = sizeof(double) * GAL_NVAR * (size_t)n_nodes, nh = sizeof(double) * (size_t)n_nodes,
               ne = sizeof(int) * (size_t)n_events;
  int rc = GAL_OK;
  if (hipMalloc(&d_kind, ne) != hipSuccess || hipMalloc(&d_a, ne) != hipSuccess || hipMalloc(&d_b, ne) != hipSuccess ||
      hipMalloc(&d_y, ny) != hipSuccess || hipMalloc(&d_h, nh) != hipSuccess) {
    rc = GAL_E_NOMEM;
  } else if (hipMemcpy(d_kind, kind, ne, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_a, node, ne, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_b, target, ne, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_y, y, ny, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_h, h_step, nh, hipMemcpyHostToDevice) != hipSuccess) {
    rc = GAL_E_HIP;
  } else {
    hipLaunchKernelGGL(event_kernel, dim3((n_events + 255) / 256), dim3(256), 0, 0, n_events, d_kind, d_a, d_b,
                       n_nodes, d_y, d_h, major_ratio);
    if (hipGetLastError() != hipSuccess || hipMemcpy(y, d_y, ny, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h_step, d_h, nh, hipMemcpyDeviceToHost) != hipSuccess)
      rc = GAL_E_HIP;
  }
  hipFree(d_kind); hipFree(d_a); hipFree(d_b); hipFree(d_y); hipFree(d_h);
  return rc;
}

extern "C" int gal_tree_evolve_d(gal_ode_batch* h, const gal_schedule* sched, int ld, double* y, double* scale,
                                 int* status, double major_ratio, void* stream) {
  return tree_evolve_graph(h, sched, ld, y, scale, status, major_ratio, stream);
}

extern "C" int gal_tree_evolve_waves_d(gal_ode_batch* h, const gal_schedule* sched, int ld, double* y, double* scale,
                                       int* status, double major_ratio, void* stream) {
  return tree_evolve(h, sched, ld, y, scale, status, major_ratio, stream, nullptr);
}

extern "C" int gal_tree_evolve_surrogate_d(gal_ode_batch* h, gal_mlp* m, const float* norm,
                                           const gal_surrogate_gate* gate, const gal_schedule* sched, int ld,
                                           double* y, double* scale, int* status, double major_ratio,
                                           int64_t* n_accepted, void* stream) {
  SurrogateRoute sr{m, norm, gate, n_accepted};
  return tree_evolve(h, sched, ld, y, scale, status, major_ratio, stream, &sr);
}

extern "C" int gal_tree_evolve_record_d(gal_ode_batch* h, const gal_schedule* sched, int ld, double* y, double* scale,
                                        int* status, double major_ratio, double* rec_y0, double* rec_y1,
                                        double* rec_t0, double* rec_t1, double* rec_node_par, int64_t capacity,
                                        int64_t* n_recorded, void* stream) {
  if (capacity < 1 || capacity > INT32_MAX) return GAL_E_INVALID;
  Recording rec{rec_y0, rec_y1, rec_t0, rec_t1, rec_node_par, capacity, n_recorded};
  return tree_evolve(h, sched, ld, y, scale, status, major_ratio, stream, nullptr, &rec);
}
