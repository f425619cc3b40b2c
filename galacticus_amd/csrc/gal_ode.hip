// gal_ode.hip -- batched node-ODE evolution on MI355X (gfx950) and its C ABI.
//
// One lane owns one node.  A persistent grid of workgroups claims nodes from
// a global counter (one atomic per wave per refill) and integrates each from
// t_start to t_end with the Cash-Karp (4,5) pair, the scaled error norm and
// step controller, the driver loop and the node-evolver trial loop of the
// reference:
//   trial loop    node_evolver/standard.F90:587-740
//   solve         numerical/ODE_solver/solver.F90:492-691
//   driver        external/gslODEInitVal2/driver2.c:148-250
//   evolve/step   GSL 2.6 gsl_odeiv2_evolve_apply + rkck (published method)
//   controller    external/gslODEInitVal2/cscal2.c:93-169
// When a lane's node finishes it claims the next one, so lanes stay busy
// while step counts differ by two orders of magnitude between nodes.
//
// Schedule.  The lanes of a wave advance in lockstep through a six-phase
// loop, one Cash-Karp attempt per trip:
//   phases 0-4  stage inputs for k2..k6, RHS, store (phase 4: combine,
//               error norm, controller, accept / reject / fail)
//   phase 5     claim nodes for lanes that finished, then one shared RHS
//               slot at (t, y): dydt at the new point (the reference's
//               dydt_out, reused as the next k1), or k1 of a freshly claimed
//               node, or k1 after a trial restart.
// There is a single RHS call site, every lane spends six RHS evaluations per
// trip whatever it is doing, and the RHS never diverges by phase.
//
// Memory.  State is structure-of-arrays in HBM (y[v*ld + i]); it is read once
// when a node is claimed and written once when it finishes.  Within a node,
// the RHS input and output stay in registers, and the five stage
// derivatives k1..k5, the state y and the error scales live in a
// lane-private, bank-conflict-free layout L[(vec*30 + v)*BLOCK + lane]
// (7*30*8 = 1680 bytes per lane): all in LDS (105 KiB per 64-lane
// workgroup), or part of them in an HBM workspace of the same layout so that
// two or four workgroups share a CU (see evolve_kernel).  Nothing is
// addressable private memory, so the long dependent chain of a node never
// waits on scratch.
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <mutex>

#include "../../include/gal_ode.h"
#include "../../include/gal_tree.h"
#include "constants.hpp"
#include "physics.hpp"
#include "integrator.hpp"
#include "batch.hpp"
#include "events.hpp"

#include <hipcub/device/device_radix_sort.hpp>

using namespace gal;

namespace {

constexpr int kStages = 5;  // k1..k5 kept; k6 is consumed as it is produced
constexpr int kLdsVectors = kStages + 2;  // k1..k5, y, scales

struct EvolveArgs {
  int n, ld;
  int n_claim;  // claims are limited to n_claim entries of order[] (n when order is null)
  const int* n_claim_dev;  // if set, the limit is read from device memory instead
  double* y;
  const double* np;
  const double* t_start;
  const double* t_end;
  double* h_step;
  double* scale;
  int* status;
  long long* counters;
  const int* order;
  unsigned int* work;
  double* ywork;  // per-lane state and scales in HBM (variant without them in LDS)
  int ystride;    // lanes in the grid (ywork holds 2*30 doubles per lane)
  double* kwork;  // per-lane stage derivatives in HBM (variant without them in LDS)
  const double* P;
  const double* tab;
  DerivedConstants D;
  CoolingDesc cd;
  double eps_abs, eps_rel;
  int flags, mode;
  // Wall-time limit of the launch in ticks of the GPU's constant-rate
  // realtime counter (0: none).  An attempt that would start after it
  // returns the node with errorStatusXCPU and its state untouched.
  unsigned long long wall_ticks;
  // Operation-graph mode (kernel parameter GRAPH): lanes claim entries of
  // the graph's ready list (a.work counts the claims) instead of nodes.  A
  // lane whose claimed entry stays unwritten while no operation is appended
  // for stall_ticks gives up, sets *graph_error and leaves, so that a
  // malformed graph ends the launch instead of holding it.
  OpGraph g;
  int* graph_error;
  unsigned long long stall_ticks;
};

// Claim one node index per requesting lane with a single atomic per wave.
__device__ __forceinline__ int wave_claim(bool want, unsigned int* work, int n, const int* order) {
  const unsigned long long mask = __ballot(want);
  int idx = -1;
  if (mask) {
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    const int cnt = __popcll(mask);
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(work, (unsigned int)cnt);
    base = __shfl(base, leader);
    const unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
    const unsigned int mine = base + (unsigned int)__popcll(below);
    if (want && mine < (unsigned int)n) idx = order ? order[mine] : (int)mine;
  }
  return idx;
}

__device__ __forceinline__ void write_counters(const EvolveArgs& a, int node, long long na, long long nc,
                                               long long nr) {
  if (a.counters) {
    a.counters[node] = na;
    a.counters[(size_t)a.n + node] = nc;
    a.counters[2 * (size_t)a.n + node] = nr;
  }
}

// Three placements of the lane-private vectors (each [vector][variable][lane]
// per workgroup, so a wave touches 64 consecutive doubles per variable):
//  YLDS && KLDS:  stage derivatives, state and error scales in LDS (105 KiB
//                 per 64 lanes: one workgroup per CU);
//  !YLDS && KLDS: state and scales in an HBM workspace, stages in LDS
//                 (75 KiB: two workgroups per CU);
//  YLDS && !KLDS: stages in an HBM workspace, state and scales in LDS
//                 (30 KiB: four workgroups per CU, one wave on every SIMD).
// The arithmetic is the same in all three; only where the operands wait
// between phases differs.
//
// WAVES = 2 gives each group of BLOCK nodes a second wave, on another SIMD of
// the CU, that computes the hot-halo half of every RHS while the first wave
// computes the galaxy half and runs the integrator.  One wave64 instruction
// costs the same issue time however many of its lanes are active, so a
// node's chain of RHS evaluations -- which sets the launch time when a few
// stiff nodes take thousands of attempts -- shortens by the share of the
// RHS the second wave takes.  Per RHS the first wave publishes the stage
// state in LDS and a command (compute, skip, quit); barrier; both halves;
// barrier; the first wave adds the second's derivatives.  Only the first
// wave touches the integrator's state; the exchange area is written by one
// wave and read by the other on opposite sides of a barrier.
constexpr int kXchgVectors = kNV + 1 + 7 + kNV;   // stage state and time, node parameters, hot-half derivatives
enum : int { kCmdCompute = 1, kCmdSkip = 2, kCmdQuit = 3 };

template <int BLOCK, bool YLDS, bool KLDS>
constexpr size_t lds_vector_bytes() {
  return sizeof(double) * ((YLDS ? 2 : 0) + (KLDS ? kStages : 0)) * kNV * BLOCK;
}

// The kernel's one argument starts the kernarg segment: the derived constants
// and table descriptor are read from there through an offset the compiler
// cannot prove constant, so their scalar loads stay inside each RHS instead
// of being hoisted and held (and spilled) across the whole integration.
struct UniformRefs {
  const double* P;
  const DerivedConstants* D;
  const CoolingDesc* C;
};
__device__ __forceinline__ UniformRefs uniform_refs(const double* P) {
  int opaque = 0;
  asm volatile("" : "+s"(opaque));
  opaque = __builtin_amdgcn_readfirstlane(opaque);
  const char* kargs = (const char*)__builtin_amdgcn_kernarg_segment_ptr() + opaque;
  return UniformRefs{P + opaque, (const DerivedConstants*)(kargs + offsetof(EvolveArgs, D)),
                     (const CoolingDesc*)(kargs + offsetof(EvolveArgs, cd))};
}

template <int BLOCK>
__device__ __forceinline__ void rhs_coprocessor(const double* P, const double* __restrict__ tab, int mode,
                                                double* __restrict__ xchg, volatile int* cmd) {
  const int lane = threadIdx.x - BLOCK;
  double* __restrict__ X = xchg + lane;
  for (;;) {
    __syncthreads();   // stage state published
    const int c = __builtin_amdgcn_readfirstlane(*cmd);
    if (c == kCmdQuit) break;
    if (c == kCmdCompute) {
      double yy[kNV], fh[kNV];
#pragma unroll
      for (int v = 0; v < kNV; ++v) yy[v] = X[v * BLOCK];
      const double ts = X[kNV * BLOCK];
      NodePar np;
      np.t_ref = X[(kNV + 1) * BLOCK];
      np.m_ref = X[(kNV + 2) * BLOCK];
      np.m_rate = X[(kNV + 3) * BLOCK];
      np.spin_rate = X[(kNV + 4) * BLOCK];
      np.ram_force = X[(kNV + 5) * BLOCK];
      np.v_host = X[(kNV + 6) * BLOCK];
      np.satellite = X[(kNV + 7) * BLOCK] != 0.0;
      const UniformRefs u = uniform_refs(P);
      rhs_part<kRhsHalo>(u.P, *u.D, tab, *u.C, mode, np, ts, yy, fh);
#pragma unroll
      for (int v = 0; v < kNV; ++v) X[(kNV + 8 + v) * BLOCK] = fh[v];
    }
    __syncthreads();   // hot-half derivatives published
  }
}

template <int BLOCK, bool YLDS, bool KLDS, int WAVES = 1, bool GRAPH = false>
__global__ __launch_bounds__(BLOCK * WAVES, 1) void evolve_kernel(EvolveArgs a) {
  extern __shared__ double lds_k[];
  __shared__ int xchg_cmd;
  double* __restrict__ xchg = lds_k + lds_vector_bytes<BLOCK, YLDS, KLDS>() / sizeof(double);
  if constexpr (WAVES == 2) {
    if (threadIdx.x >= BLOCK) {
      rhs_coprocessor<BLOCK>(a.P, a.tab, a.mode, xchg, &xchg_cmd);
      return;
    }
  }
  double* __restrict__ XC = xchg + threadIdx.x;    // this lane's column of the exchange area
  double* __restrict__ K = lds_k + threadIdx.x;   // K[(s*kNV + v)*BLOCK] when KLDS
  // In the HBM workspace the stage vectors are reached through a buffer
  // resource over the workgroup's slab: the lane offset sits in one VGPR and
  // the per-vector offsets in SGPRs, so 150 addresses cost no registers.
  const __amdgpu_buffer_rsrc_t krs = __builtin_amdgcn_make_buffer_rsrc(
      KLDS ? nullptr : a.kwork + (size_t)blockIdx.x * (kStages * kNV * BLOCK), (short)0,
      (int)(sizeof(double) * kStages * kNV * BLOCK), 0x00020000);
  const int k_lane = (int)(threadIdx.x * sizeof(double));
  using B64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(krs, 0, 0, 0));
  auto kget = [&](int st, int v) -> double {
    if constexpr (KLDS) {
      return K[(st * kNV + v) * BLOCK];
    } else {
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                            krs, k_lane, (int)sizeof(double) * (st * kNV + v) * BLOCK, 0));
    }
  };
  auto kset = [&](int st, int v, double x) {
    if constexpr (KLDS) {
      K[(st * kNV + v) * BLOCK] = x;
    } else {
      const int su = __builtin_amdgcn_readfirstlane(st);   // the phase is uniform over the wave
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B64, x), krs, k_lane,
                                            (int)sizeof(double) * (su * kNV + v) * BLOCK, 0);
    }
  };
  const double* __restrict__ P = a.P;
  const int mode = a.mode;
  const bool reuse = (a.flags & GAL_FLAG_REUSE_STEP_SIZE) != 0;
  const bool enforce = (a.flags & GAL_FLAG_ENFORCE_NON_NEGATIVITY) != 0;
  const bool scales_given = (a.flags & GAL_FLAG_SCALES_PROVIDED) != 0;
#define KS(s, v) kget((s), (v))

  // Both layouts are [vector][variable][lane] per workgroup: a wave reads
  // 64 consecutive doubles per variable.
  constexpr int ys = BLOCK;
  double* __restrict__ YL = YLDS ? lds_k + (KLDS ? kStages * kNV * BLOCK : 0) + threadIdx.x
                                 : a.ywork + (size_t)blockIdx.x * (2 * kNV * BLOCK) + threadIdx.x;
  double* __restrict__ SL = YL + kNV * BLOCK;
#define Y(v) YL[(v) * ys]
#define SC(v) SL[(v) * ys]
  // Partial sums over the stage derivatives already in memory, formed while
  // an RHS is in flight: for phases 0-3 the next stage's input sum without
  // its newest term; for phase 4 the fifth-order solution's sum without k6.
  // Each is completed, in the order of the full sum, once the newest
  // derivative is known.
  // Coefficients of the newest term of the stage input that follows phase ph (0-3).
  auto stage_b_new = [](int ph) {
    return ph == 0 ? ck::b32 : (ph == 1 ? ck::b43 : (ph == 2 ? ck::b54 : ck::b65));
  };
  auto stage_a_new = [](int ph) { return ph == 0 ? ck::a3 : (ph == 1 ? ck::a4 : (ph == 2 ? ck::a5 : ck::a6)); };
  // The LDS operands are gathered ten variables at a time before the
  // arithmetic, so that their reads are in flight together rather than one
  // read-and-wait per variable.
  constexpr int kChunk = 10;
  static_assert(kNV % kChunk == 0, "chunks must tile the state");
  auto stage_partials = [&](double (&sp)[kNV], int ph, bool act) {
    if (!act) return;
#pragma unroll
    for (int c = 0; c < kNV; c += kChunk) {
      double g0[kChunk], g1[kChunk], g2[kChunk], g3[kChunk];
      switch (ph) {
        case 0:
#pragma unroll
          for (int i = 0; i < kChunk; ++i) g0[i] = KS(0, c + i);
#pragma unroll
          for (int i = 0; i < kChunk; ++i) sp[c + i] = ck::b31 * g0[i];
          break;
        case 1:
#pragma unroll
          for (int i = 0; i < kChunk; ++i) { g0[i] = KS(0, c + i); g1[i] = KS(1, c + i); }
#pragma unroll
          for (int i = 0; i < kChunk; ++i) sp[c + i] = ck::b41 * g0[i] + ck::b42 * g1[i];
          break;
        case 2:
#pragma unroll
          for (int i = 0; i < kChunk; ++i) { g0[i] = KS(0, c + i); g1[i] = KS(1, c + i); g2[i] = KS(2, c + i); }
#pragma unroll
          for (int i = 0; i < kChunk; ++i) sp[c + i] = ck::b51 * g0[i] + ck::b52 * g1[i] + ck::b53 * g2[i];
          break;
        case 3:
#pragma unroll
          for (int i = 0; i < kChunk; ++i) {
            g0[i] = KS(0, c + i); g1[i] = KS(1, c + i); g2[i] = KS(2, c + i); g3[i] = KS(3, c + i);
          }
#pragma unroll
          for (int i = 0; i < kChunk; ++i)
            sp[c + i] = ck::b61 * g0[i] + ck::b62 * g1[i] + ck::b63 * g2[i] + ck::b64 * g3[i];
          break;
        case 4:
#pragma unroll
          for (int i = 0; i < kChunk; ++i) { g0[i] = KS(0, c + i); g2[i] = KS(2, c + i); g3[i] = KS(3, c + i); }
#pragma unroll
          for (int i = 0; i < kChunk; ++i) sp[c + i] = ck::c1 * g0[i] + ck::c3 * g2[i] + ck::c4 * g3[i];
          break;
        default:
          break;
      }
    }
  };
  NodePar np{};
  int node = -1;
  int op = -1, pend = -1;                  // graph mode: operation in hand, claimed ready-list entry
  unsigned int wait_tail = ~0u;            // graph mode: appends seen while waiting on the entry
  unsigned long long wait_since = 0ull;
  bool active = false, want = true, eval_pending = false, post_pending = false, final_step = false;
  double t = 0.0, t0 = 0.0, t1 = 0.0, h = 0.0, h0 = 0.0, h_node = 0.0, sign = 1.0;
  int trial = 0;
  long long n_attempt = 0, n_accept = 0, n_rhs = 0;
  const int n_claim = a.n_claim_dev ? *a.n_claim_dev : a.n_claim;

  int phase = 5;
  double ts = 0.0;   // time of the stage input ytmp (carried over for phases 1-4)
  const unsigned long long t_launch = a.wall_ticks ? wall_clock64() : 0ull;
  for (;;) {
    // Stage input, stage derivative and partial sums: rewritten before every
    // use within an iteration, so nothing of them is carried round the loop.
    double ytmp[kNV], kout[kNV], sp[kNV];
    bool do_rhs = false;
    if (phase == 5) {
      // ---- claim nodes for lanes that need one (convergent point) ------
      bool again = true;
      while (again) {
        bool begun = false;   // this lane takes a node's interval
        if constexpr (GRAPH) {
          // claim an entry of the ready list, then wait for it to be written;
          // events are applied on the spot and their successors released
          const bool need = want && pend < 0;
          if (__any(need)) {
            const int j = wave_claim(need, a.work, a.g.n, nullptr);
            if (need) {
              if (j < 0) {
                want = false;
                active = false;
              } else {
                pend = j;
                wait_tail = ~0u;
              }
            }
          }
          if (want && pend >= 0) {
            const int k = graph_ready_entry(a.g, pend);
            if (k >= 0) {
              pend = -1;
              const gal_schedule_op* o = a.g.ops + k;
              const int kind = o->kind;
              if (kind != GAL_OP_EVOLVE) {
                apply_event(kind, (size_t)o->a, (size_t)o->b, a.ld, a.y, a.h_step, a.g.major_ratio);
                graph_release(a.g, k);
              } else {
                op = k;
                node = o->a;
                np = load_node_par(o->node_par, 1, 0);
                t0 = o->t0;
                t1 = o->t1;
                begun = true;
              }
            } else {
              const unsigned int tl = __hip_atomic_load(a.g.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              const unsigned long long now = wall_clock64();
              if (tl != wait_tail) {
                wait_tail = tl;
                wait_since = now;
              } else if (now - wait_since > a.stall_ticks) {
                __hip_atomic_store(a.graph_error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                want = false;
                pend = -1;
              }
            }
          }
        } else {
          // an order entry outside the batch (a plan the caller did not let
          // finish) ends this lane's claims instead of addressing past it
          const int claimed = wave_claim(want, a.work, n_claim, a.order);
          const int idx = claimed < a.n ? claimed : -1;
          if (want) {
            if (idx < 0) {
              want = false;
              active = false;
            } else {
              node = idx;
              np = load_node_par(a.np, a.ld, node);
              t0 = a.t_start[node];
              t1 = a.t_end[node];
              begun = true;
            }
          }
        }
        if (begun) {
          if constexpr (WAVES == 2) {
            XC[(kNV + 1) * BLOCK] = np.t_ref;
            XC[(kNV + 2) * BLOCK] = np.m_ref;
            XC[(kNV + 3) * BLOCK] = np.m_rate;
            XC[(kNV + 4) * BLOCK] = np.spin_rate;
            XC[(kNV + 5) * BLOCK] = np.ram_force;
            XC[(kNV + 6) * BLOCK] = np.v_host;
            XC[(kNV + 7) * BLOCK] = np.satellite ? 1.0 : 0.0;
          }
          h_node = a.h_step[node];
#pragma unroll
          for (int v = 0; v < kNV; ++v) ytmp[v] = a.y[v * (size_t)a.ld + node];
          if (scales_given) {
#pragma unroll
            for (int v = 0; v < kNV; ++v) kout[v] = a.scale[v * (size_t)a.ld + node];
          } else if (mode == GAL_PHYSICS_TEST) {
#pragma unroll
            for (int v = 0; v < kNV; ++v) kout[v] = 1.0;
          } else {
            scales(P, a.D, np, t0, ytmp, kout);
          }
          if (!scales_given && a.scale) {
#pragma unroll
            for (int v = 0; v < kNV; ++v) a.scale[v * (size_t)a.ld + node] = kout[v];
          }
#pragma unroll
          for (int v = 0; v < kNV; ++v) {
            Y(v) = ytmp[v];
            SC(v) = kout[v];
          }
          n_attempt = n_accept = n_rhs = 0;
          trial = 0;
          if (t0 == t1) {
            // nothing to integrate (standard.F90:580): success, timeStep -1
            a.h_step[node] = -1.0;
            a.status[node] = GAL_STATUS_SUCCESS;
            write_counters(a, node, 0, 0, 0);
            if constexpr (GRAPH) graph_release(a.g, op);
          } else {
            want = false;
            active = true;
            // odeSolverSolve initial step (solver.F90:537-555)
            const double step_size = reuse ? h_node : -1.0;
            double xs = t1 - t0;
            if (step_size > 0.0) xs = fmin(step_size, xs);
            h = xs;
            h0 = h;
            sign = h > 0.0 ? 1.0 : -1.0;
            t = t0;
            eval_pending = true;
          }
        }
        if constexpr (GRAPH) again = __any(want && pend < 0);
        else again = __any(want);
      }
      if (eval_pending) {
        do_rhs = true;
        ts = t;
#pragma unroll
        for (int v = 0; v < kNV; ++v) ytmp[v] = Y(v);
      }
    } else if (active) {
      do_rhs = true;
      switch (phase) {
        case 0: {
          // Wall-time limit (node_evolver/standard.F90:693-705 and the RHS's
          // own check, standard.F90:860-866): the node returns
          // errorStatusXCPU with its state as it came in.
          if (a.wall_ticks && wall_clock64() - t_launch > a.wall_ticks) {
            a.status[node] = GAL_STATUS_XCPU;
            write_counters(a, node, n_attempt, n_accept, n_rhs);
            if constexpr (GRAPH) graph_release(a.g, op);
            active = false;
            want = true;
            do_rhs = false;
            break;
          }
          // gsl_odeiv2_evolve_apply: clip the step to reach t1 exactly
          const double dt = t1 - t;
          final_step = false;
          if ((dt >= 0.0 && h0 > dt) || (dt < 0.0 && h0 < dt)) {
            h0 = dt;
            final_step = true;
          }
          ++n_attempt;
          ts = t + ck::a2 * h0;
#pragma unroll
          for (int c = 0; c < kNV; c += kChunk) {
            double gy[kChunk], gk[kChunk];
#pragma unroll
            for (int i = 0; i < kChunk; ++i) { gy[i] = Y(c + i); gk[i] = KS(0, c + i); }
#pragma unroll
            for (int i = 0; i < kChunk; ++i) ytmp[c + i] = gy[i] + h0 * (ck::b21 * gk[i]);
          }
          break;
        }
        default:
          // phases 1-4: y + h (sum of the older stage terms + the newest)
          if constexpr (WAVES == 2) {
            // formed at the end of the previous iteration, in the exchange area
            ts = XC[kNV * BLOCK];
#pragma unroll
            for (int v = 0; v < kNV; ++v) ytmp[v] = XC[v * BLOCK];
          } else {
            stage_partials(sp, phase - 1, true);
            const double b_new = stage_b_new(phase - 1);
            ts = t + stage_a_new(phase - 1) * h0;
#pragma unroll
            for (int c = 0; c < kNV; c += kChunk) {
              double gy[kChunk], gk[kChunk];
#pragma unroll
              for (int i = 0; i < kChunk; ++i) { gy[i] = Y(c + i); gk[i] = KS(phase, c + i); }
#pragma unroll
              for (int i = 0; i < kChunk; ++i) ytmp[c + i] = gy[i] + h0 * (sp[c + i] + b_new * gk[i]);
            }
          }
          break;
      }
    }

    if constexpr (WAVES == 2) {
      const bool any_rhs = __any(do_rhs);
      if (phase == 0 || phase == 5) {
#pragma unroll
        for (int v = 0; v < kNV; ++v) XC[v * BLOCK] = ytmp[v];
        XC[kNV * BLOCK] = ts;
      }
      if (threadIdx.x == 0) xchg_cmd = any_rhs ? kCmdCompute : kCmdSkip;
      __syncthreads();   // stage state published
      if (do_rhs) {
        const UniformRefs u = uniform_refs(P);
        rhs_part<kRhsGalaxy>(u.P, *u.D, a.tab, *u.C, mode, np, ts, ytmp, kout);
        ++n_rhs;
      }
      stage_partials(sp, phase, active);   // while the hot half is still in flight
      __syncthreads();   // hot-half derivatives published
      if (do_rhs) {
        double kh[kNV];
#pragma unroll
        for (int v = 0; v < kNV; ++v) kh[v] = XC[(kNV + 8 + v) * BLOCK];
#pragma unroll
        for (int v = 0; v < kNV; ++v) kout[v] += kh[v];
      }
    } else {
      if (do_rhs) {
        const UniformRefs u = uniform_refs(P);
        rhs(u.P, *u.D, a.tab, *u.C, mode, np, ts, ytmp, kout);
        ++n_rhs;
      }
      if (phase == 4) stage_partials(sp, phase, active);
    }

    if (phase < 4) {
      if (active) {
#pragma unroll
        for (int v = 0; v < kNV; ++v) kset(phase + 1, v, kout[v]);
        if constexpr (WAVES == 2) {
          // the next stage's input, straight into the exchange area (the
          // other wave is past its reads and waits for the next publish)
          const double b_new = stage_b_new(phase);
          XC[kNV * BLOCK] = t + stage_a_new(phase) * h0;
#pragma unroll
          for (int c = 0; c < kNV; c += kChunk) {
            double gy[kChunk];
#pragma unroll
            for (int i = 0; i < kChunk; ++i) gy[i] = Y(c + i);
#pragma unroll
            for (int i = 0; i < kChunk; ++i) XC[(c + i) * BLOCK] = gy[i] + h0 * (sp[c + i] + b_new * kout[c + i]);
          }
        }
      }
      ++phase;
      continue;
    }
    if (phase == 5) {
      if (eval_pending) {
#pragma unroll
        for (int v = 0; v < kNV; ++v) kset(0, v, kout[v]);
        eval_pending = false;
        if (post_pending) {
          post_step(P, a.D, np, mode, t, &Y(0), ys);   // GSL_CONTINUE: no evolver reset
          post_pending = false;
        }
      }
      if (!__any(active)) {
        if constexpr (GRAPH) {
          // lanes wait on ready-list entries: poll again after a pause
          if (__any(want)) {
            __builtin_amdgcn_s_sleep(8);
            continue;
          }
        }
        if constexpr (WAVES == 2) {
          if (threadIdx.x == 0) xchg_cmd = kCmdQuit;
          __syncthreads();
        }
        break;
      }
      phase = 0;
      continue;
    }

    // ---- phase 4: k6 is in kout; combine, error norm, controller --------
    phase = 5;
    if (!active) continue;
    ErrorNorm en;
#pragma unroll
    for (int c = 0; c < kNV; c += kChunk) {
      double gy[kChunk], g1[kChunk], g3[kChunk], g4[kChunk], g5[kChunk], gs[kChunk];
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        gy[i] = Y(c + i); g1[i] = KS(0, c + i); g3[i] = KS(2, c + i); g4[i] = KS(3, c + i); g5[i] = KS(4, c + i);
        gs[i] = SC(c + i);
      }
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        const int v = c + i;
        const double yn = gy[i] + h0 * (sp[v] + ck::c6 * kout[v]);
        const double ye = h0 * (ck::ec1 * g1[i] + ck::ec3 * g3[i] + ck::ec4 * g4[i] + ck::ec5 * g5[i] +
                                ck::ec6 * kout[v]);
        ytmp[v] = yn;
        error_norm_add(en, a.eps_abs, a.eps_rel, 1.0, 0.0, h0, yn, ye, 0.0, gs[i], enforce && var_non_negative(v));
      }
    }
    double h_adj = h0;
    const int decision = hadjust_decide(en, kRkckOrder, h_adj);
    const double t_adv = final_step ? t1 : t + h0;
    bool failed = false;
    if (decision == kHadjDec) {
      if (fabs(h_adj) < fabs(h0) && t_adv + h_adj != t_adv) h0 = h_adj;  // reject: retry, same k1
      else failed = true;                                                 // step size underflow
    } else {
      ++n_accept;
      t = t_adv;
      if (!final_step) h = h_adj;
      h0 = h;
#pragma unroll
      for (int v = 0; v < kNV; ++v) Y(v) = ytmp[v];
      if (sign * (t1 - t) > 0.0) {
        eval_pending = true;   // dydt at the new point, then the post-step
        post_pending = true;
      } else {
        post_step(P, a.D, np, mode, t, &Y(0), ys);
        // odeSolverSolve returns xStep = driver h (solver.F90:593)
#pragma unroll
        for (int v = 0; v < kNV; ++v) a.y[v * (size_t)a.ld + node] = Y(v);
        a.h_step[node] = h;
        a.status[node] = GAL_STATUS_SUCCESS;
        write_counters(a, node, n_attempt, n_accept, n_rhs);
        if constexpr (GRAPH) graph_release(a.g, op);
        active = false;
        want = true;
      }
    }
    if (failed) {
      // node-evolver trial loop (standard.F90:657-692)
      bool rescued = false;
      if (enforce) {
        bool any = false;
#pragma unroll
        for (int v = 0; v < kNV; ++v) any |= (Y(v) < 0.0 && var_non_negative(v));
        if (any) {
#pragma unroll
          for (int v = 0; v < kNV; ++v)
            if (Y(v) < 0.0 && var_non_negative(v)) Y(v) = 0.0;
          rescued = true;
        }
      }
      if (rescued) {
#pragma unroll
        for (int v = 0; v < kNV; ++v) a.y[v * (size_t)a.ld + node] = Y(v);
        a.h_step[node] = h_node / (double)(1 << trial);
        a.status[node] = GAL_STATUS_SUCCESS;
        write_counters(a, node, n_attempt, n_accept, n_rhs);
        if constexpr (GRAPH) graph_release(a.g, op);
        active = false;
        want = true;
      } else if (++trial < 8) {
#pragma unroll
        for (int v = 0; v < kNV; ++v) Y(v) = a.y[v * (size_t)a.ld + node];
        const double step_size = reuse ? h_node / (double)(1 << trial) : -1.0;
        double xs = t1 - t0;
        if (step_size > 0.0) xs = fmin(step_size, xs);
        h = xs;
        h0 = h;
        t = t0;
        eval_pending = true;
      } else {
        // all trials failed: state restored, errorStatusUnderflow (standard.F90:715-718)
        a.status[node] = GAL_STATUS_UNDERFLOW;
        write_counters(a, node, n_attempt, n_accept, n_rhs);
        if constexpr (GRAPH) graph_release(a.g, op);
        active = false;
        want = true;
      }
    }
  }
#undef KS
#undef Y
#undef SC
}

__global__ void reset_counter_kernel(unsigned int* w) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *w = 0u;
}

__global__ void rhs_kernel(int n, int ld, const double* __restrict__ t, const double* __restrict__ y,
                           const double* __restrict__ npa, double* __restrict__ dydt, const double* __restrict__ P,
                           const double* __restrict__ tab, DerivedConstants D, CoolingDesc cd, int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double yy[kNV], f[kNV];
#pragma unroll
  for (int v = 0; v < kNV; ++v) yy[v] = y[v * (size_t)ld + i];
  const NodePar np = load_node_par(npa, ld, i);
  rhs(P, D, tab, cd, mode, np, t[i], yy, f);
#pragma unroll
  for (int v = 0; v < kNV; ++v) dydt[v * (size_t)ld + i] = f[v];
}

__global__ void hadjust_kernel(int n, int dim, int ord, double eps_abs, double eps_rel, double a_y, double a_dydt,
                               const double* __restrict__ y, const double* __restrict__ yerr,
                               const double* __restrict__ yp, const double* __restrict__ scale,
                               const int* __restrict__ nonneg, double* __restrict__ h, int* __restrict__ result) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ErrorNorm en;
  const double h_old = h[i];
  for (int v = 0; v < dim; ++v) {
    const size_t k = v * (size_t)n + i;
    error_norm_add(en, eps_abs, eps_rel, a_y, a_dydt, h_old, y[k], yerr[k], yp[k], scale[k], nonneg[k] != 0);
  }
  double hh = h_old;
  result[i] = hadjust_decide(en, ord, hh);
  h[i] = hh;
}

// Stiffness probe for the claim order: one RHS evaluation per node at its
// start point; cost key = |t_end - t_start| / sqrt(tau), tau = the shortest
// e-folding time min_v (|y_v| + scale_v) / |dy_v/dt|.  Only the order in
// which lanes claim nodes depends on it, never a result.
__global__ void cost_probe_kernel(int n_items, const int* __restrict__ subset, int ld, const double* __restrict__ y,
                                  const double* __restrict__ npa, const double* __restrict__ t_start,
                                  const double* __restrict__ t_end, const double* __restrict__ scale_in,
                                  const double* __restrict__ P, const double* __restrict__ tab, DerivedConstants D,
                                  CoolingDesc cd, int mode, int scales_given, const int* __restrict__ n_valid,
                                  float* __restrict__ key, int* __restrict__ index) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_items) return;
  if (n_valid && j >= *n_valid) {
    // beyond the count the caller's list is unset: sorted last, never claimed
    key[j] = -1.0f;
    index[j] = 0;
    return;
  }
  const int i = subset ? subset[j] : j;
  double yy[kNV], f[kNV], sc[kNV];
#pragma unroll
  for (int v = 0; v < kNV; ++v) yy[v] = y[v * (size_t)ld + i];
  const NodePar np = load_node_par(npa, ld, i);
  const double t0 = t_start[i];
  if (scales_given) {
#pragma unroll
    for (int v = 0; v < kNV; ++v) sc[v] = scale_in[v * (size_t)ld + i];
  } else if (mode == GAL_PHYSICS_TEST) {
#pragma unroll
    for (int v = 0; v < kNV; ++v) sc[v] = 1.0;
  } else {
    scales(P, D, np, t0, yy, sc);
  }
  rhs(P, D, tab, cd, mode, np, t0, yy, f);
  double tau = 1.0e30;
#pragma unroll
  for (int v = 0; v < kNV; ++v) {
    const double r = fabs(f[v]);
    const double tv = (fabs(yy[v]) + fabs(sc[v])) / r;
    tau = (r > 0.0 && tv < tau) ? tv : tau;
  }
  const double k = fabs(t_end[i] - t0) * rsqrt(fmax(tau, 1.0e-30));
  key[j] = (float)fmin(k, 3.0e38);
  index[j] = i;
}

template <int BLOCK, bool YLDS, bool KLDS, int WAVES = 1>
constexpr size_t lds_bytes() {
  return lds_vector_bytes<BLOCK, YLDS, KLDS>() + (WAVES == 2 ? sizeof(double) * kXchgVectors * BLOCK : 0);
}

// The dynamic LDS of every launch shape of the evolve kernel, in node mode
// and in operation-graph mode.
template <bool GRAPH>
hipError_t set_lds_attributes() {
  const struct { const void* f; size_t bytes; } shapes[] = {
      {(const void*)evolve_kernel<64, true, true, 2, GRAPH>, lds_bytes<64, true, true, 2>()},
      {(const void*)evolve_kernel<64, true, false, 2, GRAPH>, lds_bytes<64, true, false, 2>()},
      {(const void*)evolve_kernel<64, true, true, 1, GRAPH>, lds_bytes<64, true, true>()},
      {(const void*)evolve_kernel<64, false, true, 1, GRAPH>, lds_bytes<64, false, true>()},
      {(const void*)evolve_kernel<64, true, false, 1, GRAPH>, lds_bytes<64, true, false>()},
  };
  for (const auto& sh : shapes) {
    const hipError_t e = hipFuncSetAttribute(sh.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}


}  // namespace

// ============================================================================
// C ABI
// ============================================================================
static int half_power_code(double alpha) {
  const double a2 = 2.0 * alpha;
  if (a2 == std::floor(a2) && std::fabs(a2) <= 16.0) return (int)a2;
  return kNoHalfPower;
}

static void derive_constants(const double* p, DerivedConstants& D) {
  const double om = p[GAL_P_OMEGA_MATTER], ol = p[GAL_P_OMEGA_DARK_ENERGY], h0 = p[GAL_P_HUBBLE_CONSTANT];
  D.om = om;
  D.ol = ol;
  D.a3_norm = om / ol;
  D.sinh_arg = 1.5 * std::sqrt(ol) * h0 / kMpcPerKmsToGyr;
  const double rho_crit0 = 3.0 * h0 * h0 / (8.0 * kPi * kGravity);
  D.rvir_coeff = 3.0 / (4.0 * kPi * rho_crit0);
  const double ar = 1.0 / (1.0 + p[GAL_P_REDSHIFT_REIONIZATION]);
  D.a3_reion = ar * ar * ar;
  D.f_baryon = p[GAL_P_OMEGA_BARYON] / om;
  D.inv_200 = 1.0 / 200.0;
  D.inv_sqrt_200 = 1.0 / std::sqrt(200.0);
  D.ln_200 = std::log(200.0);
  D.v_fb_disk = p[GAL_P_FB_DISK_VELOCITY];
  D.sqrt_v_fb_disk = std::sqrt(D.v_fb_disk);
  D.ln_v_fb_disk = std::log(D.v_fb_disk);
  D.v_fb_sph = p[GAL_P_FB_SPHEROID_VELOCITY];
  D.sqrt_v_fb_sph = std::sqrt(D.v_fb_sph);
  D.ln_v_fb_sph = std::log(D.v_fb_sph);
  D.n2_sf_disk = half_power_code(p[GAL_P_SF_DISK_EXPONENT_VELOCITY]);
  D.n2_sf_sph = half_power_code(p[GAL_P_SF_SPHEROID_EXPONENT_VELOCITY]);
  D.n2_fb_disk = half_power_code(p[GAL_P_FB_DISK_EXPONENT]);
  D.n2_fb_sph = half_power_code(p[GAL_P_FB_SPHEROID_EXPONENT]);
  D.t_vir_per_v2 = 0.5 * kAtomicMassUnit * kMeanAtomicMassPrimordial * kKilo * kKilo / kBoltzmann;
  D.cs2_per_t = 5.0 / 3.0 * kBoltzmann / (kMeanAtomicMassPrimordial * kAtomicMassUnit) / (kKilo * kKilo);
  D.cs2_sph = D.cs2_per_t * p[GAL_P_BH_TEMPERATURE_SPHEROID];
  D.m_edd_per_mass = 4.0 * kPi * kGravitySI * kMassHydrogenAtom * kGigaYear / kThomson / kSpeedLight;
  const double unit = kErgs * kGigaYear / kMassSolar / (kKilo * kKilo);
  // (M/1e9)^0.9 (Mdot/(k M))^1.2 = 1e9^-0.9 k^-1.2 M^-0.3 Mdot^1.2
  const double ln_common = -0.9 * std::log(1.0e9) - 1.2 * std::log(D.m_edd_per_mass);
  D.ln_jet_kerr = 42.7 * std::log(10.0) + std::log(unit) + ln_common;
  D.ln_jet_schw = 41.7 * std::log(10.0) + std::log(unit) + ln_common;
  D.cool_k_pre = p[GAL_P_DEGREES_OF_FREEDOM] / 2.0 * kBoltzmann / kErgs / kGigaYear;
  D.unit_energy = kMassSolar * kKilo * kKilo / kErgs / kGigaYear;
  const double mpc_cm = kHecto * kMegaParsec;
  D.nh_coeff = kMassSolar / kMassHydrogenAtom / (mpc_cm * mpc_cm * mpc_cm);
  D.inv_eff_sf_disk = p[GAL_P_SF_DISK_EFFICIENCY] != 0.0 ? 1.0 / p[GAL_P_SF_DISK_EFFICIENCY] : 0.0;
  D.inv_eff_sf_sph = p[GAL_P_SF_SPHEROID_EFFICIENCY] != 0.0 ? 1.0 / p[GAL_P_SF_SPHEROID_EFFICIENCY] : 0.0;
  D.inv_tmin_fb_disk = 1.0 / p[GAL_P_FB_DISK_TIMESCALE_MINIMUM];
  D.inv_tmin_fb_sph = 1.0 / p[GAL_P_FB_SPHEROID_TIMESCALE_MINIMUM];
  D.inv_core = 1.0 / p[GAL_P_CORE_RADIUS_OVER_VIRIAL];
  D.inv_cs2_sph = D.cs2_sph > 0.0 ? 1.0 / D.cs2_sph : 0.0;
  const double bhl_unit = kKilo * kGigaYear / kMegaParsec * 4.0 * kPi;
  D.bhl_sph_coeff = D.cs2_sph > 0.0 ? p[GAL_P_BH_ENHANCEMENT_SPHEROID] * bhl_unit * kGravity * kGravity /
                                          (D.cs2_sph * std::sqrt(D.cs2_sph))
                                    : 0.0;
  D.jeans_coeff = D.cs2_sph > 0.0 ? kGravity / D.cs2_sph : 0.0;
  D.h0_gyr = h0 / kMpcPerKmsToGyr;
  D.inv_one_minus_j_loss = 1.0 / (1.0 - p[GAL_P_FRACTION_LOSS_ANGULAR_MOMENTUM]);
  // Ciotti & Ostriker wind coupling (black_holes/winds/Ciotti2009.F90:156-175):
  // P_ISM / P_wind = 3 (3/2) k T_ISM v_wind Gyr / (m_H c^2 Mpc) * M_gas / (r eff Mdot)
  {
    constexpr double kVelocityWind = 1.0e4, kTemperatureISM = 1.0e4;  // km/s, K
    D.wind_pressure_ratio = 3.0 * 1.5 * kBoltzmann * kTemperatureISM * kGigaYear * (kVelocityWind * kKilo) /
                            (kMassHydrogenAtom * kSpeedLight * kSpeedLight * kMegaParsec);
  }
}

extern "C" {

int gal_ode_version(void) { return 10000; }

int gal_ode_device_check(int* n_devices, int* n_cus) {
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) {
    if (n_devices) *n_devices = 0;
    return GAL_E_NODEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return GAL_E_HIP;
  if (n_devices) *n_devices = nd;
  if (n_cus) *n_cus = prop.multiProcessorCount;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GAL_E_NODEVICE;
  return GAL_OK;
}

int gal_ode_batch_create(int n_max, double tol_abs, double tol_rel, int flags, gal_ode_batch** out) {
  if (!out || n_max < 0) return GAL_E_INVALID;
  if (tol_abs < 0.0 || tol_rel < 0.0 || (tol_abs <= 0.0 && tol_rel <= 0.0)) return GAL_E_INVALID;
  *out = nullptr;
  int nd = 0, ncu = 0;
  int rc = gal_ode_device_check(&nd, &ncu);
  if (rc != GAL_OK) return rc;
  gal_ode_batch* h = new gal_ode_batch();
  h->n_max = n_max;
  h->tol_abs = tol_abs;
  h->tol_rel = tol_rel;
  h->flags = flags;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return GAL_E_HIP;
  }
  h->n_cus = ncu;
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device) == hipSuccess && khz > 0)
      h->wall_clock_hz = 1.0e3 * khz;
    else
      (void)hipGetLastError();
  }
  const size_t nm = (size_t)std::max(n_max, 1);
  if (hipMalloc(&h->d_params, sizeof(double) * GAL_NPARAM) != hipSuccess ||
      hipMalloc(&h->d_table, sizeof(double) * GAL_COOLING_NT_MAX * GAL_COOLING_NZ_MAX) != hipSuccess ||
      hipMalloc(&h->d_work, sizeof(unsigned int)) != hipSuccess ||
      hipMalloc(&h->d_order, sizeof(int) * nm) != hipSuccess || hipMalloc(&h->d_index, sizeof(int) * nm) != hipSuccess ||
      hipMalloc(&h->d_key, sizeof(float) * nm) != hipSuccess ||
      hipMalloc(&h->d_key_sorted, sizeof(float) * nm) != hipSuccess) {
    gal_ode_batch_destroy(h);
    return GAL_E_NOMEM;
  }
  if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, h->sort_temp_bytes, h->d_key, h->d_key_sorted,
                                                   h->d_index, h->d_order, (int)nm) != hipSuccess ||
      hipMalloc(&h->d_sort_temp, std::max<size_t>(h->sort_temp_bytes, 16)) != hipSuccess) {
    gal_ode_batch_destroy(h);
    return GAL_E_NOMEM;
  }
  h->ywork_lanes = ncu * 2 * 64;
  h->kwork_lanes = ncu * 4 * 64;
  if (hipMalloc(&h->d_ywork, sizeof(double) * 2 * kNV * (size_t)h->ywork_lanes) != hipSuccess ||
      hipMalloc(&h->d_kwork, sizeof(double) * kStages * kNV * (size_t)h->kwork_lanes) != hipSuccess) {
    gal_ode_batch_destroy(h);
    return GAL_E_NOMEM;
  }
  // Set on every handle creation (on the current device): no process-wide
  // state, and cheap.
  {
    static_assert(lds_bytes<64, true, true>() <= 163840, "lane-private LDS vectors must fit 160 KiB");
    static_assert(2 * lds_bytes<64, false, true>() <= 163840, "two workgroups per CU must fit 160 KiB");
    static_assert(4 * lds_bytes<64, true, false>() <= 163840, "four workgroups per CU must fit 160 KiB");
    static_assert(lds_bytes<64, true, true, 2>() + 16 <= 163840, "two-wave groups with LDS stages must fit 160 KiB");
    static_assert(2 * (lds_bytes<64, true, false, 2>() + 16) <= 163840, "two two-wave groups per CU must fit");
    if (set_lds_attributes<false>() != hipSuccess || set_lds_attributes<true>() != hipSuccess) {
      (void)hipGetLastError();
      gal_ode_batch_destroy(h);
      return GAL_E_HIP;
    }
  }
  *out = h;
  return GAL_OK;
}

static void free_staging(gal_ode_batch* h) {
  hipFree(h->d_y); hipFree(h->d_np); hipFree(h->d_t0); hipFree(h->d_t1); hipFree(h->d_h);
  hipFree(h->d_scale); hipFree(h->d_status); hipFree(h->d_counters);
  h->d_y = h->d_np = h->d_t0 = h->d_t1 = h->d_h = h->d_scale = nullptr;
  h->d_status = nullptr;
  h->d_counters = nullptr;
}

int gal_ode_batch_destroy(gal_ode_batch* h) {
  if (!h) return GAL_E_INVALID;
  hipFree(h->d_params);
  hipFree(h->d_table);
  hipFree(h->d_work);
  hipFree(h->d_ywork);
  hipFree(h->d_kwork);
  hipFree(h->d_order);
  hipFree(h->d_index);
  hipFree(h->d_key);
  hipFree(h->d_key_sorted);
  hipFree(h->d_sort_temp);
  free_staging(h);
  delete h;
  return GAL_OK;
}

static bool physics_valid(const double* params) {
  const int mode = (int)params[GAL_P_PHYSICS_MODE];
  if (mode < GAL_PHYSICS_STANDARD || mode > GAL_PHYSICS_TEST) return false;
  if (!(params[GAL_P_OMEGA_MATTER] > 0.0) || !(params[GAL_P_OMEGA_DARK_ENERGY] > 0.0) ||
      !(params[GAL_P_HUBBLE_CONSTANT] > 0.0) || !(params[GAL_P_FB_DISK_VELOCITY] > 0.0) ||
      !(params[GAL_P_FB_SPHEROID_VELOCITY] > 0.0) || !(params[GAL_P_FB_DISK_TIMESCALE_MINIMUM] > 0.0) ||
      !(params[GAL_P_FB_SPHEROID_TIMESCALE_MINIMUM] > 0.0) || !(params[GAL_P_CORE_RADIUS_OVER_VIRIAL] > 0.0) ||
      !(params[GAL_P_FRACTION_LOSS_ANGULAR_MOMENTUM] < 1.0) || !(params[GAL_P_OPERATORS] >= 0.0) ||
      !(params[GAL_P_OPERATORS] <= 15.0))
    return false;
  return true;
}

int gal_ode_batch_set_physics(gal_ode_batch* h, const double* params, int n_params) {
  if (!h || !params || n_params != GAL_NPARAM) return GAL_E_INVALID;
  if (!physics_valid(params)) return GAL_E_INVALID;
  std::memcpy(h->params, params, sizeof(double) * GAL_NPARAM);
  derive_constants(h->params, h->D);
  HIP_TRY(hipMemcpy(h->d_params, h->params, sizeof(double) * GAL_NPARAM, hipMemcpyHostToDevice));
  h->have_physics = true;
  return GAL_OK;
}

static bool cooling_table_valid(const double* ln_lambda, int n_t, double d_ln_t, int n_z, double d_ln_z) {
  return ln_lambda && n_t >= 2 && n_z >= 2 && n_t <= GAL_COOLING_NT_MAX && n_z <= GAL_COOLING_NZ_MAX &&
         d_ln_t > 0.0 && d_ln_z > 0.0;
}

static void cooling_desc(CoolingDesc& c, int n_t, double ln_t_min, double d_ln_t, int n_z, double ln_z_min,
                         double d_ln_z) {
  c.n_t = n_t;
  c.n_z = n_z;
  c.ln_t_min = ln_t_min;
  c.d_ln_t = d_ln_t;
  c.inv_d_ln_t = 1.0 / d_ln_t;
  c.ln_z_min = ln_z_min;
  c.d_ln_z = d_ln_z;
  c.inv_d_ln_z = 1.0 / d_ln_z;
  c.t_min = std::exp(ln_t_min);
  c.t_max = std::exp(ln_t_min + d_ln_t * (n_t - 1));
  c.z_min = std::exp(ln_z_min);
  c.z_max = std::exp(ln_z_min + d_ln_z * (n_z - 1));
}

int gal_ode_batch_set_cooling_table(gal_ode_batch* h, const double* ln_lambda, int n_t, double ln_t_min,
                                    double d_ln_t, int n_z, double ln_z_min, double d_ln_z) {
  if (!h || !cooling_table_valid(ln_lambda, n_t, d_ln_t, n_z, d_ln_z)) return GAL_E_INVALID;
  cooling_desc(h->cd, n_t, ln_t_min, d_ln_t, n_z, ln_z_min, d_ln_z);
  HIP_TRY(hipMemcpy(h->d_table, ln_lambda, sizeof(double) * n_t * n_z, hipMemcpyHostToDevice));
  h->have_table = true;
  return GAL_OK;
}

int gal_ode_batch_set_launch(gal_ode_batch* h, int nodes_per_block, int blocks_per_cu) {
  if (!h) return GAL_E_INVALID;
  if (nodes_per_block != 64) return GAL_E_INVALID;
  // one workgroup per CU keeps every lane vector in LDS; two keep state and
  // scales in the HBM workspace; three or four keep the stage derivatives
  // there (the LDS of a CU is 160 KiB)
  if (blocks_per_cu < 1 || blocks_per_cu > 4) return GAL_E_INVALID;
  h->nodes_per_block = nodes_per_block;
  h->blocks_per_cu = blocks_per_cu;
  return GAL_OK;
}

int gal_ode_batch_set_wall_limit(gal_ode_batch* h, double seconds) {
  if (!h || !(seconds == seconds)) return GAL_E_INVALID;
  if (seconds > 0.0 && !(h->wall_clock_hz > 0.0)) return GAL_E_HIP;
  h->wall_limit_s = seconds > 0.0 ? seconds : 0.0;
  return GAL_OK;
}

int gal_ode_batch_set_graph_stall_limit(gal_ode_batch* h, double seconds) {
  if (!h || !(seconds == seconds)) return GAL_E_INVALID;
  h->graph_stall_s = seconds > 0.0 ? seconds : 60.0;
  return GAL_OK;
}

int gal_ode_batch_set_rhs_waves(gal_ode_batch* h, int waves) {
  if (!h || (waves != 1 && waves != 2)) return GAL_E_INVALID;
  h->rhs_waves = waves;
  return GAL_OK;
}

int gal_ode_batch_get_launch(gal_ode_batch* h, int* nodes_per_block, int* blocks_per_cu, int* rhs_waves) {
  if (!h) return GAL_E_INVALID;
  if (nodes_per_block) *nodes_per_block = h->nodes_per_block;
  if (blocks_per_cu) *blocks_per_cu = h->blocks_per_cu;
  if (rhs_waves) *rhs_waves = h->rhs_waves;
  return GAL_OK;
}

extern "C++" {
template <bool GRAPH = false>
static int launch_evolve(gal_ode_batch* h, const EvolveArgs& args, hipStream_t s) {
  const int blocks = std::max(1, h->n_cus * h->blocks_per_cu);
  if (h->rhs_waves == 2 && h->blocks_per_cu > 2) return GAL_E_INVALID;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, args.work);
  if (h->rhs_waves == 2) {
    // two waves per group of 64 nodes: one group per CU with the stage
    // derivatives in LDS, or two with them in the HBM workspace
    if (h->blocks_per_cu == 1) {
      constexpr size_t lds = lds_bytes<64, true, true, 2>();
      hipLaunchKernelGGL((evolve_kernel<64, true, true, 2, GRAPH>), dim3(blocks), dim3(128), lds, s, args);
    } else {
      EvolveArgs b = args;
      b.kwork = h->d_kwork;
      if (blocks * 64 > h->kwork_lanes) return GAL_E_INVALID;
      constexpr size_t lds = lds_bytes<64, true, false, 2>();
      hipLaunchKernelGGL((evolve_kernel<64, true, false, 2, GRAPH>), dim3(blocks), dim3(128), lds, s, b);
    }
    HIP_TRY(hipGetLastError());
    return GAL_OK;
  }
  constexpr size_t lds_all = lds_bytes<64, true, true>(), lds_k = lds_bytes<64, false, true>(),
                   lds_y = lds_bytes<64, true, false>();
  if (h->blocks_per_cu == 1) {
    hipLaunchKernelGGL((evolve_kernel<64, true, true, 1, GRAPH>), dim3(blocks), dim3(64), lds_all, s, args);
  } else if (h->blocks_per_cu == 2) {
    EvolveArgs b = args;
    b.ywork = h->d_ywork;
    b.ystride = blocks * 64;
    if (b.ystride > h->ywork_lanes) return GAL_E_INVALID;
    hipLaunchKernelGGL((evolve_kernel<64, false, true, 1, GRAPH>), dim3(blocks), dim3(64), lds_k, s, b);
  } else {
    EvolveArgs b = args;
    b.kwork = h->d_kwork;
    if (blocks * 64 > h->kwork_lanes) return GAL_E_INVALID;
    hipLaunchKernelGGL((evolve_kernel<64, true, false, 1, GRAPH>), dim3(blocks), dim3(64), lds_y, s, b);
  }
  HIP_TRY(hipGetLastError());
  return GAL_OK;
}
}  // extern "C++"

static int make_args(gal_ode_batch* h, int n, int ld, double* y, const double* node_par, const double* t_start,
                     const double* t_end, double* h_step, double* scale, int* status, int64_t* counters,
                     EvolveArgs& a) {
  if (!h || n < 0 || n > h->n_max || ld < n) return GAL_E_INVALID;
  if (!h->have_physics || !h->have_table) return GAL_E_INVALID;
  if (n > 0 && (!y || !node_par || !t_start || !t_end || !h_step || !status)) return GAL_E_INVALID;
  if ((h->flags & GAL_FLAG_SCALES_PROVIDED) && !scale) return GAL_E_INVALID;
  a.n = n;
  a.n_claim = n;
  a.ld = ld;
  a.y = y;
  a.np = node_par;
  a.t_start = t_start;
  a.t_end = t_end;
  a.h_step = h_step;
  a.scale = scale;
  a.status = status;
  a.counters = reinterpret_cast<long long*>(counters);
  a.order = nullptr;
  a.n_claim_dev = nullptr;
  a.work = h->d_work;
  a.ywork = nullptr;
  a.ystride = 0;
  a.kwork = nullptr;
  a.P = h->d_params;
  a.tab = h->d_table;
  a.D = h->D;
  a.cd = h->cd;
  a.eps_abs = h->tol_abs;
  a.eps_rel = h->tol_rel;
  a.flags = h->flags;
  a.mode = (int)h->params[GAL_P_PHYSICS_MODE];
  a.wall_ticks = 0;
  if (h->wall_limit_s > 0.0)
    a.wall_ticks = (unsigned long long)std::max(1.0, std::min(h->wall_limit_s * h->wall_clock_hz, 1.0e18));
  return GAL_OK;
}

int gal_ode_batch_evolve_d(gal_ode_batch* h, int n, int ld, double* y, const double* node_par,
                           const double* t_start, const double* t_end, double* h_step, double* scale, int* status,
                           int64_t* counters, void* stream) {
  return gal_ode_evolve_subset(h, n, ld, y, node_par, t_start, t_end, h_step, scale, status, counters, nullptr, n,
                               (hipStream_t)stream);
}

int gal_ode_batch_evolve_h(gal_ode_batch* h, int n, double* y, const double* node_par, const double* t_start,
                           const double* t_end, double* h_step, double* scale, int* status, int64_t* counters) {
  if (!h || n < 0 || n > h->n_max) return GAL_E_INVALID;
  if (n == 0) return GAL_OK;
  if (!y || !node_par || !t_start || !t_end || !h_step || !status) return GAL_E_INVALID;
  if (!h->d_y) {
    const size_t nm = (size_t)h->n_max;
    if (hipMalloc(&h->d_y, sizeof(double) * GAL_NVAR * nm) != hipSuccess ||
        hipMalloc(&h->d_np, sizeof(double) * GAL_NPARAM_NODE * nm) != hipSuccess ||
        hipMalloc(&h->d_t0, sizeof(double) * nm) != hipSuccess ||
        hipMalloc(&h->d_t1, sizeof(double) * nm) != hipSuccess ||
        hipMalloc(&h->d_h, sizeof(double) * nm) != hipSuccess ||
        hipMalloc(&h->d_scale, sizeof(double) * GAL_NVAR * nm) != hipSuccess ||
        hipMalloc(&h->d_status, sizeof(int) * nm) != hipSuccess ||
        hipMalloc(&h->d_counters, sizeof(long long) * 3 * nm) != hipSuccess) {
      free_staging(h);
      return GAL_E_NOMEM;
    }
  }
  const size_t nn = (size_t)n;
  HIP_TRY(hipMemcpy(h->d_y, y, sizeof(double) * GAL_NVAR * nn, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_np, node_par, sizeof(double) * GAL_NPARAM_NODE * nn, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_t0, t_start, sizeof(double) * nn, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_t1, t_end, sizeof(double) * nn, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_h, h_step, sizeof(double) * nn, hipMemcpyHostToDevice));
  if ((h->flags & GAL_FLAG_SCALES_PROVIDED) && scale)
    HIP_TRY(hipMemcpy(h->d_scale, scale, sizeof(double) * GAL_NVAR * nn, hipMemcpyHostToDevice));
  int rc = gal_ode_batch_evolve_d(h, n, n, h->d_y, h->d_np, h->d_t0, h->d_t1, h->d_h, h->d_scale, h->d_status,
                                  counters ? reinterpret_cast<int64_t*>(h->d_counters) : nullptr, nullptr);
  if (rc != GAL_OK) return rc;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(y, h->d_y, sizeof(double) * GAL_NVAR * nn, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(h_step, h->d_h, sizeof(double) * nn, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(status, h->d_status, sizeof(int) * nn, hipMemcpyDeviceToHost));
  if (scale) HIP_TRY(hipMemcpy(scale, h->d_scale, sizeof(double) * GAL_NVAR * nn, hipMemcpyDeviceToHost));
  if (counters) HIP_TRY(hipMemcpy(counters, h->d_counters, sizeof(long long) * 3 * nn, hipMemcpyDeviceToHost));
  return GAL_OK;
}

int gal_ode_rhs_d(gal_ode_batch* h, int n, int ld, const double* t, const double* y, const double* node_par,
                  double* dydt, void* stream) {
  if (!h || n < 0 || ld < n || !h->have_physics || !h->have_table) return GAL_E_INVALID;
  if (n == 0) return GAL_OK;
  if (!t || !y || !node_par || !dydt) return GAL_E_INVALID;
  const int threads = 256;
  hipLaunchKernelGGL(rhs_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, (hipStream_t)stream, n, ld, t,
                     y, node_par, dydt, (const double*)h->d_params, (const double*)h->d_table, h->D, h->cd,
                     (int)h->params[GAL_P_PHYSICS_MODE]);
  HIP_TRY(hipGetLastError());
  return GAL_OK;
}

int gal_ode_hadjust_d(int n, int dim, int ord, double eps_abs, double eps_rel, double a_y, double a_dydt,
                      const double* y, const double* yerr, const double* yp, const double* scale,
                      const int* is_non_negative, double* h, int* result, void* stream) {
  if (n < 0 || dim < 1 || ord < 1) return GAL_E_INVALID;
  if (n == 0) return GAL_OK;
  if (!y || !yerr || !yp || !scale || !is_non_negative || !h || !result) return GAL_E_INVALID;
  const int threads = 256;
  hipLaunchKernelGGL(hadjust_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, (hipStream_t)stream, n, dim,
                     ord, eps_abs, eps_rel, a_y, a_dydt, y, yerr, yp, scale, is_non_negative, h, result);
  HIP_TRY(hipGetLastError());
  return GAL_OK;
}

int gal_ode_batch_autotune(gal_ode_batch* h, int n, int ld, double* y, const double* node_par,
                           const double* t_start, const double* t_end, double* h_step, double* scale, int* status,
                           void* stream, int* best_nodes_per_block, int* best_blocks_per_cu) {
  // Times each admissible launch shape on copies of the inputs (the caller's
  // arrays are untouched), keeps the fastest and leaves it set on the handle.
  if (!h || n <= 0 || ld < n) return GAL_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  double *wy = nullptr, *wh = nullptr, *wsc = nullptr;
  int* wst = nullptr;
  const size_t nn = (size_t)ld;
  if (hipMalloc(&wy, sizeof(double) * GAL_NVAR * nn) != hipSuccess ||
      hipMalloc(&wh, sizeof(double) * nn) != hipSuccess ||
      hipMalloc(&wsc, sizeof(double) * GAL_NVAR * nn) != hipSuccess ||
      hipMalloc(&wst, sizeof(int) * nn) != hipSuccess) {
    hipFree(wy); hipFree(wh); hipFree(wsc); hipFree(wst);
    return GAL_E_NOMEM;
  }
  // (nodes per workgroup, workgroups per CU, waves per group of nodes)
  const int shapes[][3] = {{64, 1, 1}, {64, 2, 1}, {64, 4, 1}, {64, 1, 2}, {64, 2, 2}};
  float best = 1e30f;
  int bn = h->nodes_per_block, bb = h->blocks_per_cu, bw = h->rhs_waves;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int rc = GAL_OK;
  for (auto& sh : shapes) {
    if (gal_ode_batch_set_launch(h, sh[0], sh[1]) != GAL_OK || gal_ode_batch_set_rhs_waves(h, sh[2]) != GAL_OK)
      continue;
    for (int rep = 0; rep < 2 && rc == GAL_OK; ++rep) {
      hipMemcpyAsync(wy, y, sizeof(double) * GAL_NVAR * nn, hipMemcpyDeviceToDevice, s);
      hipMemcpyAsync(wh, h_step, sizeof(double) * nn, hipMemcpyDeviceToDevice, s);
      if (scale && (h->flags & GAL_FLAG_SCALES_PROVIDED))
        hipMemcpyAsync(wsc, scale, sizeof(double) * GAL_NVAR * nn, hipMemcpyDeviceToDevice, s);
      hipEventRecord(e0, s);
      rc = gal_ode_batch_evolve_d(h, n, ld, wy, node_par, t_start, t_end, wh, wsc, wst, nullptr, stream);
      hipEventRecord(e1, s);
      hipEventSynchronize(e1);
      float ms = 0.0f;
      hipEventElapsedTime(&ms, e0, e1);
      if (rep == 1 && ms < best) {
        best = ms;
        bn = sh[0];
        bb = sh[1];
        bw = sh[2];
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(wy); hipFree(wh); hipFree(wsc); hipFree(wst);
  (void)status;
  gal_ode_batch_set_launch(h, bn, bb);
  gal_ode_batch_set_rhs_waves(h, bw);
  if (best_nodes_per_block) *best_nodes_per_block = bn;
  if (best_blocks_per_cu) *best_blocks_per_cu = bb;
  return rc;
}

}  // extern "C"

// ============================================================================
// Allocation-free evolve
// ============================================================================
// The handle path keeps physics parameters, the cooling table and the claim
// counter in device buffers that gal_ode_batch_create allocates.  Callers
// that cannot allocate (code under graph capture, serving loops, harnesses
// that forbid hipMalloc) use gal_ode_evolve_static_d instead: its scratch is
// this file-scope array, one copy per GPU, held for the life of the process:
//   [0, 8)                         claim counter
//   [8, 8 + 8*GAL_NPARAM)          physics parameters
//   then                           cooling table, up to NT_MAX x NZ_MAX doubles
// Every call on a GPU reuses that GPU's copy; calls that overlap on one GPU
// race on it.
namespace {
constexpr size_t kStaticTableDoubles = (size_t)GAL_COOLING_NT_MAX * GAL_COOLING_NZ_MAX;
constexpr size_t kStaticScratchDoubles = 1 + GAL_NPARAM + kStaticTableDoubles;
constexpr int kMaxDevices = 64;
std::mutex g_static_mutex;
void* g_static_scratch[kMaxDevices] = {};
int g_static_cus[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kStaticScratchDoubles];

// Once per device, with that device current: the device address of this
// device's copy of the scratch array (the array's name is a host-side
// stand-in, not a device address), the CU count and the LDS attribute of the
// kernel the entry point launches.
static int static_scratch(int device, void** ptr, int* n_cus) {
  if (device < 0 || device >= kMaxDevices) return GAL_E_INVALID;
  std::lock_guard<std::mutex> lock(g_static_mutex);
  if (!g_static_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return GAL_E_HIP;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
      return GAL_E_HIP;
    if (hipFuncSetAttribute((const void*)evolve_kernel<64, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes<64, true, true>()) != hipSuccess) {
      (void)hipGetLastError();
      return GAL_E_HIP;
    }
    g_static_cus[device] = cus;
    g_static_scratch[device] = p;
  }
  *ptr = g_static_scratch[device];
  *n_cus = g_static_cus[device];
  return GAL_OK;
}

extern "C" int gal_ode_evolve_static_d(const double* params, int n_params, const double* ln_lambda, int n_t,
                                       double ln_t_min, double d_ln_t, int n_z, double ln_z_min, double d_ln_z,
                                       double tol_abs, double tol_rel, int flags, int n, int ld, double* y,
                                       const double* node_par, const double* t_start, const double* t_end,
                                       double* h_step, double* scale, int* status, int64_t* counters, void* stream) {
  if (!params || n_params != GAL_NPARAM || !physics_valid(params)) return GAL_E_INVALID;
  if (!cooling_table_valid(ln_lambda, n_t, d_ln_t, n_z, d_ln_z)) {
    if (ln_lambda && ((size_t)n_t * (size_t)n_z > kStaticTableDoubles))
      fprintf(stderr, "gal_ode_evolve_static_d: cooling table %d x %d exceeds the %d x %d of the static scratch\n",
              n_t, n_z, GAL_COOLING_NT_MAX, GAL_COOLING_NZ_MAX);
    return GAL_E_INVALID;
  }
  if (tol_abs < 0.0 || tol_rel < 0.0 || (tol_abs <= 0.0 && tol_rel <= 0.0)) return GAL_E_INVALID;
  if (n < 0 || ld < n) return GAL_E_INVALID;
  if (n == 0) return GAL_OK;
  if (!y || !node_par || !t_start || !t_end || !h_step || !status) return GAL_E_INVALID;
  if ((flags & GAL_FLAG_SCALES_PROVIDED) && !scale) return GAL_E_INVALID;
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  void* base = nullptr;
  int n_cus = 0;
  int rc = static_scratch(device, &base, &n_cus);
  if (rc != GAL_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned int* work = reinterpret_cast<unsigned int*>(base);
  double* d_params = reinterpret_cast<double*>(base) + 1;
  double* d_table = d_params + GAL_NPARAM;
  HIP_TRY(hipMemcpyAsync(d_params, params, sizeof(double) * GAL_NPARAM, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d_table, ln_lambda, sizeof(double) * (size_t)n_t * n_z, hipMemcpyHostToDevice, s));
  EvolveArgs a{};
  a.n = n;
  a.n_claim = n;
  a.ld = ld;
  a.y = y;
  a.np = node_par;
  a.t_start = t_start;
  a.t_end = t_end;
  a.h_step = h_step;
  a.scale = scale;
  a.status = status;
  a.counters = reinterpret_cast<long long*>(counters);
  a.order = nullptr;             // claims in node order
  a.n_claim_dev = nullptr;
  a.work = work;
  a.P = d_params;
  a.tab = d_table;
  derive_constants(params, a.D);
  cooling_desc(a.cd, n_t, ln_t_min, d_ln_t, n_z, ln_z_min, d_ln_z);
  a.eps_abs = tol_abs;
  a.eps_rel = tol_rel;
  a.flags = flags;
  a.mode = (int)params[GAL_P_PHYSICS_MODE];
  a.wall_ticks = 0;
  // every lane vector in LDS, one workgroup per CU: no workspace beyond the scratch
  constexpr size_t lds_all = lds_bytes<64, true, true>();
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, work);
  hipLaunchKernelGGL((evolve_kernel<64, true, true>), dim3(n_cus), dim3(64), lds_all, s, a);
  HIP_TRY(hipGetLastError());
  return GAL_OK;
}

// Claim order of an evolve: every node's cost estimate, sorted so that the
// costliest nodes are claimed first (their chains then overlap the bulk).
// Writes h->d_order on stream s.
static int plan_claim_order(gal_ode_batch* h, int n_subset, const int* subset, int ld, const double* y,
                            const double* node_par, const double* t_start, const double* t_end, const double* scale,
                            const int* n_subset_dev, hipStream_t s) {
  const int threads = 256;
  hipLaunchKernelGGL(cost_probe_kernel, dim3((n_subset + threads - 1) / threads), dim3(threads), 0, s, n_subset,
                     subset, ld, y, node_par, t_start, t_end, scale, (const double*)h->d_params,
                     (const double*)h->d_table, h->D, h->cd, (int)h->params[GAL_P_PHYSICS_MODE],
                     (h->flags & GAL_FLAG_SCALES_PROVIDED) ? 1 : 0, n_subset_dev, h->d_key, h->d_index);
  HIP_TRY(hipGetLastError());
  size_t tb = h->sort_temp_bytes;
  HIP_TRY(hipcub::DeviceRadixSort::SortPairsDescending(h->d_sort_temp, tb, h->d_key, h->d_key_sorted, h->d_index,
                                                       h->d_order, n_subset, 0, 32, s));
  return GAL_OK;
}

int gal_ode_evolve_subset(gal_ode_batch* h, int n, int ld, double* y, const double* node_par, const double* t_start,
                          const double* t_end, double* h_step, double* scale, int* status, int64_t* counters,
                          const int* subset, int n_subset, hipStream_t s, const int* n_subset_dev) {
  EvolveArgs a;
  int rc = make_args(h, n, ld, y, node_par, t_start, t_end, h_step, scale, status, counters, a);
  if (rc != GAL_OK) return rc;
  if (n_subset < 0 || n_subset > n || (subset == nullptr && n_subset != n)) return GAL_E_INVALID;
  if (n_subset == 0) return GAL_OK;
  a.n_claim = n_subset;
  a.n_claim_dev = n_subset_dev;   // when set, n_subset is the list's capacity
  a.order = subset;
  h->plan_y = nullptr;            // an evolve of its own replaces any pending plan
  if (h->flags & GAL_FLAG_COST_ORDER) {
    rc = plan_claim_order(h, n_subset, subset, ld, y, node_par, t_start, t_end, scale, n_subset_dev, s);
    if (rc != GAL_OK) return rc;
    a.order = h->d_order;
  }
  return launch_evolve(h, a, s);
}

int gal_ode_batch_plan_d(gal_ode_batch* h, int n, int ld, const double* y, const double* node_par,
                         const double* t_start, const double* t_end, const double* scale, void* stream) {
  if (!h || n < 0 || n > h->n_max || ld < n) return GAL_E_INVALID;
  if (!h->have_physics || !h->have_table || !(h->flags & GAL_FLAG_COST_ORDER)) return GAL_E_INVALID;
  if (n > 0 && (!y || !node_par || !t_start || !t_end || ((h->flags & GAL_FLAG_SCALES_PROVIDED) && !scale)))
    return GAL_E_INVALID;
  h->plan_y = nullptr;
  if (n > 0) {
    const int rc = plan_claim_order(h, n, nullptr, ld, y, node_par, t_start, t_end, scale, nullptr, (hipStream_t)stream);
    if (rc != GAL_OK) return rc;
  }
  h->plan_y = y;
  h->plan_n = n;
  h->plan_ld = ld;
  return GAL_OK;
}

int gal_ode_batch_evolve_planned_d(gal_ode_batch* h, int n, int ld, double* y, const double* node_par,
                                   const double* t_start, const double* t_end, double* h_step, double* scale,
                                   int* status, int64_t* counters, void* stream) {
  if (!h || !y || h->plan_y != y || h->plan_n != n || h->plan_ld != ld) return GAL_E_INVALID;
  h->plan_y = nullptr;   // a plan serves one evolve
  EvolveArgs a;
  const int rc = make_args(h, n, ld, y, node_par, t_start, t_end, h_step, scale, status, counters, a);
  if (rc != GAL_OK) return rc;
  if (n == 0) return GAL_OK;
  a.n_claim = n;
  a.n_claim_dev = nullptr;
  a.order = h->d_order;
  return launch_evolve(h, a, (hipStream_t)stream);
}

int gal_ode_evolve_graph(gal_ode_batch* h, int ld, double* y, double* h_step, double* scale, int* status,
                         const gal::OpGraph& g, int* graph_error, hipStream_t s) {
  if (!h || ld < 1 || ld > h->n_max || g.n < 0 || !graph_error) return GAL_E_INVALID;
  if (g.n == 0) return GAL_OK;
  if (!g.ops || !g.next_a || !g.next_b || !g.npred || !g.ready || !g.tail) return GAL_E_INVALID;
  EvolveArgs a;
  // the per-node arrays of the node mode are unused: intervals and halo
  // parameters come with each operation
  const double* unused = &g.ops->t0;
  int rc = make_args(h, ld, ld, y, unused, unused, unused, h_step, scale, status, nullptr, a);
  if (rc != GAL_OK) return rc;
  a.g = g;
  a.graph_error = graph_error;
  // a claimed entry may wait for the whole remaining critical path, but
  // not for the handle's stall limit (a minute by default) in which no
  // operation anywhere becomes ready
  const double hz = h->wall_clock_hz > 0.0 ? h->wall_clock_hz : 1.0e8;
  a.stall_ticks = (unsigned long long)std::min(std::max(h->graph_stall_s * hz, 1.0), 1.0e18);
  return launch_evolve<true>(h, a, s);
}
