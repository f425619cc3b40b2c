// surrogate.hip -- the learned node-interval surrogate on MI355X (gfx950):
// a bf16 MLP on the MFMA units (forward, backward, Adam), the node encode
// and gate kernels, and the surrogate evolve that routes rejected nodes to
// the FP64 integrator.  C ABI of include/gal_surrogate.h.
//
// Matrix products.  One kernel template computes C = A B for the three
// shapes of training: activations x weights (forward), gradients x W^T
// (backward through a layer) and activations^T x gradients (weight
// gradients, split over the batch so that every split writes its own fp32
// partial sum; the Adam kernel adds the partials in a fixed order, so
// training is deterministic).  Each 256-thread workgroup owns a 64 x 64 tile
// of C, each of its four waves a 32 x 32 quarter, built from 2 x 2
// v_mfma_f32_16x16x32_bf16 tiles with fp32 accumulators.  Operand tiles of
// 64 x 32 are staged in LDS as [row][k] with a padded row stride, so that
// each lane reads its eight consecutive k values with one 16-byte load.
// Epilogues fuse the bias, SiLU and the bf16 rounding of the activations
// (forward) and the SiLU derivative (backward).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>

#include "../../include/gal_ode.h"
#include "../../include/gal_surrogate.h"
#include "constants.hpp"
#include "physics.hpp"
#include "batch.hpp"

using namespace gal;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = GAL_MLP_INPUTS, kHid = GAL_MLP_HIDDEN, kOut = GAL_MLP_OUTPUTS, kOutLd = GAL_MLP_OUT_LD;
constexpr int kLayers = 4;
constexpr int kFanIn[kLayers] = {kIn, kHid, kHid, kHid};
constexpr int kFanOut[kLayers] = {kHid, kHid, kHid, kOut};
constexpr int kLdW[kLayers] = {kHid, kHid, kHid, kOutLd};  // bf16 weight row stride (fan_out padded to 8)

constexpr int64_t param_count() {
  int64_t p = 0;
  for (int l = 0; l < kLayers; ++l) p += (int64_t)kFanIn[l] * kFanOut[l] + kFanOut[l];
  return p;
}
constexpr int64_t kParams = param_count();

// ---- bf16 ------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even; NaN stays quiet NaN
  const uint32_t u = __float_as_uint(f);
  const uint32_t rounded = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  const uint32_t quiet = (u >> 16) | 0x40u;
  return (uint16_t)(((u & 0x7fffffffu) > 0x7f800000u) ? quiet : rounded);   // a select, not a branch
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ---- MFMA GEMM ---------------------------------------------------------------
constexpr int TM = 64, TN = 64, TK = 32, LDT = TK + 8;
enum Epi { EPI_SILU = 0, EPI_LINEAR = 1, EPI_DSILU = 2, EPI_PARTIAL = 3 };

struct GemmArgs {
  const uint16_t* A;  // A(m,k) = AT ? A[k*lda + m] : A[m*lda + k]
  const uint16_t* B;  // B(k,n) = BT ? B[n*ldb + k] : B[k*ldb + n]
  int M, N, K, lda, ldb, ldc;
  int n_valid;        // epilogue writes columns n < n_valid
  int k_split;        // K rows per split (multiple of TK); grid.z splits
  const float* bias;  // SILU, LINEAR
  uint16_t* out_bf;   // SILU: activation; DSILU: gradient wrt pre-activation
  uint16_t* z_bf;     // SILU: pre-activation (may be null)
  const uint16_t* z_in;  // DSILU: pre-activation of this layer
  float* out_f;       // LINEAR: output; PARTIAL: [split][M][ldc]
};

template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[TM * LDT];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[TN * LDT];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int kbeg = blockIdx.z * g.k_split;
  const int kend = min(g.K, kbeg + g.k_split);
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    if (!AT) {  // rows of A are contiguous in k
      const int r = tid >> 2, c = (tid & 3) * 8;
      uint4 v = zero;
      if (m0 + r < g.M && k0 + c < kend) v = *reinterpret_cast<const uint4*>(g.A + (size_t)(m0 + r) * g.lda + k0 + c);
      *reinterpret_cast<uint4*>(As + r * LDT + c) = v;
    } else {    // A^T: contiguous in m, transposed into LDS
      const int kk = tid >> 3, c = (tid & 7) * 8;
      uint4 v = zero;
      if (k0 + kk < kend && m0 + c < g.M) v = *reinterpret_cast<const uint4*>(g.A + (size_t)(k0 + kk) * g.lda + m0 + c);
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
      for (int j = 0; j < 8; ++j) As[(c + j) * LDT + kk] = e[j];
    }
    if (BT) {   // B^T rows are contiguous in k
      const int r = tid >> 2, c = (tid & 3) * 8;
      uint4 v = zero;
      if (n0 + r < g.N && k0 + c < kend) v = *reinterpret_cast<const uint4*>(g.B + (size_t)(n0 + r) * g.ldb + k0 + c);
      *reinterpret_cast<uint4*>(Bs + r * LDT + c) = v;
    } else {    // B: contiguous in n, transposed into LDS
      const int kk = tid >> 3, c = (tid & 7) * 8;
      uint4 v = zero;
      if (k0 + kk < kend && n0 + c < g.N) v = *reinterpret_cast<const uint4*>(g.B + (size_t)(k0 + kk) * g.ldb + n0 + c);
      const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[(c + j) * LDT + kk] = e[j];
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint4 v = *reinterpret_cast<const uint4*>(As + (wm * 32 + i * 16 + (lane & 15)) * LDT + (lane >> 4) * 8);
      a[i] = __builtin_bit_cast(bf16x8, v);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint4 v = *reinterpret_cast<const uint4*>(Bs + (wn * 32 + j * 16 + (lane & 15)) * LDT + (lane >> 4) * 8);
      b[j] = __builtin_bit_cast(bf16x8, v);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // C(m, n) of accumulator register r: m = 4*(lane/16) + r, n = lane % 16 within the 16 x 16 tile.
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.n_valid) continue;
        const float c = acc[i][j][r];
        const size_t o = (size_t)m * g.ldc + n;
        if (EPI == EPI_SILU) {
          const float z = c + g.bias[n];
          if (g.z_bf) g.z_bf[o] = f2bf(z);
          g.out_bf[o] = f2bf(silu(z));
        } else if (EPI == EPI_LINEAR) {
          g.out_f[o] = c + g.bias[n];
        } else if (EPI == EPI_DSILU) {
          g.out_bf[o] = f2bf(c * dsilu(bf2f(g.z_in[o])));
        } else {
          g.out_f[((size_t)blockIdx.z * g.M + m) * g.ldc + n] = c;
        }
      }
}

// ---- fused inference forward ---------------------------------------------
// The whole network for 128 rows per workgroup in one kernel: the rows'
// activations stay in LDS from layer to layer (76 KiB, so two workgroups
// share a CU).  Wave w computes all 128 rows x columns [64w, 64w+64) of a
// hidden layer (8 x 4 tiles of 16 x 16); for the 32-column output layer waves
// 0 and 1 take one column tile each.  A wave's columns are its own, so it
// reads its weight fragments straight from the transposed bf16 weights
// ([n][k], k contiguous) into registers, four k chunks at a time, with no LDS
// staging and no workgroup barrier inside a layer.  The arithmetic is the
// per-layer GEMMs' (same MFMA, same k order, same bias + SiLU epilogue and
// bf16 rounding), so the outputs are bit-identical to them.
constexpr int FR = 128;                 // rows per workgroup
constexpr int LDH = kHid + 8;           // activation row stride (bf16)
constexpr int LDX = kIn + 8;            // input row stride (bf16)
constexpr size_t kFusedLds = sizeof(uint16_t) * ((size_t)FR * LDH + (size_t)FR * LDX);

struct FusedArgs {
  const uint16_t* x;                 // [n][kIn] bf16
  const uint16_t* wt[kLayers];       // [kLdW[l]][kFanIn[l]] bf16
  const float* bias[kLayers];
  float* out;                        // [n][kOutLd]
  int n;
};

template <int K, bool LAST>
__device__ __forceinline__ void fused_layer(const FusedArgs& a, int l, const uint16_t* __restrict__ A, int lda,
                                            uint16_t* __restrict__ H, int r0) {
  constexpr int NK = K / TK;
  constexpr int KG = NK < 4 ? NK : 4;                    // k chunks per register group
  constexpr int JT = LAST ? 1 : 4;                       // column tiles per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool active = LAST ? wave < 2 : true;
  const int nbase = LAST ? wave * 16 : wave * 64;
  const uint16_t* __restrict__ wt = a.wt[l];
  f32x4 acc[8][JT];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < JT; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  if (active) {
    for (int g0 = 0; g0 < NK; g0 += KG) {
      uint4 bq[KG][JT];
#pragma unroll
      for (int kk = 0; kk < KG; ++kk)
#pragma unroll
        for (int j = 0; j < JT; ++j)
          bq[kk][j] = *reinterpret_cast<const uint4*>(wt + (size_t)(nbase + j * 16 + (lane & 15)) * K +
                                                      (g0 + kk) * TK + (lane >> 4) * 8);
#pragma unroll
      for (int kk = 0; kk < KG; ++kk) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const uint4 v = *reinterpret_cast<const uint4*>(A + (i * 16 + (lane & 15)) * lda + (g0 + kk) * TK +
                                                          (lane >> 4) * 8);
          const bf16x8 av = __builtin_bit_cast(bf16x8, v);
#pragma unroll
          for (int j = 0; j < JT; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, __builtin_bit_cast(bf16x8, bq[kk][j]), acc[i][j],
                                                                0, 0, 0);
        }
      }
    }
  }
  // every wave has read this layer's input before any epilogue overwrites H
  __syncthreads();
  if (!active) return;
  const float* __restrict__ bias = a.bias[l];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < JT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i * 16 + (lane >> 4) * 4 + r;
        const int col = nbase + j * 16 + (lane & 15);
        const float z = acc[i][j][r] + bias[LAST ? (col < kOut ? col : 0) : col];
        if (LAST) {
          if (col < kOut && r0 + row < a.n) a.out[(size_t)(r0 + row) * kOutLd + col] = z;
        } else {
          H[row * LDH + col] = f2bf(silu(z));
        }
      }
}

__global__ __launch_bounds__(256, 2) void mlp_forward_fused(FusedArgs a) {
  extern __shared__ __attribute__((aligned(16))) uint16_t fused_lds[];
  uint16_t* H = fused_lds;
  uint16_t* X = H + FR * LDH;
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * FR;
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  for (int e = tid; e < FR * (kIn / 8); e += 256) {
    const int r = e / (kIn / 8), c = (e % (kIn / 8)) * 8;
    uint4 v = zero;
    if (r0 + r < a.n) v = *reinterpret_cast<const uint4*>(a.x + (size_t)(r0 + r) * kIn + c);
    *reinterpret_cast<uint4*>(X + r * LDX + c) = v;
  }
  __syncthreads();
  fused_layer<kIn, false>(a, 0, X, LDX, H, r0);
  __syncthreads();
  fused_layer<kHid, false>(a, 1, H, LDH, H, r0);
  __syncthreads();
  fused_layer<kHid, false>(a, 2, H, LDH, H, r0);
  __syncthreads();
  fused_layer<kHid, true>(a, 3, H, LDH, H, r0);
}

// Transposed bf16 copy of one layer's weights: wt[n][k] = wbf[k][n].
__global__ __launch_bounds__(256) void transpose_weights_kernel(const uint16_t* __restrict__ wbf, int K, int N,
                                                                uint16_t* __restrict__ wt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= K * N) return;
  const int k = e / N, n = e % N;
  wt[(size_t)n * K + k] = wbf[e];
}

// ---- training support kernels ----------------------------------------------
// Network outputs enter sinh(); beyond +-20 they are clipped (sinh(20) = 2.4e8).
constexpr float kOutClip = 20.0f;

struct LossArgs {
  const float* out;
  const float* target;
  uint16_t* dout;
  float* loss_partial;
  int B;
  float w[32], sigma[32];
  unsigned head_mask;
  float head_w, inv_wsum, inv_B;
};

__global__ __launch_bounds__(256) void loss_kernel(LossArgs a) {
  __shared__ float red[256];
  const int r = blockIdx.x * 256 + threadIdx.x;
  float l = 0.0f;
  if (r < a.B) {
    const float* o = a.out + (size_t)r * kOutLd;
    const float* t = a.target + (size_t)r * kOutLd;
    uint16_t* d = a.dout + (size_t)r * kOutLd;
    float mse = 0.0f, emax = 0.0f;
    for (int v = 0; v < 30; ++v) {
      const float res = o[v] - t[v];
      mse += a.w[v] * res * res;
      d[v] = f2bf(2.0f * a.w[v] * res * a.inv_wsum * a.inv_B);
      // targets are asinh((r - m)/sigma) of the residual r: the error of the
      // residual itself is sigma |sinh(o) - sinh(t)|
      if ((a.head_mask >> v) & 1u)
        emax = fmaxf(emax, a.sigma[v] * fabsf(sinhf(fminf(fmaxf(o[v], -kOutClip), kOutClip)) - sinhf(t[v])));
    }
    // a sample whose target was clipped carries a flag in column 31: its true
    // error is at least 1 (asinh units), so the head learns to reject it
    if (t[31] != 0.0f) emax = fmaxf(emax, 1.0f);
    const float rh = o[GAL_MLP_HEAD] - logf(emax + 1.0e-7f);
    l = mse * a.inv_wsum + a.head_w * rh * rh;
    d[GAL_MLP_HEAD] = f2bf(2.0f * a.head_w * rh * a.inv_B);
    d[31] = 0;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.loss_partial[blockIdx.x] = red[0];
}

// Column sums of a bf16 gradient, one fp32 partial per batch split.  A
// workgroup covers 64 columns of one split: each wave reads whole 128-byte
// row segments (64 consecutive columns) of every fourth row, and the four
// waves' sums are added in LDS in a fixed order (deterministic).
__global__ __launch_bounds__(256) void colsum_kernel(const uint16_t* __restrict__ dz, int B, int N, int ld,
                                                     int rows_per_split, float* __restrict__ partial) {
  __shared__ float red[4][64];
  const int col = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + col;
  const int s = blockIdx.y;
  const int r0 = s * rows_per_split, r1 = min(B, r0 + rows_per_split);
  float acc = 0.0f;
  if (n < N) {
#pragma unroll 8
    for (int r = r0 + q; r < r1; r += 4) acc += bf2f(dz[(size_t)r * ld + n]);
  }
  red[q][col] = acc;
  __syncthreads();
  if (q == 0 && n < N) partial[(size_t)s * ld + n] = (red[0][col] + red[1][col]) + (red[2][col] + red[3][col]);
}

struct AdamArgs {
  float* param;
  float* m;
  float* v;
  uint16_t* wbf[kLayers];
  float* grad_out;
  const float* pw;
  const float* pb;
  int64_t w_off[kLayers], b_off[kLayers], pw_off[kLayers], pb_off[kLayers];
  int splits;
  float lr0, lr_min, beta1, beta2, eps;
  int64_t total_steps;
  const long long* step;
  int apply;
};

// Sum the split partials into the gradient (fixed order) and take one Adam
// step; keeps the bf16 copy of each weight in step with its fp32 master.
__global__ __launch_bounds__(256) void adam_kernel(AdamArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  int l = 0, k = 0, n = 0;
  bool is_b = false;
#pragma unroll
  for (int j = 0; j < kLayers; ++j) {
    if (p >= a.w_off[j] && p < a.b_off[j]) {
      const int64_t q = p - a.w_off[j];
      l = j; k = (int)(q / kFanOut[j]); n = (int)(q % kFanOut[j]);
    } else if (p >= a.b_off[j] && p < a.b_off[j] + kFanOut[j]) {
      l = j; n = (int)(p - a.b_off[j]); is_b = true;
    }
  }
  float g = 0.0f;
  if (is_b) {
    for (int s = 0; s < a.splits; ++s) g += a.pb[a.pb_off[l] + (int64_t)s * kLdW[l] + n];
  } else {
    for (int s = 0; s < a.splits; ++s) g += a.pw[a.pw_off[l] + ((int64_t)s * kFanIn[l] + k) * kLdW[l] + n];
  }
  if (a.grad_out) a.grad_out[p] = g;
  if (!a.apply) return;
  const double t = (double)(*a.step + 1);
  const double frac = fmin(t, (double)a.total_steps) / (double)a.total_steps;
  const float lr = (float)(a.lr_min + 0.5 * (a.lr0 - a.lr_min) * (1.0 + cos(kPi * frac)));
  const float bc1 = (float)(1.0 - pow((double)a.beta1, t));
  const float bc2 = (float)(1.0 - pow((double)a.beta2, t));
  const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * g;
  const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * g * g;
  a.m[p] = m;
  a.v[p] = v;
  const float w = a.param[p] - lr * (m / bc1) / (sqrtf(v / bc2) + a.eps);
  a.param[p] = w;
  if (!is_b) a.wbf[l][(size_t)k * kLdW[l] + n] = f2bf(w);
}

__global__ void weights_to_bf16_kernel(const float* __restrict__ param, int64_t w_off, int fan_in, int fan_out,
                                       int ldw, uint16_t* __restrict__ wbf) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)fan_in * fan_out) return;
  const int k = (int)(q / fan_out), n = (int)(q % fan_out);
  wbf[(size_t)k * ldw + n] = f2bf(param[w_off + q]);
}

__global__ void gather_kernel(const uint16_t* __restrict__ ds_x, const float* __restrict__ ds_t,
                              const int* __restrict__ perm, long long ds_n, const long long* __restrict__ cursor,
                              int B, uint16_t* __restrict__ x, float* __restrict__ t) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const long long c = (*cursor + r) % ds_n;
  const long long src = perm[c];
  const uint4* xs = reinterpret_cast<const uint4*>(ds_x + src * kIn);
  uint4* xd = reinterpret_cast<uint4*>(x + (size_t)r * kIn);
#pragma unroll
  for (int j = 0; j < kIn * 2 / 16; ++j) xd[j] = xs[j];
  const uint4* ts = reinterpret_cast<const uint4*>(ds_t + src * kOutLd);
  uint4* td = reinterpret_cast<uint4*>(t + (size_t)r * kOutLd);
#pragma unroll
  for (int j = 0; j < kOutLd * 4 / 16; ++j) td[j] = ts[j];
}

__global__ void reset_train_kernel(long long* cursor, int* local) {
  if (threadIdx.x == 0) {
    *cursor = 0;
    *local = 0;
  }
}

__global__ void advance_kernel(long long* step, long long* cursor, int* local, const float* __restrict__ loss_partial,
                               int nb, float inv_B, float* loss_hist, int hist_cap, int B) {
  if (threadIdx.x != 0) return;
  float s = 0.0f;
  for (int i = 0; i < nb; ++i) s += loss_partial[i];
  const int ls = *local;
  if (ls < hist_cap) loss_hist[ls] = s * inv_B;
  *local = ls + 1;
  *step += 1;
  *cursor += B;
}

__global__ void loss_sum_kernel(const float* __restrict__ loss_partial, int nb, float inv_B, float* loss) {
  if (threadIdx.x != 0) return;
  float s = 0.0f;
  for (int i = 0; i < nb; ++i) s += loss_partial[i];
  *loss = s * inv_B;
}

// ---- node encoding ------------------------------------------------------------
enum Role : int { R_MASS, R_METAL, R_ANGMOM, R_ENERGY, R_BH, R_UNIT, R_CONST, R_SPIN, R_FIRST_ORDER };

__device__ __forceinline__ int var_role(int v) {
  switch (v) {
    case GAL_V_HOT_MASS: case GAL_V_HOT_OUTFLOWED_MASS: case GAL_V_HOT_UNACCRETED_MASS:
    case GAL_V_DISK_MASS_STELLAR: case GAL_V_DISK_MASS_STELLAR_FORMED: case GAL_V_DISK_MASS_GAS:
    case GAL_V_SPHEROID_MASS_STELLAR: case GAL_V_SPHEROID_MASS_STELLAR_FORMED: case GAL_V_SPHEROID_MASS_GAS:
      return R_MASS;
    case GAL_V_HOT_ABUNDANCES: case GAL_V_HOT_OUTFLOWED_ABUNDANCES: case GAL_V_HOT_UNACCRETED_ABUNDANCES:
    case GAL_V_DISK_ABUNDANCES_STELLAR: case GAL_V_DISK_ABUNDANCES_GAS: case GAL_V_SPHEROID_ABUNDANCES_STELLAR:
    case GAL_V_SPHEROID_ABUNDANCES_GAS:
      return R_METAL;
    case GAL_V_HOT_ANGULAR_MOMENTUM: case GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM: case GAL_V_DISK_ANGULAR_MOMENTUM:
    case GAL_V_SPHEROID_ANGULAR_MOMENTUM:
      return R_ANGMOM;
    case GAL_V_HOT_ENERGY_RADIATED: return R_ENERGY;
    case GAL_V_BH_MASS: return R_BH;
    case GAL_V_BH_SPIN: return R_UNIT;
    case GAL_V_SPIN_ANGULAR_MOMENTUM: return R_SPIN;
    // not fed to the network: advanced by the first-order step of the rate at
    // the interval's start (outer radius, bound mass and stripped gas)
    case GAL_V_HOT_OUTER_RADIUS: case GAL_V_SATELLITE_BOUND_MASS: case GAL_V_HOT_STRIPPED_MASS:
    case GAL_V_HOT_STRIPPED_ABUNDANCES:
      return R_FIRST_ORDER;
    default: return R_CONST;
  }
}

// Variables fed to the network, in state order.
__constant__ int kInputVars[GAL_SURR_N_INPUT_VARS] = {
    GAL_V_HOT_MASS, GAL_V_HOT_ABUNDANCES, GAL_V_HOT_ANGULAR_MOMENTUM, GAL_V_HOT_OUTFLOWED_MASS,
    GAL_V_HOT_OUTFLOWED_ANGULAR_MOMENTUM, GAL_V_HOT_OUTFLOWED_ABUNDANCES, GAL_V_HOT_UNACCRETED_MASS,
    GAL_V_HOT_UNACCRETED_ABUNDANCES, GAL_V_DISK_MASS_STELLAR, GAL_V_DISK_MASS_STELLAR_FORMED,
    GAL_V_DISK_ABUNDANCES_STELLAR, GAL_V_DISK_MASS_GAS, GAL_V_DISK_ABUNDANCES_GAS, GAL_V_DISK_ANGULAR_MOMENTUM,
    GAL_V_SPHEROID_MASS_STELLAR, GAL_V_SPHEROID_MASS_STELLAR_FORMED, GAL_V_SPHEROID_ABUNDANCES_STELLAR,
    GAL_V_SPHEROID_MASS_GAS, GAL_V_SPHEROID_ABUNDANCES_GAS, GAL_V_SPHEROID_ANGULAR_MOMENTUM, GAL_V_BH_MASS,
    GAL_V_BH_SPIN};

struct NodeFrame {
  Halo H;
  double unit;      // f_b M_halo(t0)
  double s_mass, s_metal, s_ang, s_energy;
  double failed;
};

__device__ __forceinline__ NodeFrame node_frame(const double* __restrict__ P, const DerivedConstants& D,
                                                const NodePar& np, double t0, double spin_j) {
  NodeFrame f;
  f.H = halo(P, D, np, t0, spin_j);
  f.unit = D.f_baryon * f.H.mass;
  f.s_mass = 1.0e-4 * f.unit;
  f.s_metal = 1.0e-6 * f.unit;
  f.s_ang = 1.0e-4 * f.unit * f.H.r_vir * f.H.v_vir;
  f.s_energy = D.unit_energy * f.H.mass * f.H.v_vir * f.H.v_vir * 1.0e-3;
  f.failed = (f.H.a3 > D.a3_reion && f.H.v_vir < P[GAL_P_VELOCITY_SUPPRESSION_REIONIZATION]) ? 1.0 : 0.0;
  return f;
}

__device__ __forceinline__ double var_scale(const NodeFrame& f, int role) {
  switch (role) {
    case R_MASS: return f.s_mass;
    case R_METAL: return f.s_metal;
    case R_ANGMOM: return f.s_ang;
    case R_ENERGY: return f.s_energy;
    default: return 1.0;
  }
}

// The first-order increment of each trained variable in asinh space from one
// RHS evaluation at t0: for a growing quantity the Euler step taken in y,
// asinh((y + f dt)/s) - asinh(y/s) (exact for a constant rate); for a
// decaying one the Euler step taken in u, f dt / sqrt(s^2 + y^2) (exact for
// exponential decay, and never overshoots through zero).  The network
// learns only what this misses.
__device__ __forceinline__ void euler_baseline(const double* __restrict__ P, const DerivedConstants& D,
                                               const double* __restrict__ tab, const CoolingDesc& cd, int mode,
                                               const NodePar& np, const NodeFrame& f, double t0, double dt,
                                               const double (&y0)[kNV], double (&du)[kNV]) {
  double rate[kNV];
  rhs(P, D, tab, cd, mode, np, t0, y0, rate);
#pragma unroll
  for (int v = 0; v < kNV; ++v) {
    const int role = var_role(v);
    double d = 0.0;
    if (role == R_ENERGY) {
      d = asinh(rate[v] * dt / f.s_energy);
    } else if (role != R_CONST && role != R_SPIN && role != R_FIRST_ORDER) {
      const double s = var_scale(f, role);
      if (rate[v] >= 0.0)
        d = asinh((y0[v] + rate[v] * dt) / s) - asinh(y0[v] / s);          // linear growth
      else if (y0[v] > 0.0)
        d = asinh(y0[v] * exp(rate[v] * dt / y0[v]) / s) - asinh(y0[v] / s);  // decay on y/|f|
      else
        d = rate[v] * dt * rsqrt(s * s + y0[v] * y0[v]);
    }
    du[v] = d;
  }
}

__global__ __launch_bounds__(256) void encode_kernel(int n, int ld, const double* __restrict__ y0,
                                                     const double* __restrict__ y1, const double* __restrict__ npa,
                                                     const double* __restrict__ t0a, const double* __restrict__ t1a,
                                                     const double* __restrict__ P, DerivedConstants D,
                                                     const double* __restrict__ tab, CoolingDesc cd, int mode,
                                                     const float* __restrict__ norm, uint16_t* __restrict__ x,
                                                     float* __restrict__ x_raw, float* __restrict__ target) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const NodePar np = load_node_par(npa, ld, i);
  const double t0 = t0a[i];
  const double dt = t1a[i] - t0;
  const NodeFrame f = node_frame(P, D, np, t0, y0[GAL_V_SPIN_ANGULAR_MOMENTUM * (size_t)ld + i]);
  float in[kIn];
#pragma unroll
  for (int j = 0; j < GAL_SURR_N_INPUT_VARS; ++j) {
    const int v = kInputVars[j];
    in[j] = (float)asinh(y0[v * (size_t)ld + i] / var_scale(f, var_role(v)));
  }
  const double dtp = fmax(dt, 1.0e-12);
  in[22] = (float)log(dtp);
  in[23] = (float)(log10(f.H.mass) - 11.0);
  in[24] = (float)(0.1 * t0);
  in[25] = (float)((np.satellite ? 0.0 : np.m_rate) * dt / f.H.mass);
  in[26] = (float)log(fmax(f.H.lambda, 1.0e-4));
  in[27] = (float)(y0[GAL_V_HOT_OUTER_RADIUS * (size_t)ld + i] * f.H.inv_r_vir);
  in[28] = (float)(f.H.v_vir / 200.0);
  in[29] = (float)f.failed;
  in[30] = (float)log(dtp * f.H.inv_t_dyn);
  in[31] = 0.0f;
  if (x_raw) {
#pragma unroll
    for (int j = 0; j < kIn; ++j) x_raw[(size_t)i * kIn + j] = in[j];
  }
  if (x) {
#pragma unroll
    for (int j = 0; j < kIn; ++j) {
      const float z = norm ? (in[j] - norm[j]) / norm[kIn + j] : in[j];
      x[(size_t)i * kIn + j] = f2bf(z);
    }
  }
  if (target && y1) {
    double ys[kNV], du[kNV];
#pragma unroll
    for (int v = 0; v < kNV; ++v) ys[v] = y0[v * (size_t)ld + i];
    euler_baseline(P, D, tab, cd, mode, np, f, t0, dt, ys, du);
    float* tg = target + (size_t)i * kOutLd;
    for (int v = 0; v < GAL_NVAR; ++v) {
      const int role = var_role(v);
      const double a = ys[v], b = y1[v * (size_t)ld + i];
      double d = 0.0;
      if (role == R_ENERGY) {
        d = asinh((b - a) / f.s_energy) - du[v];
      } else if (role != R_CONST && role != R_SPIN && role != R_FIRST_ORDER) {
        const double s = var_scale(f, role);
        d = asinh(b / s) - asinh(a / s) - du[v];
      }
      float z = (float)d;
      if (norm) z = (float)asinh((d - (double)norm[2 * kIn + v]) / (double)norm[3 * kIn + v]);
      tg[v] = z;
    }
    tg[GAL_MLP_HEAD] = 0.0f;
    tg[31] = 0.0f;
  }
}

struct GateArgs {
  int n, ld;
  double* y;
  const double* np;
  const double* t0;
  const double* t1;
  const float* out;
  const float* norm;
  double* scale;
  int* status;
  int* accepted;
  int* reject_list;
  int* reject_count;
  int index_offset;   // global index of node 0 of this call (chunking)
  const double* P;
  const double* tab;
  DerivedConstants D;
  CoolingDesc cd;
  int mode;
  gal_surrogate_gate gate;
};

// Decode the network's increments for one node, check them and either
// commit the new state or queue the node for the integrator.
__global__ __launch_bounds__(256) void apply_kernel(GateArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool reject = false;
  if (i < a.n) {
    const size_t ld = (size_t)a.ld;
    const NodePar np = load_node_par(a.np, a.ld, i);
    const double t0 = a.t0[i];
    const double dt = a.t1[i] - t0;
    double y0[kNV], y1[kNV];
#pragma unroll
    for (int v = 0; v < kNV; ++v) y0[v] = a.y[v * ld + i];
    if (dt > 0.0) {
      const NodeFrame f = node_frame(a.P, a.D, np, t0, y0[GAL_V_SPIN_ANGULAR_MOMENTUM]);
      const float* o = a.out + (size_t)i * kOutLd;
      const float* nm = a.norm;
      bool ok = (double)o[GAL_MLP_HEAD] <= a.gate.log_error_max;
      double rate[kNV];
      rhs(a.P, a.D, a.tab, a.cd, a.mode, np, t0, y0, rate);
#pragma unroll
      for (int v = 0; v < kNV; ++v) {
        const int role = var_role(v);
        if (role == R_CONST) {
          y1[v] = y0[v];
        } else if (role == R_FIRST_ORDER) {
          y1[v] = y0[v] + rate[v] * dt;
        } else if (role == R_SPIN) {
          y1[v] = y0[v] + (np.satellite ? 0.0 : np.spin_rate) * dt;
        } else {
          // u1 = asinh(y0/s) + du + c with du the first-order increment of
          // euler_baseline and c = m + sigma sinh(o) the network's residual.
          // Written as u1 = asinh(Y/s) + C for the first-order end point Y,
          // y1 = s sinh(u1) = Y cosh C + sqrt(s^2 + Y^2) sinh C: no asinh,
          // and sinh C, cosh C from one expm1 (accurate for small C).
          const double ov = fmin(fmax((double)o[v], -(double)kOutClip), (double)kOutClip);
          double C = (double)nm[2 * kIn + v] + (double)nm[3 * kIn + v] * sinh(ov);
          const double s = var_scale(f, role);
          const double step = rate[v] * dt;
          double Y;
          if (role == R_ENERGY) {
            Y = step;                                  // energy: y1 = y0 + s sinh(asinh(f dt/s) + c)
          } else if (rate[v] >= 0.0) {
            Y = y0[v] + step;                          // linear growth
          } else if (y0[v] > 0.0) {
            Y = y0[v] * exp(step / y0[v]);             // decay on y/|f|
          } else {
            Y = y0[v];                                 // Euler step in u
            C += step * rsqrt(s * s + y0[v] * y0[v]);
          }
          const double em = expm1(C);
          const double e = 1.0 + em;
          const double sinh_c = 0.5 * em * (1.0 + 1.0 / e);
          const double cosh_c = 1.0 + 0.5 * em * em / e;
          const double w = Y * cosh_c + sqrt(s * s + Y * Y) * sinh_c;
          y1[v] = role == R_ENERGY ? y0[v] + w : w;
        }
        // a small negative value of a non-negative quantity is set to zero,
        // as the reference's post-step fixes do after an integrator step
        if (var_non_negative(v) && y1[v] < 0.0 && role != R_ENERGY &&
            y1[v] >= -a.gate.negative_tolerance * var_scale(f, role))
          y1[v] = 0.0;
        ok = ok && isfinite(y1[v]);
        ok = ok && !(var_non_negative(v) && y1[v] < 0.0);
      }
      // Baryon budget: accretion f_b Mdot dt is the only source; black-hole
      // growth is the only sink (radiated and jet energy).
      const double accreted = np.satellite ? 0.0 : a.D.f_baryon * np.m_rate * dt;
      auto baryons = [](const double* s) {
        return s[GAL_V_HOT_MASS] + s[GAL_V_HOT_OUTFLOWED_MASS] + s[GAL_V_HOT_UNACCRETED_MASS] +
               s[GAL_V_HOT_STRIPPED_MASS] + s[GAL_V_DISK_MASS_GAS] + s[GAL_V_DISK_MASS_STELLAR] +
               s[GAL_V_SPHEROID_MASS_GAS] + s[GAL_V_SPHEROID_MASS_STELLAR] + s[GAL_V_BH_MASS];
      };
      auto metals = [](const double* s) {
        return s[GAL_V_HOT_ABUNDANCES] + s[GAL_V_HOT_OUTFLOWED_ABUNDANCES] + s[GAL_V_HOT_UNACCRETED_ABUNDANCES] +
               s[GAL_V_HOT_STRIPPED_ABUNDANCES] + s[GAL_V_DISK_ABUNDANCES_GAS] + s[GAL_V_DISK_ABUNDANCES_STELLAR] +
               s[GAL_V_SPHEROID_ABUNDANCES_GAS] + s[GAL_V_SPHEROID_ABUNDANCES_STELLAR];
      };
      const double mb0 = baryons(y0), mb1 = baryons(y1);
      const double dbh = fmax(y1[GAL_V_BH_MASS] - y0[GAL_V_BH_MASS], 0.0);
      const double rm = mb1 - mb0 - accreted;
      const double tol_m = a.gate.mass_tolerance * (mb0 + accreted);
      ok = ok && rm <= tol_m && rm >= -(tol_m + a.gate.bh_slack * dbh);
      // Metal budget: the yield of newly formed stars is the only source.
      const double formed = (y1[GAL_V_DISK_MASS_STELLAR_FORMED] - y0[GAL_V_DISK_MASS_STELLAR_FORMED]) +
                            (y1[GAL_V_SPHEROID_MASS_STELLAR_FORMED] - y0[GAL_V_SPHEROID_MASS_STELLAR_FORMED]);
      const double src = a.P[GAL_P_METAL_YIELD] * formed;
      const double z0 = metals(y0), z1 = metals(y1);
      const double rz = z1 - z0 - src;
      const double tol_z = a.gate.metal_tolerance * (z0 + fabs(src)) + 1.0e-3 * f.s_metal;
      ok = ok && rz <= tol_z && rz >= -(tol_z + 0.1 * a.gate.bh_slack * dbh);
      if (ok) {
        double sc[kNV];
        scales(a.P, a.D, np, t0, y0, sc);
#pragma unroll
        for (int v = 0; v < kNV; ++v) {
          a.y[v * ld + i] = y1[v];
          if (a.scale) a.scale[v * ld + i] = sc[v];
        }
      }
      reject = !ok;
    }
    if (!reject) a.status[i] = GAL_STATUS_SUCCESS;
    a.accepted[i] = reject ? 0 : 1;
  }
  // Wave-aggregated append of the rejected nodes.
  const unsigned long long mask = __ballot(reject);
  if (mask) {
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(a.reject_count, __popcll(mask));
    base = __shfl(base, leader);
    if (reject) {
      const unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
      a.reject_list[base + __popcll(below)] = a.index_offset + i;
    }
  }
}

}  // namespace

// ============================================================================
// The network handle
// ============================================================================
struct gal_mlp {
  int device = 0;
  int batch = 0, splits = 0, rows_per_split = 0;
  int64_t w_off[kLayers] = {}, b_off[kLayers] = {}, pw_off[kLayers] = {}, pb_off[kLayers] = {};
  float* d_param = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  uint16_t* d_wbf[kLayers] = {};
  uint16_t* d_wt[kLayers] = {};   // transposed bf16 weights for the fused inference kernel
  uint16_t* d_x = nullptr;
  float* d_t = nullptr;
  uint16_t* d_z[3] = {};
  uint16_t* d_h[3] = {};
  float* d_out = nullptr;
  uint16_t* d_dout = nullptr;
  uint16_t* d_dz[2] = {};
  float* d_pw = nullptr;
  float* d_pb = nullptr;
  float* d_loss_partial = nullptr;
  long long* d_step = nullptr;
  long long* d_cursor = nullptr;
  int* d_local = nullptr;
  float* d_loss_hist = nullptr;
  int loss_hist_cap = 0;
  float lr0 = 1.0e-3f, lr_min = 1.0e-5f, beta1 = 0.9f, beta2 = 0.999f, eps = 1.0e-8f;
  int64_t total_steps = 1;
  float var_w[32] = {}, sigma[32] = {};
  unsigned head_mask = 0;
  float head_w = 0.1f, inv_wsum = 0.0f;
  long long ds_n = 0;
  const uint16_t* ds_x = nullptr;
  const float* ds_t = nullptr;
  const int* ds_perm = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
  int inf_rows = 0;
  uint16_t* d_ix = nullptr;
  uint16_t* d_ih[2] = {};
  float* d_iout = nullptr;
  int rej_cap = 0;
  int* d_rej = nullptr;
  int* d_rej_count = nullptr;
};

namespace {

constexpr int kInferChunk = 1 << 17;

template <bool AT, bool BT, int EPI>
int gemm(const GemmArgs& g, int splits, hipStream_t s) {
  // Operand loads are 16-byte vectors along the contiguous dimension.
  if ((!AT && (g.K % 8 || g.lda % 8)) || (AT && (g.M % 8 || g.lda % 8)) || (BT && (g.K % 8 || g.ldb % 8)) ||
      (!BT && (g.N % 8 || g.ldb % 8)) || g.k_split % TK)
    return GAL_E_INVALID;
  const dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, splits);
  if (grid.y > 65535u) return GAL_E_INVALID;
  hipLaunchKernelGGL((gemm_kernel<AT, BT, EPI>), grid, dim3(256), 0, s, g);
  return hipGetLastError() == hipSuccess ? GAL_OK : GAL_E_HIP;
}

GemmArgs gemm_args(const uint16_t* A, const uint16_t* B, int M, int N, int K, int lda, int ldb, int ldc) {
  GemmArgs g{};
  g.A = A; g.B = B; g.M = M; g.N = N; g.K = K; g.lda = lda; g.ldb = ldb; g.ldc = ldc;
  g.n_valid = N;
  g.k_split = ((K + TK - 1) / TK) * TK;
  return g;
}

#define GAL_TRY(x)                 \
  do {                             \
    const int rc_ = (x);           \
    if (rc_ != GAL_OK) return rc_; \
  } while (0)

// Forward through the four layers for n rows of x; training keeps the
// pre-activations and activations of each hidden layer.
int forward(gal_mlp* m, int n, const uint16_t* x, uint16_t* const z[3], uint16_t* const h[3], float* out,
            hipStream_t s) {
  const uint16_t* in = x;
  for (int l = 0; l < 3; ++l) {
    GemmArgs g = gemm_args(in, m->d_wbf[l], n, kFanOut[l], kFanIn[l], kFanIn[l], kLdW[l], kFanOut[l]);
    g.bias = m->d_param + m->b_off[l];
    g.out_bf = h[l];
    g.z_bf = z ? z[l] : nullptr;
    GAL_TRY((gemm<false, false, EPI_SILU>(g, 1, s)));
    in = h[l];
  }
  GemmArgs g = gemm_args(in, m->d_wbf[3], n, kOutLd, kHid, kHid, kLdW[3], kOutLd);
  g.n_valid = kOut;
  g.bias = m->d_param + m->b_off[3];
  g.out_f = out;
  return gemm<false, false, EPI_LINEAR>(g, 1, s);
}

// Backward pass from the output gradient already in d_dout; writes the
// per-split partial gradients of every layer.
int backward(gal_mlp* m, int n, const uint16_t* x, hipStream_t s) {
  const int rps = ((n + m->splits - 1) / m->splits + TK - 1) / TK * TK;
  const int splits = (n + rps - 1) / rps;
  const uint16_t* grad = m->d_dout;  // gradient wrt this layer's output
  int ld_grad = kOutLd;
  int pp = 0;
  for (int l = kLayers - 1; l >= 0; --l) {
    const uint16_t* act_in = l == 0 ? x : m->d_h[l - 1];
    // weight gradient: act_in^T (fan_in x n) . grad (n x ldW)
    GemmArgs gw = gemm_args(act_in, grad, kFanIn[l], kLdW[l], n, kFanIn[l], ld_grad, kLdW[l]);
    gw.k_split = rps;
    gw.out_f = m->d_pw + m->pw_off[l];
    GAL_TRY((gemm<true, false, EPI_PARTIAL>(gw, splits, s)));
    hipLaunchKernelGGL(colsum_kernel, dim3((kLdW[l] + 63) / 64, splits), dim3(256), 0, s, grad, n, kLdW[l], ld_grad,
                       rps, m->d_pb + m->pb_off[l]);
    if (l == 0) break;
    // gradient wrt the pre-activation of layer l-1: (grad . W_l^T) * silu'(z_{l-1})
    GemmArgs gx = gemm_args(grad, m->d_wbf[l], n, kHid, kLdW[l], ld_grad, kLdW[l], kHid);
    gx.z_in = m->d_z[l - 1];
    gx.out_bf = m->d_dz[pp];
    GAL_TRY((gemm<false, true, EPI_DSILU>(gx, 1, s)));
    grad = m->d_dz[pp];
    ld_grad = kHid;
    pp ^= 1;
  }
  m->rows_per_split = rps;
  return hipGetLastError() == hipSuccess ? GAL_OK : GAL_E_HIP;
}

LossArgs loss_args(gal_mlp* m, int n, const float* target) {
  LossArgs a{};
  a.out = m->d_out;
  a.target = target;
  a.dout = m->d_dout;
  a.loss_partial = m->d_loss_partial;
  a.B = n;
  for (int v = 0; v < 32; ++v) { a.w[v] = m->var_w[v]; a.sigma[v] = m->sigma[v]; }
  a.head_mask = m->head_mask;
  a.head_w = m->head_w;
  a.inv_wsum = m->inv_wsum;
  a.inv_B = 1.0f / (float)n;
  return a;
}

AdamArgs adam_args(gal_mlp* m, int n, float* grad_out, int apply) {
  AdamArgs a{};
  a.param = m->d_param; a.m = m->d_m; a.v = m->d_v;
  for (int l = 0; l < kLayers; ++l) {
    a.wbf[l] = m->d_wbf[l];
    a.w_off[l] = m->w_off[l]; a.b_off[l] = m->b_off[l]; a.pw_off[l] = m->pw_off[l]; a.pb_off[l] = m->pb_off[l];
  }
  const int rps = ((n + m->splits - 1) / m->splits + TK - 1) / TK * TK;
  a.splits = (n + rps - 1) / rps;
  a.grad_out = grad_out;
  a.pw = m->d_pw; a.pb = m->d_pb;
  a.lr0 = m->lr0; a.lr_min = m->lr_min; a.beta1 = m->beta1; a.beta2 = m->beta2; a.eps = m->eps;
  a.total_steps = m->total_steps;
  a.step = m->d_step;
  a.apply = apply;
  return a;
}

int grad_pass(gal_mlp* m, int n, const uint16_t* x, const float* target, hipStream_t s) {
  GAL_TRY(forward(m, n, x, m->d_z, m->d_h, m->d_out, s));
  hipLaunchKernelGGL(loss_kernel, dim3((n + 255) / 256), dim3(256), 0, s, loss_args(m, n, target));
  return backward(m, n, x, s);
}

int record_train_step(gal_mlp* m, hipStream_t s) {
  const int B = m->batch;
  hipLaunchKernelGGL(gather_kernel, dim3((B + 255) / 256), dim3(256), 0, s, m->ds_x, m->ds_t, m->ds_perm, m->ds_n,
                     (const long long*)m->d_cursor, B, m->d_x, m->d_t);
  GAL_TRY(grad_pass(m, B, m->d_x, m->d_t, s));
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((kParams + 255) / 256)), dim3(256), 0, s,
                     adam_args(m, B, nullptr, 1));
  hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, s, m->d_step, m->d_cursor, m->d_local,
                     (const float*)m->d_loss_partial, (B + 255) / 256, 1.0f / (float)B, m->d_loss_hist,
                     m->loss_hist_cap, B);
  return hipGetLastError() == hipSuccess ? GAL_OK : GAL_E_HIP;
}

void release_graph(gal_mlp* m) {
  if (m->exec) hipGraphExecDestroy(m->exec);
  if (m->graph) hipGraphDestroy(m->graph);
  m->exec = nullptr;
  m->graph = nullptr;
  m->graph_valid = false;
}

int ensure_inference(gal_mlp* m, int rows) {
  if (rows <= m->inf_rows) return GAL_OK;
  hipFree(m->d_ix); hipFree(m->d_ih[0]); hipFree(m->d_ih[1]); hipFree(m->d_iout);
  m->d_ix = nullptr; m->d_ih[0] = m->d_ih[1] = nullptr; m->d_iout = nullptr; m->inf_rows = 0;
  if (hipMalloc(&m->d_ix, sizeof(uint16_t) * kIn * (size_t)rows) != hipSuccess ||
      hipMalloc(&m->d_ih[0], sizeof(uint16_t) * kHid * (size_t)rows) != hipSuccess ||
      hipMalloc(&m->d_ih[1], sizeof(uint16_t) * kHid * (size_t)rows) != hipSuccess ||
      hipMalloc(&m->d_iout, sizeof(float) * kOutLd * (size_t)rows) != hipSuccess)
    return GAL_E_NOMEM;
  m->inf_rows = rows;
  return GAL_OK;
}

// Per-layer GEMM path (the training kernels' forward).
int infer_layers(gal_mlp* m, int n, const uint16_t* x, float* out, hipStream_t s) {
  uint16_t* h[3] = {m->d_ih[0], m->d_ih[1], m->d_ih[0]};
  return forward(m, n, x, nullptr, h, out, s);
}

// Refresh the transposed weights from the current bf16 weights.
int prepare_fused(gal_mlp* m, hipStream_t s) {
  // the LDS limit of the kernel on the current device (a cheap host call; no
  // process-wide state)
  static_assert(kFusedLds <= 163840, "fused forward LDS must fit a CU");
  HIP_TRY(hipFuncSetAttribute((const void*)mlp_forward_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kFusedLds));
  for (int l = 0; l < kLayers; ++l) {
    const int e = kFanIn[l] * kLdW[l];
    hipLaunchKernelGGL(transpose_weights_kernel, dim3((e + 255) / 256), dim3(256), 0, s, m->d_wbf[l], kFanIn[l],
                       kLdW[l], m->d_wt[l]);
  }
  return hipGetLastError() == hipSuccess ? GAL_OK : GAL_E_HIP;
}

// The fused one-kernel forward (after prepare_fused).
int infer(gal_mlp* m, int n, const uint16_t* x, float* out, hipStream_t s) {
  if (n <= 0) return GAL_OK;
  FusedArgs a{};
  a.x = x;
  for (int l = 0; l < kLayers; ++l) { a.wt[l] = m->d_wt[l]; a.bias[l] = m->d_param + m->b_off[l]; }
  a.out = out;
  a.n = n;
  hipLaunchKernelGGL(mlp_forward_fused, dim3((n + FR - 1) / FR), dim3(256), kFusedLds, s, a);
  return hipGetLastError() == hipSuccess ? GAL_OK : GAL_E_HIP;
}

}  // namespace

extern "C" {

int64_t gal_mlp_param_count(void) { return kParams; }

int gal_mlp_create(int batch, gal_mlp** out) {
  if (!out || batch < 1 || batch > 65535 * TM) return GAL_E_INVALID;
  *out = nullptr;
  gal_mlp* m = new gal_mlp();
  hipGetDevice(&m->device);
  m->batch = batch;
  m->splits = std::max(1, std::min(32, batch / 256));
  int64_t off = 0, pw = 0, pb = 0;
  for (int l = 0; l < kLayers; ++l) {
    m->w_off[l] = off;
    off += (int64_t)kFanIn[l] * kFanOut[l];
    m->b_off[l] = off;
    off += kFanOut[l];
    m->pw_off[l] = pw;
    pw += (int64_t)m->splits * kFanIn[l] * kLdW[l];
    m->pb_off[l] = pb;
    pb += (int64_t)m->splits * kLdW[l];
  }
  const size_t B = (size_t)batch;
  bool ok = hipMalloc(&m->d_param, sizeof(float) * kParams) == hipSuccess &&
            hipMalloc(&m->d_m, sizeof(float) * kParams) == hipSuccess &&
            hipMalloc(&m->d_v, sizeof(float) * kParams) == hipSuccess &&
            hipMalloc(&m->d_x, sizeof(uint16_t) * kIn * B) == hipSuccess &&
            hipMalloc(&m->d_t, sizeof(float) * kOutLd * B) == hipSuccess &&
            hipMalloc(&m->d_out, sizeof(float) * kOutLd * B) == hipSuccess &&
            hipMalloc(&m->d_dout, sizeof(uint16_t) * kOutLd * B) == hipSuccess &&
            hipMalloc(&m->d_pw, sizeof(float) * pw) == hipSuccess &&
            hipMalloc(&m->d_pb, sizeof(float) * pb) == hipSuccess &&
            hipMalloc(&m->d_loss_partial, sizeof(float) * ((B + 255) / 256)) == hipSuccess &&
            hipMalloc(&m->d_step, sizeof(long long)) == hipSuccess &&
            hipMalloc(&m->d_cursor, sizeof(long long)) == hipSuccess &&
            hipMalloc(&m->d_local, sizeof(int)) == hipSuccess;
  for (int l = 0; l < kLayers && ok; ++l)
    ok = hipMalloc(&m->d_wbf[l], sizeof(uint16_t) * kFanIn[l] * kLdW[l]) == hipSuccess &&
         hipMalloc(&m->d_wt[l], sizeof(uint16_t) * kFanIn[l] * kLdW[l]) == hipSuccess;
  for (int l = 0; l < 3 && ok; ++l)
    ok = hipMalloc(&m->d_z[l], sizeof(uint16_t) * kHid * B) == hipSuccess &&
         hipMalloc(&m->d_h[l], sizeof(uint16_t) * kHid * B) == hipSuccess;
  for (int l = 0; l < 2 && ok; ++l) ok = hipMalloc(&m->d_dz[l], sizeof(uint16_t) * kHid * B) == hipSuccess;
  // the training stream is created on the first training call: every stream
  // holds a hardware queue for the rest of the process, and inference-only
  // handles never need one
  ok = ok && hipEventCreateWithFlags(&m->ev_in, hipEventDisableTiming) == hipSuccess &&
       hipEventCreateWithFlags(&m->ev_out, hipEventDisableTiming) == hipSuccess;
  if (ok) {
    ok = hipMemset(m->d_param, 0, sizeof(float) * kParams) == hipSuccess &&
         hipMemset(m->d_m, 0, sizeof(float) * kParams) == hipSuccess &&
         hipMemset(m->d_v, 0, sizeof(float) * kParams) == hipSuccess &&
         hipMemset(m->d_step, 0, sizeof(long long)) == hipSuccess &&
         hipMemset(m->d_dout, 0, sizeof(uint16_t) * kOutLd * B) == hipSuccess;
    for (int l = 0; l < kLayers && ok; ++l)
      ok = hipMemset(m->d_wbf[l], 0, sizeof(uint16_t) * kFanIn[l] * kLdW[l]) == hipSuccess;
    ok = ok && hipDeviceSynchronize() == hipSuccess;
  }
  if (!ok) {
    gal_mlp_destroy(m);
    return GAL_E_NOMEM;
  }
  for (int v = 0; v < 30; ++v) { m->var_w[v] = 1.0f; m->sigma[v] = 1.0f; }
  m->inv_wsum = 1.0f / 30.0f;
  *out = m;
  return GAL_OK;
}

int gal_mlp_destroy(gal_mlp* m) {
  if (!m) return GAL_OK;
  release_graph(m);
  hipFree(m->d_param); hipFree(m->d_m); hipFree(m->d_v);
  for (int l = 0; l < kLayers; ++l) { hipFree(m->d_wbf[l]); hipFree(m->d_wt[l]); }
  hipFree(m->d_x); hipFree(m->d_t);
  for (int l = 0; l < 3; ++l) { hipFree(m->d_z[l]); hipFree(m->d_h[l]); }
  hipFree(m->d_out); hipFree(m->d_dout); hipFree(m->d_dz[0]); hipFree(m->d_dz[1]);
  hipFree(m->d_pw); hipFree(m->d_pb); hipFree(m->d_loss_partial);
  hipFree(m->d_step); hipFree(m->d_cursor); hipFree(m->d_local); hipFree(m->d_loss_hist);
  hipFree(m->d_ix); hipFree(m->d_ih[0]); hipFree(m->d_ih[1]); hipFree(m->d_iout);
  hipFree(m->d_rej); hipFree(m->d_rej_count);
  if (m->ev_in) hipEventDestroy(m->ev_in);
  if (m->ev_out) hipEventDestroy(m->ev_out);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return GAL_OK;
}

int gal_mlp_set_params(gal_mlp* m, const float* params, int64_t n) {
  if (!m || !params || n != kParams) return GAL_E_INVALID;
  HIP_TRY(hipMemcpy(m->d_param, params, sizeof(float) * kParams, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(m->d_m, 0, sizeof(float) * kParams));
  HIP_TRY(hipMemset(m->d_v, 0, sizeof(float) * kParams));
  HIP_TRY(hipMemset(m->d_step, 0, sizeof(long long)));
  for (int l = 0; l < kLayers; ++l) {
    const int64_t cnt = (int64_t)kFanIn[l] * kFanOut[l];
    hipLaunchKernelGGL(weights_to_bf16_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, 0,
                       (const float*)m->d_param, m->w_off[l], kFanIn[l], kFanOut[l], kLdW[l], m->d_wbf[l]);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return GAL_OK;
}

int gal_mlp_get_params(gal_mlp* m, float* params, int64_t n) {
  if (!m || !params || n != kParams) return GAL_E_INVALID;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(params, m->d_param, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return GAL_OK;
}

int gal_mlp_set_adam(gal_mlp* m, double lr, double lr_min, double beta1, double beta2, double eps,
                     int64_t total_steps) {
  if (!m || !(lr > 0.0) || lr_min < 0.0 || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) ||
      !(eps > 0.0) || total_steps < 1)
    return GAL_E_INVALID;
  m->lr0 = (float)lr; m->lr_min = (float)lr_min; m->beta1 = (float)beta1; m->beta2 = (float)beta2;
  m->eps = (float)eps; m->total_steps = total_steps;
  release_graph(m);
  return GAL_OK;
}

int gal_mlp_set_loss(gal_mlp* m, const float* var_weight, const float* target_sigma, const int* head_vars,
                     double head_weight) {
  if (!m || !var_weight || !target_sigma || !head_vars || head_weight < 0.0) return GAL_E_INVALID;
  double ws = 0.0;
  unsigned mask = 0;
  for (int v = 0; v < 30; ++v) {
    if (var_weight[v] < 0.0f) return GAL_E_INVALID;
    m->var_w[v] = var_weight[v];
    m->sigma[v] = target_sigma[v];
    ws += var_weight[v];
    if (head_vars[v]) mask |= 1u << v;
  }
  if (!(ws > 0.0)) return GAL_E_INVALID;
  m->inv_wsum = (float)(1.0 / ws);
  m->head_mask = mask;
  m->head_w = (float)head_weight;
  release_graph(m);
  return GAL_OK;
}

int gal_mlp_forward_d(gal_mlp* m, int64_t n, const uint16_t* x, float* out, void* stream) {
  if (!m || n < 0 || (n > 0 && (!x || !out))) return GAL_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int chunk = (int)std::min<int64_t>(n, kInferChunk);
  if (n == 0) return GAL_OK;
  GAL_TRY(prepare_fused(m, s));
  for (int64_t c0 = 0; c0 < n; c0 += chunk) {
    const int nc = (int)std::min<int64_t>(chunk, n - c0);
    GAL_TRY(infer(m, nc, x + c0 * kIn, out + c0 * kOutLd, s));
  }
  return GAL_OK;
}

int gal_mlp_forward_layers_d(gal_mlp* m, int64_t n, const uint16_t* x, float* out, void* stream) {
  if (!m || n < 0 || (n > 0 && (!x || !out))) return GAL_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int chunk = (int)std::min<int64_t>(n, kInferChunk);
  if (n == 0) return GAL_OK;
  GAL_TRY(ensure_inference(m, chunk));
  for (int64_t c0 = 0; c0 < n; c0 += chunk) {
    const int nc = (int)std::min<int64_t>(chunk, n - c0);
    GAL_TRY(infer_layers(m, nc, x + c0 * kIn, out + c0 * kOutLd, s));
  }
  return GAL_OK;
}

int gal_mlp_gradient_d(gal_mlp* m, int n, const uint16_t* x, const float* target, float* grad, float* loss,
                       void* stream) {
  if (!m || n < 1 || n > m->batch || !x || !target) return GAL_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  GAL_TRY(grad_pass(m, n, x, target, s));
  if (grad)
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((kParams + 255) / 256)), dim3(256), 0, s,
                       adam_args(m, n, grad, 0));
  if (loss)
    hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, s, (const float*)m->d_loss_partial, (n + 255) / 256,
                       1.0f / (float)n, loss);
  HIP_TRY(hipGetLastError());
  return GAL_OK;
}

int gal_mlp_set_dataset(gal_mlp* m, int64_t n, const uint16_t* x, const float* target, const int* perm) {
  if (!m || n < 1 || !x || !target || !perm) return GAL_E_INVALID;
  if ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(target)) & 15u) return GAL_E_INVALID;
  m->ds_n = n; m->ds_x = x; m->ds_t = target; m->ds_perm = perm;
  release_graph(m);
  return GAL_OK;
}

int gal_mlp_train_d(gal_mlp* m, int n_steps, int use_graph, float* loss_trace, void* stream) {
  if (!m || n_steps < 0 || !m->ds_x) return GAL_E_INVALID;
  if (n_steps == 0) return GAL_OK;
  if (!m->stream) HIP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  if (n_steps > m->loss_hist_cap) {
    release_graph(m);
    hipFree(m->d_loss_hist);
    m->d_loss_hist = nullptr;
    m->loss_hist_cap = 0;
    HIP_TRY(hipMalloc(&m->d_loss_hist, sizeof(float) * (size_t)n_steps));
    m->loss_hist_cap = n_steps;
  }
  hipStream_t cs = (hipStream_t)stream;
  HIP_TRY(hipEventRecord(m->ev_in, cs));
  HIP_TRY(hipStreamWaitEvent(m->stream, m->ev_in, 0));
  hipLaunchKernelGGL(reset_train_kernel, dim3(1), dim3(64), 0, m->stream, m->d_cursor, m->d_local);
  HIP_TRY(hipGetLastError());
  if (use_graph) {
    if (!m->graph_valid) {
      release_graph(m);
      HIP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
      const int rc = record_train_step(m, m->stream);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(m->stream, &g);
      if (rc != GAL_OK || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        hipGetLastError();
        return rc != GAL_OK ? rc : GAL_E_HIP;
      }
      m->graph = g;
      HIP_TRY(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
      m->graph_valid = true;
    }
    for (int i = 0; i < n_steps; ++i) HIP_TRY(hipGraphLaunch(m->exec, m->stream));
  } else {
    for (int i = 0; i < n_steps; ++i) GAL_TRY(record_train_step(m, m->stream));
  }
  if (loss_trace)
    HIP_TRY(hipMemcpyAsync(loss_trace, m->d_loss_hist, sizeof(float) * n_steps, hipMemcpyDeviceToDevice, m->stream));
  HIP_TRY(hipEventRecord(m->ev_out, m->stream));
  HIP_TRY(hipStreamWaitEvent(cs, m->ev_out, 0));
  return GAL_OK;
}

int gal_surrogate_encode_d(gal_ode_batch* h, int n, int ld, const double* y0, const double* y1,
                           const double* node_par, const double* t0, const double* t1, const float* norm,
                           uint16_t* x, float* x_raw, float* target, void* stream) {
  if (!h || n < 0 || ld < n || !h->have_physics) return GAL_E_INVALID;
  if (n == 0) return GAL_OK;
  if (!y0 || !node_par || !t0 || !t1 || (target && !y1)) return GAL_E_INVALID;
  if (target && !h->have_table) return GAL_E_INVALID;
  hipLaunchKernelGGL(encode_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, ld, y0, y1, node_par,
                     t0, t1, (const double*)h->d_params, h->D, (const double*)h->d_table, h->cd,
                     (int)h->params[GAL_P_PHYSICS_MODE], norm, x, x_raw, target);
  HIP_TRY(hipGetLastError());
  return GAL_OK;
}

static int apply_chunk(gal_ode_batch* h, const gal_surrogate_gate* gate, int n, int ld, double* y,
                       const double* node_par, const double* t0, const double* t1, const float* mlp_out,
                       const float* norm, double* scale, int* status, int* accepted, int* reject_list,
                       int* reject_count, int index_offset, hipStream_t s) {
  GateArgs a{};
  a.n = n; a.ld = ld; a.y = y; a.np = node_par; a.t0 = t0; a.t1 = t1; a.out = mlp_out; a.norm = norm;
  a.scale = scale; a.status = status; a.accepted = accepted; a.reject_list = reject_list;
  a.reject_count = reject_count; a.index_offset = index_offset;
  a.P = h->d_params; a.tab = h->d_table; a.D = h->D; a.cd = h->cd; a.mode = (int)h->params[GAL_P_PHYSICS_MODE];
  a.gate = *gate;
  hipLaunchKernelGGL(apply_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a);
  HIP_TRY(hipGetLastError());
  return GAL_OK;
}

int gal_surrogate_apply_d(gal_ode_batch* h, const gal_surrogate_gate* gate, int n, int ld, double* y,
                          const double* node_par, const double* t0, const double* t1, const float* mlp_out,
                          const float* norm, double* scale, int* status, int* accepted, int* reject_list,
                          int* reject_count, void* stream) {
  if (!h || !gate || n < 0 || ld < n || !h->have_physics || !h->have_table) return GAL_E_INVALID;
  if (n == 0) return GAL_OK;
  if (!y || !node_par || !t0 || !t1 || !mlp_out || !norm || !status || !accepted || !reject_list || !reject_count)
    return GAL_E_INVALID;
  return apply_chunk(h, gate, n, ld, y, node_par, t0, t1, mlp_out, norm, scale, status, accepted, reject_list,
                     reject_count, 0, (hipStream_t)stream);
}

int gal_surrogate_evolve_d(gal_ode_batch* h, gal_mlp* m, const float* norm, const gal_surrogate_gate* gate,
                           int n, int ld, double* y, const double* node_par, const double* t_start,
                           const double* t_end, double* h_step, double* scale, int* status, int* accepted,
                           int* n_rejected, void* stream) {
  if (!h || !m || !norm || !gate || n < 0 || ld < n || n > h->n_max || !h->have_physics || !h->have_table)
    return GAL_E_INVALID;
  if (n_rejected) *n_rejected = 0;
  if (n == 0) return GAL_OK;
  if (!y || !node_par || !t_start || !t_end || !h_step || !status || !accepted) return GAL_E_INVALID;
  if ((h->flags & GAL_FLAG_SCALES_PROVIDED) && !scale) return GAL_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  if (n > m->rej_cap) {
    hipFree(m->d_rej);
    m->d_rej = nullptr;
    m->rej_cap = 0;
    HIP_TRY(hipMalloc(&m->d_rej, sizeof(int) * (size_t)n));
    m->rej_cap = n;
  }
  if (!m->d_rej_count) HIP_TRY(hipMalloc(&m->d_rej_count, sizeof(int)));
  const int chunk = std::min(n, kInferChunk);
  GAL_TRY(ensure_inference(m, chunk));
  GAL_TRY(prepare_fused(m, s));
  HIP_TRY(hipMemsetAsync(m->d_rej_count, 0, sizeof(int), s));
  for (int c0 = 0; c0 < n; c0 += chunk) {
    const int nc = std::min(chunk, n - c0);
    GAL_TRY(gal_surrogate_encode_d(h, nc, ld, y + c0, nullptr, node_par + c0, t_start + c0, t_end + c0, norm,
                                   m->d_ix, nullptr, nullptr, s));
    GAL_TRY(infer(m, nc, m->d_ix, m->d_iout, s));
    // With provided scales the integrator's scales stay the caller's.
    double* sc = (h->flags & GAL_FLAG_SCALES_PROVIDED) ? nullptr : (scale ? scale + c0 : nullptr);
    GAL_TRY(apply_chunk(h, gate, nc, ld, y + c0, node_par + c0, t_start + c0, t_end + c0, m->d_iout, norm, sc,
                        status + c0, accepted + c0, m->d_rej, m->d_rej_count, c0, s));
  }
  // The integrator takes the rejected list with its count in device memory,
  // so nothing here waits for the GPU: the call only enqueues work on s
  // unless the caller asks for the count on the host.
  GAL_TRY(gal_ode_evolve_subset(h, n, ld, y, node_par, t_start, t_end, h_step, scale, status, nullptr, m->d_rej, n,
                                s, m->d_rej_count));
  if (n_rejected) {
    int n_rej = 0;
    HIP_TRY(hipMemcpyAsync(&n_rej, m->d_rej_count, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *n_rejected = n_rej;
  }
  return GAL_OK;
}

}  // extern "C"
