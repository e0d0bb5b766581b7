// Learned initialiser: a fully connected network n_in -> width^depth -> n_out
// with GELU activations, evaluated and trained on the GPU with BF16 MFMA
// products and FP32 accumulation.  It maps a voxel's normalised signal
// (measurements divided by a per-voxel scale, e.g. S0) to a starting point
// for Levenberg-Marquardt, standing in for GridSearch.  No counterpart in the
// reference.
//
// Products (all "NT": both operands K-contiguous, C = A B^T):
//   forward   H_l = gelu(H_{l-1} W_l^T + b_l)          A = H_{l-1} [B][K], B = W_l [N][K]
//   backward  dH_{l-1} = dZ_l W_l                        A = dZ_l [B][N],   B = W_l^T [K][N]
//             dW_l = dZ_l^T H_{l-1}                      A = dZ_l^T [N][B], B = H_{l-1}^T [K][B]
// so each activation is kept in both layouts (the forward epilogue writes the
// transposed copy), and each weight matrix in both layouts (the Adam step
// writes them).  One wave computes a 32x32 tile with
// v_mfma_f32_32x32x16_bf16 (operands read straight from global memory /
// L2: every matrix here is at most 256 wide); a block of 4 waves covers 64x64.
//
// Output head: each model parameter the network predicts is one entry
// (cdm_head_entry): a sigmoid onto [lo, hi] (linear or logarithmic), a
// sigmoid ratio of an earlier parameter, or an axis (three outputs, the
// direction of the output vector gives theta, phi).  Training minimises the
// soft-target cross entropy of the sigmoid outputs plus w (1 - (v.n)^2/|v|^2)
// for axes (sign-invariant: n and -n are the same fibre).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/cudimot_amd.h"
#include "../engine/philox.h"
#include "../host/model_ops.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMaxHead = 16;
constexpr int kMaxLayers = 8;
constexpr int kTile = 64;        // block tile (4 waves of 32x32)
constexpr int kInferChunk = 65536;
constexpr int kWgradSplits = 16;  // most K slices of a weight gradient

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// ------------------------------------------------------------------ math
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}
__device__ __forceinline__ float sigmoid(float o) { return 1.0f / (1.0f + __expf(-o)); }

// ------------------------------------------------------------------ GEMM
enum { EPI_PLAIN = 0, EPI_FWD_GELU = 1, EPI_BWD_GELU = 2 };

struct Epi {
  const float* bias;  // [N] or null
  float* out_f32;     // [M][ld_f32] (plain: C (+bias); forward: pre-activation z)
  int ld_f32;
  bf16* out_bf;       // [M][ld_bf]
  int ld_bf;
  bf16* out_bft;      // [N][ld_bft] transposed copy
  int ld_bft;
  const float* zin;   // [M][ld_zin] pre-activation (backward through gelu)
  int ld_zin;
  size_t split_stride;  // plain, split over K: partial z goes to out_f32 + z * split_stride
};

// C[M][N] = A[M][K] B[N][K]^T with M, N multiples of 32 and K of 16; lda, ldb
// multiples of 8.  Grid (ceil(N/64), ceil(M/64), S), block 256: with S > 1
// (plain epilogue only) block z sums the K slice [z K, (z+1) K) into its own
// partial C at out_f32 + z * split_stride, reduced afterwards in a fixed
// order.
// MFMA 32x32x16 operand layout (wave64): lane l supplies A row (l & 31) and
// B row (l & 31), elements k0 + 8 (l >> 5) .. +7; accumulator element j of
// lane l is C[8 (j >> 2) + 4 (l >> 5) + (j & 3)][l & 31].
// STEPS > 0: K == 16 STEPS, and every operand of the K range is loaded before
// the first product, with the epilogue's own inputs (bias, pre-activations),
// so one memory latency precedes the MFMA chain instead of one per chunk;
// STEPS == 0: any K, in prefetched chunks.  The products run in the same K
// order either way.
// (bx, by, bz): the block's tile column, tile row and K slice.
template <int EPI, int STEPS>
__device__ __forceinline__ void gemm_tile(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb,
                                          int M, int N, int K, const Epi& e, int bx, int by, int bz) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = by * kTile + (wave >> 1) * 32;
  const int n0 = bx * kTile + (wave & 1) * 32;
  if (m0 >= M || n0 >= N) return;
  const int r = lane & 31, half = lane >> 5;
  const size_t kb = (size_t)bz * K;
  const bf16* pa = A + (size_t)(m0 + r) * lda + kb + 8 * half;
  const bf16* pb = B + (size_t)(n0 + r) * ldb + kb + 8 * half;
  const int col = n0 + r;
  const float bias = (EPI != EPI_BWD_GELU && e.bias) ? e.bias[col] : 0.0f;
  f32x16 acc = {};
  float zpre[16];
  if constexpr (STEPS > 0) {
    bf16x8 a[STEPS], b[STEPS];
#pragma unroll
    for (int u = 0; u < STEPS; ++u) {
      a[u] = *(const bf16x8*)(pa + 16 * u);
      b[u] = *(const bf16x8*)(pb + 16 * u);
    }
    if (EPI == EPI_BWD_GELU) {
#pragma unroll
      for (int j = 0; j < 16; ++j) zpre[j] = e.zin[(size_t)(m0 + 8 * (j >> 2) + 4 * half + (j & 3)) * e.ld_zin + col];
    }
    // keep every load above ahead of the products (the scheduler would
    // otherwise interleave them to save registers)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < STEPS; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u], b[u], acc, 0, 0, 0);
  } else {
    // K in chunks of four MFMA steps, the next chunk's operands loaded before
    // the current chunk's products (each step's load would otherwise wait for
    // the memory latency alone), then the steps that remain
    constexpr int kChunk = 4 * 16;
    int k0 = 0;
    if (K >= kChunk) {
      bf16x8 a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = *(const bf16x8*)(pa + 16 * u);
        b[u] = *(const bf16x8*)(pb + 16 * u);
      }
      for (; k0 + kChunk <= K; k0 += kChunk) {
        bf16x8 an[4], bn[4];
        const bool more = k0 + 2 * kChunk <= K;
        if (more) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            an[u] = *(const bf16x8*)(pa + k0 + kChunk + 16 * u);
            bn[u] = *(const bf16x8*)(pb + k0 + kChunk + 16 * u);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u], b[u], acc, 0, 0, 0);
        if (more) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            a[u] = an[u];
            b[u] = bn[u];
          }
        }
      }
    }
    for (; k0 < K; k0 += 16) {
      const bf16x8 a = *(const bf16x8*)(pa + k0);
      const bf16x8 b = *(const bf16x8*)(pb + k0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    if (EPI == EPI_BWD_GELU) {
#pragma unroll
      for (int j = 0; j < 16; ++j) zpre[j] = e.zin[(size_t)(m0 + 8 * (j >> 2) + 4 * half + (j & 3)) * e.ld_zin + col];
    }
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int row0 = m0 + 8 * g + 4 * half;
    bf16x4 tv;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = row0 + i;
      float v = acc[4 * g + i];
      if (EPI == EPI_PLAIN) {
        e.out_f32[bz * e.split_stride + (size_t)row * e.ld_f32 + col] = v + bias;
      } else if (EPI == EPI_FWD_GELU) {
        const float z = v + bias;
        if (e.out_f32) e.out_f32[(size_t)row * e.ld_f32 + col] = z;
        v = gelu(z);
      } else {
        v *= gelu_grad(zpre[4 * g + i]);
      }
      if (EPI != EPI_PLAIN) {
        const bf16 h = (bf16)v;
        if (e.out_bf) e.out_bf[(size_t)row * e.ld_bf + col] = h;
        tv[i] = h;
      }
    }
    if (EPI != EPI_PLAIN && e.out_bft) *(bf16x4*)(e.out_bft + (size_t)col * e.ld_bft + row0) = tv;
  }
}

template <int EPI, int STEPS>
__global__ void __launch_bounds__(256) gemm_nt(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B,
                                               int ldb, int M, int N, int K, Epi e) {
  gemm_tile<EPI, STEPS>(A, lda, B, ldb, M, N, K, e, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Several independent plain products with the same K slice length in one
// launch (the weight gradients of every layer): problem p owns the blocks
// [first[p], first[p] + tiles_n tiles_m splits), tile column fastest, then
// tile row, then K slice.
struct GemmJob {
  const bf16* A;
  const bf16* B;
  int lda, ldb, M, N, tiles_n, tiles_m, first;
  Epi e;
};
// advance (a training-step graph's step index, read by no block of this
// launch): its first thread moves it on to the next step's entry.
struct GemmJobs {
  int n, K;
  int* advance;
  GemmJob j[kMaxLayers];
};

template <int STEPS>
__global__ void __launch_bounds__(256) gemm_multi(GemmJobs jobs) {
  if (jobs.advance && blockIdx.x == 0 && threadIdx.x == 0) *jobs.advance = *jobs.advance + 1;
  int b = blockIdx.x, p = 0;
  while (p + 1 < jobs.n && b >= jobs.j[p + 1].first) ++p;
  const GemmJob& J = jobs.j[p];
  b -= J.first;
  const int bx = b % J.tiles_n;
  b /= J.tiles_n;
  const int by = b % J.tiles_m, bz = b / J.tiles_m;
  gemm_tile<EPI_PLAIN, STEPS>(J.A, J.lda, J.B, J.ldb, J.M, J.N, jobs.K, J.e, bx, by, bz);
}

// ------------------------------------------------------------------ inputs
// x[v][k] = meas[v][k] / scale[v] for k < nmeas, 0 for the padding columns;
// rows past nrows are zero.  Optional transposed copy xt[k][v].
__global__ void inputs_kernel(int nrows, int rows_pad, int nmeas, int kpad, const float* __restrict__ meas,
                              const float* __restrict__ scale, bf16* x, bf16* xt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows_pad * kpad) return;
  const int v = i / kpad, k = i - v * kpad;
  float val = 0.0f;
  if (v < nrows && k < nmeas) val = meas[(size_t)v * nmeas + k] / (scale ? scale[v] : 1.0f);
  const bf16 h = (bf16)val;
  x[(size_t)v * kpad + k] = h;
  if (xt) xt[(size_t)k * rows_pad + v] = h;
}

// Training inputs with fresh Rician noise: |S + s n1 + i s n2| / scale with
// s = scale / SNR, SNR uniform in [snr_lo, snr_hi] per sample.  Draws are
// Philox blocks addressed by (sample, measurement, epoch) under the seed.
// Arguments of one training step that change from step to step: the batch's
// first sample, the epoch (noise stream), the learning rate and Adam's bias
// corrections.  A replayed graph reads them from a device table at the
// index *ctr, which its weight-gradient launch advances (the Adam step after
// it reads the entry behind); eager launches pass tab null.
struct StepArgs {
  long long first;
  unsigned ep;
  float lr, c1, c2;
};

__global__ void noisy_inputs_kernel(int rows, int nmeas, int kpad, long long first, const float* __restrict__ clean,
                                    const float* __restrict__ scale, float snr_lo, float snr_hi, uint64_t seed,
                                    uint32_t epoch, bf16* x, bf16* xt, const StepArgs* tab, const int* ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * kpad) return;
  if (tab) {
    first = tab[*ctr].first;
    epoch = tab[*ctr].ep;
  }
  const int v = i / kpad, k = i - v * kpad;
  float val = 0.0f;
  if (k < nmeas) {
    const uint64_t s = (uint64_t)(first + v);
    const float sc = scale ? scale[s] : 1.0f;
    cdm::U32x4 c;
    c.v[0] = (uint32_t)s;
    c.v[1] = (uint32_t)(s >> 32);
    c.v[2] = 0xFFFFFFFFu;
    c.v[3] = epoch;
    const cdm::U32x4 ws = cdm::philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float snr = snr_lo + (snr_hi - snr_lo) * (float)cdm::u01_open(ws.v[0]);
    c.v[2] = (uint32_t)k;
    const cdm::U32x4 w = cdm::philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float u0 = (float)cdm::u01_open(w.v[0]), u1 = (float)cdm::u01_open(w.v[1]);
    const float rad = sqrtf(-2.0f * logf(u0));
    float sn, cs;
    sincosf(6.2831853071795865f * u1, &sn, &cs);
    const float sigma = sc / snr;
    const float re = clean[(size_t)s * nmeas + k] + sigma * rad * cs, im = sigma * rad * sn;
    val = sqrtf(re * re + im * im) / sc;
  }
  const bf16 h = (bf16)val;
  x[(size_t)v * kpad + k] = h;
  xt[(size_t)k * rows + v] = h;
}

// ------------------------------------------------------------------ head
struct Head {
  int n;
  cdm_head_entry e[kMaxHead];
  int out_index[kMaxHead];  // first network output of each entry
};

// Target encoding of the true parameters: per sample, n_out values.
__global__ void encode_kernel(int n, int nparams, const float* __restrict__ params, Head h, int n_out,
                              float* __restrict__ target) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const float* p = params + (size_t)v * nparams;
  float* t = target + (size_t)v * n_out;
  for (int j = 0; j < h.n; ++j) {
    const cdm_head_entry& e = h.e[j];
    const int o = h.out_index[j];
    const float x = p[e.param];
    if (e.kind == CDM_HEAD_AXIS) {
      const float th = x, ph = p[e.param2];
      t[o] = sinf(th) * cosf(ph);
      t[o + 1] = sinf(th) * sinf(ph);
      t[o + 2] = cosf(th);
    } else {
      float u;
      if (e.kind == CDM_HEAD_SIGMOID) u = (x - e.lo) / (e.hi - e.lo);
      else if (e.kind == CDM_HEAD_SIGMOID_LOG) u = (logf(x) - logf(e.lo)) / (logf(e.hi) - logf(e.lo));
      else u = (x / p[e.param2] - e.lo) / (e.hi - e.lo);
      t[o] = fminf(fmaxf(u, 0.0f), 1.0f);
    }
  }
}

// Loss and its gradient with respect to the raw outputs (times grad_scale,
// the inverse batch size), written as BF16 in both layouts (padding columns
// zero); loss_sum accumulates the loss times loss_scale.  One thread per
// sample, one wave per block (the head's outputs [0, n_out) are covered by
// its entries in order; each gradient value is stored as it is formed).
constexpr int kLossBlock = 64;
__global__ void __launch_bounds__(kLossBlock) loss_kernel(int rows, int n_out, int npad, const float* __restrict__ out,
                                                          const float* __restrict__ target, Head h, float grad_scale,
                                                          float loss_scale, bf16* dout, bf16* doutt, float* loss_sum,
                                                          const StepArgs* tab, const int* ctr) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (tab) target += (size_t)tab[*ctr].first * n_out;
  float l = 0.0f;
  if (v < rows) {
    const float* o = out + (size_t)v * npad;
    const float* t = target + (size_t)v * n_out;
    bf16* d = dout + (size_t)v * npad;
    auto put = [&](int k, float g) {
      const bf16 q = (bf16)(g * grad_scale);
      d[k] = q;
      doutt[(size_t)k * rows + v] = q;
    };
    for (int j = 0; j < h.n; ++j) {
      const cdm_head_entry& e = h.e[j];
      const int a = h.out_index[j];
      if (e.kind == CDM_HEAD_AXIS) {
        const float vx = o[a], vy = o[a + 1], vz = o[a + 2];
        const float q = vx * t[a] + vy * t[a + 1] + vz * t[a + 2];
        const float rr = vx * vx + vy * vy + vz * vz + 1e-6f;
        l += e.weight * (1.0f - q * q / rr);
        const float c1 = -2.0f * e.weight * q / rr, c2 = q / rr;
        put(a, c1 * (t[a] - c2 * vx));
        put(a + 1, c1 * (t[a + 1] - c2 * vy));
        put(a + 2, c1 * (t[a + 2] - c2 * vz));
      } else {
        const float s = sigmoid(o[a]), tt = t[a];
        const float sc = fminf(fmaxf(s, 1e-7f), 1.0f - 1e-7f);
        // soft-target cross entropy minus its minimum (0 at s = t)
        const float tc = fminf(fmaxf(tt, 1e-7f), 1.0f - 1e-7f);
        l += e.weight * (tc * logf(tc / sc) + (1.0f - tc) * logf((1.0f - tc) / (1.0f - sc)));
        put(a, e.weight * (s - tt));
      }
    }
    for (int k = n_out; k < npad; ++k) put(k, 0.0f);
  }
  // wave sum of the loss, one atomic per block (global memory, vector path)
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if (threadIdx.x == 0) atomicAdd(loss_sum, l * loss_scale);
}

// Raw outputs -> model parameters, entries in order (a ratio entry reads the
// already-decoded parameter it refers to).  Parameters without an entry keep
// their value.
__global__ void decode_kernel(int n, int npad, const float* __restrict__ out, Head h, int nparams,
                              float* __restrict__ params) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const float* o = out + (size_t)v * npad;
  float* p = params + (size_t)v * nparams;
  for (int j = 0; j < h.n; ++j) {
    const cdm_head_entry& e = h.e[j];
    const int a = h.out_index[j];
    if (e.kind == CDM_HEAD_AXIS) {
      const float vx = o[a], vy = o[a + 1], vz = o[a + 2];
      const float r = sqrtf(vx * vx + vy * vy + vz * vz);
      if (r > 0.0f) {
        p[e.param] = acosf(fminf(fmaxf(vz / r, -1.0f), 1.0f));
        p[e.param2] = atan2f(vy, vx);
      }
    } else {
      const float s = sigmoid(o[a]);
      float x;
      if (e.kind == CDM_HEAD_SIGMOID) x = e.lo + (e.hi - e.lo) * s;
      else if (e.kind == CDM_HEAD_SIGMOID_LOG) x = expf(logf(e.lo) + (logf(e.hi) - logf(e.lo)) * s);
      else x = (e.lo + (e.hi - e.lo) * s) * p[e.param2];
      p[e.param] = fminf(fmaxf(x, e.vmin), e.vmax);
    }
  }
}

// Adam over weight matrices [nrow][ncol] (padded storage, live region
// [rows_live][cols_live]) and bias vectors (ncol = 1): FP32 master, moments,
// and the BF16 copies of the matrices in both layouts.
struct AdamTensor {
  int nrow, ncol, rows_live, cols_live;
  float *w, *m, *v;
  bf16 *wbf, *wbft;
};
struct AdamArgs {
  float lr, b1, b2, eps, c1, c2;
};

// Element i's weight and moments, loaded ahead of its gradient.
struct AdamElem {
  float w, m, v;
};
__device__ __forceinline__ AdamElem adam_load(const AdamTensor& T, int i) { return {T.w[i], T.m[i], T.v[i]}; }

// Element i of tensor T (weight and moments x) with gradient gi; padding
// stays zero.  The moment updates name their fused multiply-adds, so every
// kernel that inlines this rounds them alike.
__device__ __forceinline__ void adam_update(const AdamTensor& T, int i, float gi, AdamElem x, const AdamArgs& a) {
  const int r = i / T.ncol, c = i - r * T.ncol;
  if (r >= T.rows_live || c >= T.cols_live) return;
  const float mi = fmaf(a.b1, x.m, (1.0f - a.b1) * gi);
  const float vi = fmaf(a.b2, x.v, (1.0f - a.b2) * gi * gi);
  T.m[i] = mi;
  T.v[i] = vi;
  const float wi = x.w - a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
  T.w[i] = wi;
  if (T.wbf) {
    const bf16 h = (bf16)wi;
    T.wbf[i] = h;
    T.wbft[(size_t)c * T.nrow + r] = h;
  }
}

// A replayed graph takes lr and the bias corrections from the step table at
// the entry behind the step index (advanced by the weight-gradient launch
// before it).
__device__ __forceinline__ AdamArgs adam_args(AdamArgs a, const StepArgs* tab, const int* ctr) {
  if (tab) {
    const StepArgs& t = tab[*ctr - 1];
    a.lr = t.lr;
    a.c1 = t.c1;
    a.c2 = t.c2;
  }
  return a;
}

// The weight gradients' second stage and the bias gradients of every layer in
// one launch.  Layer p owns the blocks [first, first + nb + units): blocks
// [0, nb) of it sum the S partial products in order, gw[i] = sum_z
// part[z][i] (nb = ceil(n / 256)); block nb + u forms db[u] = sum_b
// dzt[u][b] (rows a multiple of 8: each thread sums eight-element vectors,
// then the block reduces in a fixed order).  With tw.w set, each gradient
// element also takes its Adam step there (tw: the layer's weights, tb: its
// bias), with lr and the corrections from tab[*ctr - 1] when tab is set.
struct FinishLayer {
  int n, units, nb, first;
  const float* part;
  float* gw;
  const bf16* dzt;
  float* db;
  AdamTensor tw, tb;
};
struct FinishSet {
  int nl, S, rows;
  FinishLayer L[kMaxLayers];
};

__global__ void __launch_bounds__(256) wgrad_finish_kernel(FinishSet fs, AdamArgs aa, const StepArgs* tab,
                                                           const int* ctr) {
  int p = 0;
  while (p + 1 < fs.nl && (int)blockIdx.x >= fs.L[p + 1].first) ++p;
  const FinishLayer& F = fs.L[p];
  const int blk = blockIdx.x - F.first;
  const AdamArgs a = F.tw.w ? adam_args(aa, tab, ctr) : aa;
  if (blk < F.nb) {
    const int i = blk * 256 + threadIdx.x;
    if (i < F.n) {
      AdamElem x = {0.0f, 0.0f, 0.0f};
      if (F.tw.w) x = adam_load(F.tw, i);
      float acc = F.part[i];
      for (int z = 1; z < fs.S; ++z) acc += F.part[(size_t)z * F.n + i];
      F.gw[i] = acc;
      if (F.tw.w) adam_update(F.tw, i, acc, x, a);
    }
  } else {
    const int unit = blk - F.nb;
    float s = 0.0f;
    if (unit < F.units) {
      const bf16* row = F.dzt + (size_t)unit * fs.rows;
      for (int b = 8 * threadIdx.x; b < fs.rows; b += 8 * 256) {
        const bf16x8 q = *(const bf16x8*)(row + b);
#pragma unroll
        for (int j = 0; j < 8; ++j) s += (float)q[j];
      }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0 && unit < F.units) {
      const float g = (red[0] + red[1]) + (red[2] + red[3]);
      F.db[unit] = g;
      if (F.tw.w) adam_update(F.tb, unit, g, adam_load(F.tb, unit), a);
    }
  }
}

__global__ void to_bf16_kernel(int nrow, int ncol, const float* w, bf16* wbf, bf16* wbft) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrow * ncol) return;
  const int r = i / ncol, c = i - r * ncol;
  const bf16 h = (bf16)w[i];
  wbf[i] = h;
  wbft[(size_t)c * nrow + r] = h;
}

dim3 grid1(long long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

// ------------------------------------------------------------------ handle
struct cdm_mlp {
  int n_in = 0, width = 0, depth = 0, n_out = 0;
  int kpad = 0, npad = 0;  // padded input width and output width
  Head head;
  int nlayers = 0;
  // per weight layer: out x in (padded), live sizes
  int rows[kMaxLayers], cols[kMaxLayers], rows_live[kMaxLayers], cols_live[kMaxLayers];
  float *w[kMaxLayers] = {}, *b[kMaxLayers] = {}, *gw[kMaxLayers] = {}, *gb[kMaxLayers] = {};
  float *mw[kMaxLayers] = {}, *vw[kMaxLayers] = {}, *mb[kMaxLayers] = {}, *vb[kMaxLayers] = {};
  bf16 *wbf[kMaxLayers] = {}, *wbft[kMaxLayers] = {};
  long long adam_step = 0;
  // workspaces (grown on demand)
  std::vector<void*> ws;
  int train_rows = 0, infer_rows = 0;
  bf16 *x = nullptr, *xt = nullptr;
  float* z[kMaxLayers] = {};
  bf16 *h[kMaxLayers] = {}, *ht[kMaxLayers] = {};
  float* out = nullptr;
  bf16 *dout = nullptr, *doutt = nullptr, *dz[2] = {};
  bf16* dzt[kMaxLayers] = {};  // per hidden layer: dZ^T, kept for the weight gradients
  float* loss_sum = nullptr;
  bf16 *ix = nullptr, *ih[2] = {};
  float* iout = nullptr;
  float* wpart = nullptr;  // split-K partial weight gradients (kWgradSplits x every layer)
  // training steps replayed from one captured graph (the step table and its
  // index), or launched one by one
  bool train_graph = true;
  hipStream_t capture = nullptr;
};

namespace {

int hip_err(hipError_t e, const char* what) {
  return cdm::set_error(std::string(what) + ": " + hipGetErrorString(e), -2);
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

void free_all(cdm_mlp* m) {
  for (int l = 0; l < kMaxLayers; ++l) {
    for (float* p : {m->w[l], m->b[l], m->gw[l], m->gb[l], m->mw[l], m->vw[l], m->mb[l], m->vb[l], m->z[l]})
      if (p) (void)hipFree(p);
    for (bf16* p : {m->wbf[l], m->wbft[l], m->h[l], m->ht[l], m->dzt[l]})
      if (p) (void)hipFree(p);
  }
  for (void* p : {(void*)m->x, (void*)m->xt, (void*)m->out, (void*)m->dout, (void*)m->doutt, (void*)m->dz[0],
                  (void*)m->dz[1], (void*)m->loss_sum, (void*)m->ix,
                  (void*)m->ih[0], (void*)m->ih[1], (void*)m->iout, (void*)m->wpart})
    if (p) (void)hipFree(p);
}

// splitmix64 -> uniform in [-1, 1)
struct HostRng {
  uint64_t s;
  double next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
};

int upload_layer(cdm_mlp* m, int l, const std::vector<float>& wpad, const std::vector<float>& bpad) {
  hipError_t e;
  const size_t nw = (size_t)m->rows[l] * m->cols[l];
  if ((e = hipMemcpy(m->w[l], wpad.data(), nw * 4, hipMemcpyHostToDevice)) != hipSuccess) return hip_err(e, "weights");
  if ((e = hipMemcpy(m->b[l], bpad.data(), m->rows[l] * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return hip_err(e, "weights");
  hipLaunchKernelGGL(to_bf16_kernel, grid1(nw), dim3(256), 0, 0, m->rows[l], m->cols[l], m->w[l], m->wbf[l],
                     m->wbft[l]);
  if ((e = hipGetLastError()) != hipSuccess) return hip_err(e, "weights");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_err(e, "weights");
  return 0;
}

int ensure_train(cdm_mlp* m, int rows) {
  if (m->train_rows >= rows) return 0;
  hipError_t e = hipSuccess;
  for (void* p : {(void*)m->x, (void*)m->xt, (void*)m->out, (void*)m->dout, (void*)m->doutt, (void*)m->dz[0],
                  (void*)m->dz[1]})
    if (p) (void)hipFree(p);
  for (int l = 0; l < kMaxLayers; ++l) {
    if (m->z[l]) (void)hipFree(m->z[l]);
    if (m->h[l]) (void)hipFree(m->h[l]);
    if (m->ht[l]) (void)hipFree(m->ht[l]);
    if (m->dzt[l]) (void)hipFree(m->dzt[l]);
    m->z[l] = nullptr;
    m->h[l] = m->ht[l] = m->dzt[l] = nullptr;
  }
  const size_t W = m->width;
  if (e == hipSuccess) e = dalloc(&m->x, (size_t)rows * m->kpad);
  if (e == hipSuccess) e = dalloc(&m->xt, (size_t)rows * m->kpad);
  for (int l = 0; l < m->depth && e == hipSuccess; ++l) {
    e = dalloc(&m->z[l], rows * W);
    if (e == hipSuccess) e = dalloc(&m->h[l], rows * W);
    if (e == hipSuccess) e = dalloc(&m->ht[l], rows * W);
  }
  if (e == hipSuccess) e = dalloc(&m->out, (size_t)rows * m->npad);
  if (e == hipSuccess) e = dalloc(&m->dout, (size_t)rows * m->npad);
  if (e == hipSuccess) e = dalloc(&m->doutt, (size_t)rows * m->npad);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = dalloc(&m->dz[i], rows * W);
  for (int l = 0; l < m->depth && e == hipSuccess; ++l) e = dalloc(&m->dzt[l], rows * W);
  if (!m->wpart && e == hipSuccess) {
    size_t all = 0;
    for (int l = 0; l <= m->depth; ++l) all += (size_t)m->rows[l] * m->cols[l];
    e = dalloc(&m->wpart, (size_t)kWgradSplits * all);
  }
  if (e != hipSuccess) {
    m->train_rows = 0;
    return hip_err(e, "network training workspace");
  }
  m->train_rows = rows;
  return 0;
}

int ensure_infer(cdm_mlp* m) {
  if (m->infer_rows) return 0;
  const int rows = kInferChunk;
  hipError_t e = dalloc(&m->ix, (size_t)rows * m->kpad);
  if (e == hipSuccess) e = dalloc(&m->ih[0], (size_t)rows * m->width);
  if (e == hipSuccess) e = dalloc(&m->ih[1], (size_t)rows * m->width);
  if (e == hipSuccess) e = dalloc(&m->iout, (size_t)rows * m->npad);
  if (e != hipSuccess) return hip_err(e, "network inference workspace");
  m->infer_rows = rows;
  return 0;
}

// The K range of one block picks the kernel: 2, 4, 8 or 16 MFMA steps are
// loaded whole, other lengths run in chunks.
template <int EPI>
void gemm(const bf16* A, int lda, const bf16* B, int ldb, int M, int N, int K, const Epi& e, hipStream_t s,
          int splits = 1) {
  const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile, splits);
  const int ks = K / splits;
  auto kernel = gemm_nt<EPI, 0>;
  if (ks == 32) kernel = gemm_nt<EPI, 2>;
  else if (ks == 64) kernel = gemm_nt<EPI, 4>;
  else if (ks == 128) kernel = gemm_nt<EPI, 8>;
  else if (ks == 256) kernel = gemm_nt<EPI, 16>;
  hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, A, lda, B, ldb, M, N, ks, e);
}



Epi epi0() {
  Epi e;
  std::memset(&e, 0, sizeof(e));
  return e;
}

// K slices of a weight gradient (K = the batch rows): a 256 x 256 layer is
// only 16 output tiles, so without slices 16 blocks would each run the whole
// batch.  The slices keep K a multiple of 16 and at least 64 rows each.
int wgrad_splits(int rows) {
  int sp = kWgradSplits;
  while (sp > 1 && (rows % (16 * sp) != 0 || rows / sp < 64)) sp /= 2;
  return sp;
}

// Adam's step for one training step: hyper-parameters and bias corrections,
// and, for a replayed graph, the step table, its index and the last-block
// counter of the kernel that advances the index.
struct AdamStep {
  AdamArgs a;
  const StepArgs* tab;
  int* ctr;
};

// dW_l = dZ_l^T H_{l-1} of every layer as K slices into m->wpart, in one
// launch, then their ordered sums and the bias gradients (sums over the
// batch of dZ_l^T) in another, which also takes every layer's Adam step when
// `step` is given (and advances the step index of a replayed graph).
void wgrad_all(cdm_mlp* m, int rows, const AdamStep* step, hipStream_t s) {
  const int sp = wgrad_splits(rows), ks = rows / sp;
  GemmJobs jobs;
  std::memset(&jobs, 0, sizeof(jobs));
  FinishSet fs;
  std::memset(&fs, 0, sizeof(fs));
  jobs.n = fs.nl = m->depth + 1;
  jobs.K = ks;
  fs.S = sp;
  fs.rows = rows;
  int gblocks = 0, fblocks = 0;
  size_t off = 0;
  for (int l = 0; l <= m->depth; ++l) {
    const int units = m->rows[l], cols = m->cols[l];
    const bf16* dzt = l == m->depth ? m->doutt : m->dzt[l];
    GemmJob& J = jobs.j[l];
    J.A = dzt;
    J.lda = rows;
    J.B = l == 0 ? m->xt : m->ht[l - 1];
    J.ldb = rows;
    J.M = units;
    J.N = cols;
    J.tiles_n = (cols + kTile - 1) / kTile;
    J.tiles_m = (units + kTile - 1) / kTile;
    J.first = gblocks;
    J.e = epi0();
    J.e.out_f32 = m->wpart + off;
    J.e.ld_f32 = cols;
    J.e.split_stride = (size_t)units * cols;
    gblocks += J.tiles_n * J.tiles_m * sp;
    FinishLayer& F = fs.L[l];
    F.n = units * cols;
    F.units = units;
    F.nb = (F.n + 255) / 256;
    F.first = fblocks;
    F.part = m->wpart + off;
    F.gw = m->gw[l];
    F.dzt = dzt;
    F.db = m->gb[l];
    if (step) {
      F.tw = {m->rows[l], m->cols[l], m->rows_live[l], m->cols_live[l], m->w[l], m->mw[l], m->vw[l], m->wbf[l],
              m->wbft[l]};
      F.tb = {m->rows[l], 1, m->rows_live[l], 1, m->b[l], m->mb[l], m->vb[l], nullptr, nullptr};
    }
    fblocks += F.nb + units;
    off += (size_t)sp * units * cols;
  }
  auto kernel = gemm_multi<0>;
  if (ks == 64) kernel = gemm_multi<4>;
  else if (ks == 128) kernel = gemm_multi<8>;
  else if (ks == 256) kernel = gemm_multi<16>;
  AdamArgs aa;
  std::memset(&aa, 0, sizeof(aa));
  const StepArgs* tab = nullptr;
  int* ctr = nullptr;
  if (step) {
    aa = step->a;
    tab = step->tab;
    ctr = step->ctr;
  }
  jobs.advance = tab ? ctr : nullptr;
  hipLaunchKernelGGL(kernel, dim3(gblocks), dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(wgrad_finish_kernel, dim3(fblocks), dim3(256), 0, s, fs, aa, tab, (const int*)ctr);
}

// Forward over `rows` (multiple of 32) rows of x.  Training keeps z, h, ht.
void forward_train(cdm_mlp* m, int rows, hipStream_t s) {
  const bf16* in = m->x;
  int ld_in = m->kpad, k = m->kpad;
  for (int l = 0; l < m->depth; ++l) {
    Epi e = epi0();
    e.bias = m->b[l];
    e.out_f32 = m->z[l];
    e.ld_f32 = m->width;
    e.out_bf = m->h[l];
    e.ld_bf = m->width;
    e.out_bft = m->ht[l];
    e.ld_bft = rows;
    gemm<EPI_FWD_GELU>(in, ld_in, m->wbf[l], m->cols[l], rows, m->width, k, e, s);
    in = m->h[l];
    ld_in = k = m->width;
  }
  Epi e = epi0();
  e.bias = m->b[m->depth];
  e.out_f32 = m->out;
  e.ld_f32 = m->npad;
  gemm<EPI_PLAIN>(in, ld_in, m->wbf[m->depth], m->width, rows, m->npad, k, e, s);
}

// Backward from dout/doutt: dZ_l = (dZ_{l+1} W_{l+1}) gelu'(Z_l) down the
// layers (the old weights), then every weight and bias gradient (gw, gb),
// with Adam's step on every layer when `step` is given.
void backward(cdm_mlp* m, int rows, hipStream_t s, const AdamStep* step = nullptr) {
  const int L = m->depth;
  const bf16* dcur = m->dout;
  int kcur = m->npad;
  for (int l = L - 1; l >= 0; --l) {
    const int slot = l & 1;
    Epi e = epi0();
    e.zin = m->z[l];
    e.ld_zin = m->width;
    e.out_bf = m->dz[slot];
    e.ld_bf = m->width;
    e.out_bft = m->dzt[l];
    e.ld_bft = rows;
    gemm<EPI_BWD_GELU>(dcur, kcur, m->wbft[l + 1], kcur, rows, m->width, kcur, e, s);
    dcur = m->dz[slot];
    kcur = m->width;
  }
  wgrad_all(m, rows, step, s);
}

constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;

// Adam's bias corrections at step t (1-based).
float adam_c1(long long t) { return 1.0f - std::pow(kAdamB1, (float)t); }
float adam_c2(long long t) { return 1.0f - std::pow(kAdamB2, (float)t); }

// The next Adam step's arguments; with tab (a replayed graph) lr and the
// corrections come from the step table at *ctr instead.
AdamStep adam_step_args(cdm_mlp* m, float lr, const StepArgs* tab, int* ctr) {
  ++m->adam_step;
  AdamStep st;
  st.a = {lr, kAdamB1, kAdamB2, kAdamEps, adam_c1(m->adam_step), adam_c2(m->adam_step)};
  st.tab = tab;
  st.ctr = ctr;
  return st;
}

// Inference over n voxels in chunks: raw outputs to out [n][npad] (device,
// caller's) or straight into the decode step.
int infer(cdm_mlp* m, int n, int nmeas, const float* meas, const float* scale, float* out_raw, int nparams,
          float* params, hipStream_t s) {
  int r = ensure_infer(m);
  if (r) return r;
  for (int first = 0; first < n; first += kInferChunk) {
    const int cnt = std::min(kInferChunk, n - first);
    const int rows = round_up(cnt, 32);
    hipLaunchKernelGGL(inputs_kernel, grid1((long long)rows * m->kpad), dim3(256), 0, s, cnt, rows, nmeas, m->kpad,
                       meas + (size_t)first * nmeas, scale ? scale + first : nullptr, m->ix, (bf16*)nullptr);
    const bf16* in = m->ix;
    int k = m->kpad;
    for (int l = 0; l < m->depth; ++l) {
      Epi e = epi0();
      e.bias = m->b[l];
      e.out_bf = m->ih[l & 1];
      e.ld_bf = m->width;
      gemm<EPI_FWD_GELU>(in, k, m->wbf[l], m->cols[l], rows, m->width, k, e, s);
      in = m->ih[l & 1];
      k = m->width;
    }
    Epi e = epi0();
    e.bias = m->b[m->depth];
    e.out_f32 = m->iout;
    e.ld_f32 = m->npad;
    gemm<EPI_PLAIN>(in, k, m->wbf[m->depth], m->width, rows, m->npad, k, e, s);
    if (out_raw) {
      hipError_t e2 = hipMemcpy2DAsync(out_raw + (size_t)first * m->n_out, m->n_out * sizeof(float), m->iout,
                                       m->npad * sizeof(float), m->n_out * sizeof(float), cnt, hipMemcpyDeviceToDevice,
                                       s);
      if (e2 != hipSuccess) return hip_err(e2, "network outputs");
    }
    if (params)
      hipLaunchKernelGGL(decode_kernel, grid1(cnt), dim3(256), 0, s, cnt, m->npad, m->iout, m->head, nparams,
                         params + (size_t)first * nparams);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_err(e, "network inference launch");
}

int check_params(const cdm_mlp* m, int nparams) {
  for (int j = 0; j < m->head.n; ++j) {
    const cdm_head_entry& e = m->head.e[j];
    if (e.param >= nparams || ((e.kind == CDM_HEAD_AXIS || e.kind == CDM_HEAD_SIGMOID_RATIO) && e.param2 >= nparams))
      return cdm::set_error("network head refers to a parameter beyond nparams");
  }
  return 0;
}

}  // namespace

extern "C" {

int cdm_mlp_create(int n_in, int width, int depth, const cdm_head_entry* head, int nhead, unsigned long long seed,
                   cdm_mlp** out) {
  if (!out) return cdm::set_error("null output handle");
  *out = nullptr;
  if (n_in <= 0 || width <= 0 || width % 32 != 0 || depth < 1 || depth > kMaxLayers - 1)
    return cdm::set_error("network shape: n_in > 0, width a positive multiple of 32, 1 <= depth <= 7");
  if (!head || nhead <= 0 || nhead > kMaxHead) return cdm::set_error("network head: 1..16 entries");
  cdm_mlp* m = new cdm_mlp();
  m->n_in = n_in;
  m->width = width;
  m->depth = depth;
  m->head.n = nhead;
  int n_out = 0;
  for (int j = 0; j < nhead; ++j) {
    const cdm_head_entry& e = head[j];
    if (e.kind < CDM_HEAD_SIGMOID || e.kind > CDM_HEAD_AXIS || e.param < 0 ||
        ((e.kind == CDM_HEAD_AXIS || e.kind == CDM_HEAD_SIGMOID_RATIO) && e.param2 < 0) ||
        (e.kind != CDM_HEAD_AXIS && !(e.hi > e.lo)) || (e.kind == CDM_HEAD_SIGMOID_LOG && !(e.lo > 0))) {
      delete m;
      return cdm::set_error("network head entry " + std::to_string(j) + " is invalid");
    }
    m->head.e[j] = e;
    m->head.out_index[j] = n_out;
    n_out += e.kind == CDM_HEAD_AXIS ? 3 : 1;
  }
  if (n_out > 32) {
    delete m;
    return cdm::set_error("network head: at most 32 outputs");
  }
  m->n_out = n_out;
  m->kpad = round_up(n_in, 32);
  m->npad = 32;
  m->nlayers = depth + 1;
  HostRng rng{seed};
  for (int l = 0; l <= depth; ++l) {
    m->rows[l] = l == depth ? m->npad : width;
    m->cols[l] = l == 0 ? m->kpad : width;
    m->rows_live[l] = l == depth ? n_out : width;
    m->cols_live[l] = l == 0 ? n_in : width;
    const size_t nw = (size_t)m->rows[l] * m->cols[l];
    hipError_t e = hipSuccess;
    for (float** p : {&m->w[l], &m->gw[l], &m->mw[l], &m->vw[l]})
      if (e == hipSuccess) e = dalloc(p, nw);
    for (float** p : {&m->b[l], &m->gb[l], &m->mb[l], &m->vb[l]})
      if (e == hipSuccess) e = dalloc(p, m->rows[l]);
    if (e == hipSuccess) e = dalloc(&m->wbf[l], nw);
    if (e == hipSuccess) e = dalloc(&m->wbft[l], nw);
    for (float* p : {m->mw[l], m->vw[l]})
      if (e == hipSuccess) e = hipMemset(p, 0, nw * 4);
    for (float* p : {m->mb[l], m->vb[l], m->gb[l]})
      if (e == hipSuccess) e = hipMemset(p, 0, m->rows[l] * 4);
    if (e != hipSuccess) {
      free_all(m);
      delete m;
      return hip_err(e, "network weights");
    }
    // LeCun-uniform initialisation: variance 1 / fan_in
    std::vector<float> w(nw, 0.0f), b(m->rows[l], 0.0f);
    const double a = std::sqrt(3.0 / m->cols_live[l]);
    for (int r = 0; r < m->rows_live[l]; ++r)
      for (int c = 0; c < m->cols_live[l]; ++c) w[(size_t)r * m->cols[l] + c] = (float)(a * rng.next());
    const int rc = upload_layer(m, l, w, b);
    if (rc) {
      free_all(m);
      delete m;
      return rc;
    }
  }
  *out = m;
  return 0;
}

void cdm_mlp_destroy(cdm_mlp* m) {
  if (!m) return;
  (void)hipDeviceSynchronize();
  free_all(m);
  if (m->capture) (void)hipStreamDestroy(m->capture);
  delete m;
}

int cdm_mlp_set_train_graph(cdm_mlp* m, int use_graph) {
  if (!m) return cdm::set_error("null network handle");
  m->train_graph = use_graph != 0;
  return 0;
}

int cdm_mlp_info(const cdm_mlp* m, int* n_in, int* width, int* depth, int* n_out, long long* nweights) {
  if (!m) return cdm::set_error("null network handle");
  long long nw = 0;
  for (int l = 0; l <= m->depth; ++l) nw += (long long)m->rows_live[l] * m->cols_live[l] + m->rows_live[l];
  if (n_in) *n_in = m->n_in;
  if (width) *width = m->width;
  if (depth) *depth = m->depth;
  if (n_out) *n_out = m->n_out;
  if (nweights) *nweights = nw;
  return 0;
}

// Flat host layout: for each layer, W[out][in] (live part) then b[out].
int cdm_mlp_get_weights(const cdm_mlp* m, float* w) {
  if (!m || !w) return cdm::set_error("null argument");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_err(e, "network weights");
  for (int l = 0; l <= m->depth; ++l) {
    std::vector<float> wp((size_t)m->rows[l] * m->cols[l]), bp(m->rows[l]);
    if ((e = hipMemcpy(wp.data(), m->w[l], wp.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess)
      return hip_err(e, "network weights");
    if ((e = hipMemcpy(bp.data(), m->b[l], bp.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess)
      return hip_err(e, "network weights");
    for (int r = 0; r < m->rows_live[l]; ++r)
      for (int c = 0; c < m->cols_live[l]; ++c) *w++ = wp[(size_t)r * m->cols[l] + c];
    for (int r = 0; r < m->rows_live[l]; ++r) *w++ = bp[r];
  }
  return 0;
}

int cdm_mlp_set_weights(cdm_mlp* m, const float* w) {
  if (!m || !w) return cdm::set_error("null argument");
  for (int l = 0; l <= m->depth; ++l) {
    std::vector<float> wp((size_t)m->rows[l] * m->cols[l], 0.0f), bp(m->rows[l], 0.0f);
    for (int r = 0; r < m->rows_live[l]; ++r)
      for (int c = 0; c < m->cols_live[l]; ++c) wp[(size_t)r * m->cols[l] + c] = *w++;
    for (int r = 0; r < m->rows_live[l]; ++r) bp[r] = *w++;
    const int rc = upload_layer(m, l, wp, bp);
    if (rc) return rc;
  }
  m->adam_step = 0;
  for (int l = 0; l <= m->depth; ++l) {
    const size_t nw = (size_t)m->rows[l] * m->cols[l];
    hipError_t e = hipSuccess;
    for (float* p : {m->mw[l], m->vw[l]})
      if (e == hipSuccess) e = hipMemset(p, 0, nw * 4);
    for (float* p : {m->mb[l], m->vb[l]})
      if (e == hipSuccess) e = hipMemset(p, 0, m->rows[l] * 4);
    if (e != hipSuccess) return hip_err(e, "network optimiser state");
  }
  return 0;
}

int cdm_mlp_forward(cdm_mlp* m, int n, int nmeas, const float* meas, const float* scale, float* out, void* stream) {
  if (!m) return cdm::set_error("null network handle");
  if (n < 0 || nmeas != m->n_in) return cdm::set_error("nmeas must equal the network's input width");
  if (n == 0) return 0;
  if (!meas || !out) return cdm::set_error("null device pointer");
  return infer(m, n, nmeas, meas, scale, out, 0, nullptr, (hipStream_t)stream);
}

int cdm_mlp_init_params(cdm_mlp* m, int n, int nmeas, const float* meas, const float* scale, int nparams,
                        float* params, void* stream) {
  if (!m) return cdm::set_error("null network handle");
  if (n < 0 || nmeas != m->n_in) return cdm::set_error("nmeas must equal the network's input width");
  int r = check_params(m, nparams);
  if (r) return r;
  if (n == 0) return 0;
  if (!meas || !params) return cdm::set_error("null device pointer");
  return infer(m, n, nmeas, meas, scale, nullptr, nparams, params, (hipStream_t)stream);
}

int cdm_mlp_loss_grad(cdm_mlp* m, int n, int nmeas, const float* meas, const float* scale, int nparams,
                      const float* params_true, float* loss, float* grad, void* stream) {
  if (!m) return cdm::set_error("null network handle");
  if (n <= 0 || n % 32 != 0 || nmeas != m->n_in)
    return cdm::set_error("loss_grad: n a positive multiple of 32, nmeas equal to the input width");
  int r = check_params(m, nparams);
  if (r) return r;
  if (!meas || !params_true || !loss || !grad) return cdm::set_error("null pointer");
  hipStream_t s = (hipStream_t)stream;
  if ((r = ensure_train(m, n))) return r;
  float* target = nullptr;
  hipError_t e = dalloc(&target, (size_t)n * m->n_out);
  if (e == hipSuccess && !m->loss_sum) e = dalloc(&m->loss_sum, 1);
  if (e != hipSuccess) return hip_err(e, "loss_grad workspace");
  hipLaunchKernelGGL(encode_kernel, grid1(n), dim3(256), 0, s, n, nparams, params_true, m->head, m->n_out, target);
  hipLaunchKernelGGL(inputs_kernel, grid1((long long)n * m->kpad), dim3(256), 0, s, n, n, nmeas, m->kpad, meas, scale,
                     m->x, m->xt);
  forward_train(m, n, s);
  (void)hipMemsetAsync(m->loss_sum, 0, sizeof(float), s);
  hipLaunchKernelGGL(loss_kernel, grid1(n, kLossBlock), dim3(kLossBlock), 0, s, n, m->n_out, m->npad, m->out, target, m->head,
                     1.0f / n, 1.0f / n, m->dout, m->doutt, m->loss_sum, (const StepArgs*)nullptr, (const int*)nullptr);
  backward(m, n, s);
  e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(loss, m->loss_sum, sizeof(float), hipMemcpyDeviceToHost);
  for (int l = 0; l <= m->depth && e == hipSuccess; ++l) {
    std::vector<float> gp((size_t)m->rows[l] * m->cols[l]), bp(m->rows[l]);
    e = hipMemcpy(gp.data(), m->gw[l], gp.size() * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(bp.data(), m->gb[l], bp.size() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) break;
    for (int rr = 0; rr < m->rows_live[l]; ++rr)
      for (int c = 0; c < m->cols_live[l]; ++c) *grad++ = gp[(size_t)rr * m->cols[l] + c];
    for (int rr = 0; rr < m->rows_live[l]; ++rr) *grad++ = bp[rr];
  }
  (void)hipFree(target);
  return e == hipSuccess ? 0 : hip_err(e, "loss_grad");
}

int cdm_mlp_train(cdm_mlp* m, int n, int nmeas, const float* clean, const float* scale, int nparams,
                  const float* params_true, float snr_lo, float snr_hi, int epochs, int batch, float lr,
                  unsigned long long seed, float* epoch_loss, void* stream) {
  if (!m) return cdm::set_error("null network handle");
  if (nmeas != m->n_in) return cdm::set_error("nmeas must equal the network's input width");
  if (batch <= 0 || batch % 32 != 0) return cdm::set_error("batch must be a positive multiple of 32");
  if (n < batch) return cdm::set_error("fewer training samples than one batch");
  if (epochs < 0 || !(snr_hi >= snr_lo) || !(snr_lo > 0) || !(lr > 0)) return cdm::set_error("bad training options");
  int r = check_params(m, nparams);
  if (r) return r;
  if (!clean || !params_true) return cdm::set_error("null device pointer");
  hipStream_t s = (hipStream_t)stream;
  if ((r = ensure_train(m, batch))) return r;
  float* target = nullptr;
  hipError_t e = dalloc(&target, (size_t)n * m->n_out);
  if (e == hipSuccess && !m->loss_sum) e = dalloc(&m->loss_sum, 1);
  if (e != hipSuccess) return hip_err(e, "training workspace");
  hipLaunchKernelGGL(encode_kernel, grid1(n), dim3(256), 0, s, n, nparams, params_true, m->head, m->n_out, target);
  const int nb = n / batch;  // whole batches; the remainder rotates in through the epoch offset
  const long long total = (long long)nb * epochs;
  // every step's changing arguments, as the eager loop computes them
  std::vector<StepArgs> steps((size_t)total);
  for (long long st = 0; st < total; ++st) {
    const int ep = (int)(st / nb), bi = (int)(st % nb);
    const long long shift = (long long)(ep * 7919LL % std::max(1, n - nb * batch + 1));
    const float cosw = 0.5f * (1.0f + std::cos(3.14159265f * (float)st / (float)std::max(1LL, total)));
    StepArgs& a = steps[(size_t)st];
    a.first = shift + (long long)bi * batch;
    a.ep = (unsigned)ep;
    a.lr = lr * (0.02f + 0.98f * cosw);
    a.c1 = adam_c1(m->adam_step + st + 1);
    a.c2 = adam_c2(m->adam_step + st + 1);
  }
  auto step_kernels = [&](hipStream_t q, const StepArgs& a, const StepArgs* tab, int* ctr) {
    hipLaunchKernelGGL(noisy_inputs_kernel, grid1((long long)batch * m->kpad), dim3(256), 0, q, batch, nmeas,
                       m->kpad, a.first, clean, scale, snr_lo, snr_hi, (uint64_t)seed, (uint32_t)a.ep, m->x, m->xt,
                       tab, ctr);
    forward_train(m, batch, q);
    hipLaunchKernelGGL(loss_kernel, grid1(batch, kLossBlock), dim3(kLossBlock), 0, q, batch, m->n_out, m->npad, m->out,
                       tab ? target : target + (size_t)a.first * m->n_out, m->head, 1.0f / batch,
                       1.0f / (batch * (float)nb), m->dout, m->doutt, m->loss_sum, tab, ctr);
    const AdamStep st = adam_step_args(m, a.lr, tab, ctr);
    backward(m, batch, q, &st);
  };
  // Graph: one step captured with its changing arguments read from the
  // device step table at *ctr, replayed once per step; its weight-gradient
  // launch advances *ctr (the Adam step after it reads the entry behind).
  // The table, the index and the per-epoch loss reset are written outside
  // the graph.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  StepArgs* d_tab = nullptr;
  int* d_ctr = nullptr;
  if (m->train_graph && total > 0) {
    e = dalloc(&d_tab, (size_t)total);
    if (e == hipSuccess) e = dalloc(&d_ctr, 1);
    if (e == hipSuccess) e = hipMemcpyAsync(d_tab, steps.data(), steps.size() * sizeof(StepArgs), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(d_ctr, 0, sizeof(int), s);
    if (e == hipSuccess && !m->capture) e = hipStreamCreateWithFlags(&m->capture, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(m->capture, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const long long step0 = m->adam_step;
      step_kernels(m->capture, steps[0], d_tab, d_ctr);
      m->adam_step = step0;  // the replays advance it below
      const hipError_t ec = hipStreamEndCapture(m->capture, &graph);
      e = e == hipSuccess ? ec : e;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  }
  long long step = 0;
  for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    e = hipMemsetAsync(m->loss_sum, 0, sizeof(float), s);
    for (int bi = 0; bi < nb && e == hipSuccess; ++bi, ++step) {
      if (exec) {
        e = hipGraphLaunch(exec, s);
        ++m->adam_step;
      } else {
        step_kernels(s, steps[(size_t)step], nullptr, nullptr);
      }
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && epoch_loss) e = hipMemcpyAsync(&epoch_loss[ep], m->loss_sum, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipStreamSynchronize(s);
  if (d_tab) (void)hipFree(d_tab);
  if (d_ctr) (void)hipFree(d_ctr);
  (void)hipStreamSynchronize(s);
  (void)hipFree(target);
  return e == hipSuccess ? 0 : hip_err(e, "network training");
}

}  // extern "C"
