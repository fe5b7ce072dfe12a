// lh_mlp.hip -- the learned subfilter closure: a small MLP trained and evaluated on gfx950
// with BF16 matrix cores (v_mfma_f32_16x16x32_bf16 and v_mfma_f32_16x16x16_bf16) accumulating in FP32.
//
// The reference has no counterpart: the closure is the project's canopy-layer subfilter model
// (inputs: filtered velocity gradients on a 3x3x3 stencil and the solid fraction; outputs:
// the six subfilter stresses).  Layers: nin -> nhid (tanh) -> ... -> nhid (tanh) -> nout
// (linear), `nlayers` hidden layers.  FP32 master weights and Adam moments; BF16 copies of
// the weights and of the activations feed the matrix cores; pre-activations, gradients of
// the activations and the weight gradients are FP32.
//
// Every matrix is stored padded to multiples of 16 (pad entries are zero and stay zero: their
// gradients are zero).  One wave computes one 16x16 output tile:
//   lane l holds A[i][k] for i = l % 16, k = 4 (l / 16) + q and B[k][j] for j = l % 16,
//   k = 4 (l / 16) + q (q = 0..3), and receives D[i][j] for j = l % 16, i = 4 (l / 16) + q.
// The forward products step through k 32 at a time with v_mfma_f32_16x16x32_bf16 (gfx950),
// whose lane l holds the eight k = 8 (l / 16) + 0..7 of its row (A) or column (B).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/littleharsh_amd.h"
#include "lh_internal.hpp"

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s8 __attribute__((ext_vector_type(8)));

// eight contiguous bf16 (16 bytes) as a v_mfma_f32_16x16x32_bf16 operand
__device__ inline bf8 ld8(const uint16_t* p) { return __builtin_bit_cast(bf8, *reinterpret_cast<const s8*>(p)); }
// 32-wide step: lane l holds A[i = l % 16][k = 8 (l / 16) + 0..7] and B[k = 8 (l / 16) + 0..7]
// [j = l % 16]; D as in the 16-wide instruction
__device__ inline f4 mfma32(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __host__ inline float bf2f(uint16_t b) {
  union { float f; uint32_t u; } c;
  c.u = (uint32_t)b << 16;
  return c.f;
}
__host__ __device__ inline int pad16(int n) { return (n + 15) & ~15; }

// The activation, tanh, from the exponential: tanh|x| = (1 - t) / (1 + t) with t = exp(-2|x|)
// (hardware exp2 and reciprocal), and below |x| = 1/16, where 1 - t cancels, the Taylor series
// x - x^3/3 + 2 x^5/15 - 17 x^7/315 (truncation below 1e-12).  Relative error about 1e-6, far
// below the bf16 rounding the activations take next; the forward pass, its derivative and
// inference all use this one function.
__device__ inline float tanh_act(float x) {
  const float ax = fabsf(x);
  const float t = __expf(-2.f * ax);
  const float big = copysignf((1.f - t) * __builtin_amdgcn_rcpf(1.f + t), x);
  const float x2 = x * x;
  const float small = x + x * x2 * (-1.f / 3.f + x2 * (2.f / 15.f - x2 * (17.f / 315.f)));
  return ax < 0.0625f ? small : big;
}

// counter-based hash (SplitMix64 finaliser)
__device__ __host__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// ---- kernels (one 64-thread block = one wave = one 16x16 tile) -------------------------

// Z = X W^T + b (X: B x K bf16, W: N x K bf16, Z: B x N f32); Xn = tanh(Z) as bf16 when
// act != 0 (Xn has N columns)
__global__ void __launch_bounds__(64) k_fwd(const uint16_t* __restrict__ X, const uint16_t* __restrict__ W,
                                            const float* __restrict__ bias, float* __restrict__ Z,
                                            uint16_t* __restrict__ Xn, int K, int N, int act) {
  const int l = threadIdx.x;
  const int b0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  // a lane's k are contiguous (every width is a multiple of 16): 32 at a time, then a last
  // 16-wide step when K is an odd multiple of 16
  const uint16_t* xr = X + (size_t)(b0 + l % 16) * K;
  const uint16_t* wr = W + (size_t)(n0 + l % 16) * K;
  int k0 = 0;
  for (; k0 + 32 <= K; k0 += 32) acc = mfma32(ld8(xr + k0 + 8 * (l / 16)), ld8(wr + k0 + 8 * (l / 16)), acc);
  if (k0 < K)
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(*reinterpret_cast<const bf4*>(xr + k0 + 4 * (l / 16)),
                                                    *reinterpret_cast<const bf4*>(wr + k0 + 4 * (l / 16)), acc,
                                                    0, 0, 0);
  const int n = n0 + l % 16;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = b0 + 4 * (l / 16) + q;
    const float z = acc[q] + bias[n];
    Z[(size_t)r * N + n] = z;
    if (act) Xn[(size_t)r * N + n] = f2bf(tanh_act(z));
  }
}

// dW = dZ^T X (dZ: B x N bf16, X: B x K bf16, dW: N x K f32), summed over the B rows; the
// blocks of the extra last row (blockIdx.y = K / 16) form db = the column sums of the FP32 dZf
// instead, beside the products.
// One block of kDwWaves waves per 16x16 tile: wave w takes the row tiles w, w + kDwWaves, ...;
// the partial tiles are summed in LDS in wave order (deterministic).
constexpr int kDwWaves = 16;
__device__ __forceinline__ void dw_tile(const uint16_t* __restrict__ dZ, const uint16_t* __restrict__ X,
                                        float* __restrict__ dW, int B, int K, int N, const float* __restrict__ dZf,
                                        float* __restrict__ db, int bx, int by) {
  __shared__ float part[kDwWaves][256];
  __shared__ float dbp[kDwWaves][64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n0 = bx * 16, k0 = by * 16;
  if (k0 >= K) {                                    // bias row: column sums of dZf
    float sb = 0.f;
    for (int r0 = 16 * w; r0 < B; r0 += 16 * kDwWaves)
#pragma unroll
      for (int q = 0; q < 4; ++q) sb += dZf[(size_t)(r0 + 4 * (l / 16) + q) * N + n0 + l % 16];
    dbp[w][l] = sb;
    __syncthreads();
    if (threadIdx.x < 16) {
      float s = 0.f;
      for (int v = 0; v < kDwWaves; ++v)
        for (int gq = 0; gq < 4; ++gq) s += dbp[v][gq * 16 + threadIdx.x];
      db[n0 + threadIdx.x] = s;
    }
    return;
  }
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int r0 = 16 * w; r0 < B; r0 += 16 * kDwWaves) {
    bf4 a, b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + 4 * (l / 16) + q;
      a[q] = (short)dZ[(size_t)r * N + n0 + l % 16];   // A[i = neuron][k = row]
      b[q] = (short)X[(size_t)r * K + k0 + l % 16];    // B[k = row][j = feature]
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) part[w][(4 * (l / 16) + q) * 16 + l % 16] = acc[q];
  __syncthreads();
  if (threadIdx.x < 256) {
    float s = 0.f;
    for (int v = 0; v < kDwWaves; ++v) s += part[v][threadIdx.x];
    dW[(size_t)(n0 + threadIdx.x / 16) * K + k0 + threadIdx.x % 16] = s;
  }
}
__global__ void __launch_bounds__(64 * kDwWaves) k_dw(const uint16_t* __restrict__ dZ,
                                                      const uint16_t* __restrict__ X,
                                                      float* __restrict__ dW, int B, int K, int N,
                                                      const float* __restrict__ dZf, float* __restrict__ db) {
  dw_tile(dZ, X, dW, B, K, N, dZf, db, blockIdx.x, blockIdx.y);
}

// The weight gradients of every layer in one launch: block b of layer l is the block
// (b % (N / 16), b / (N / 16)) of that layer's k_dw grid.  With tstep, block 0 advances the
// step counter (the step's gather has finished and no block here reads it).
constexpr int kMaxL = 8;
struct DwAll {
  const uint16_t* dZ[kMaxL];
  const uint16_t* X[kMaxL];
  float* dW[kMaxL];
  const float* dZf[kMaxL];
  float* db[kMaxL];
  int K[kMaxL], N[kMaxL], first[kMaxL + 1];   // first block of each layer
  int L, B;
  int* tstep;
};
__global__ void __launch_bounds__(64 * kDwWaves) k_dw_all(DwAll a) {
  int l = 0;
  while (l + 1 < a.L && (int)blockIdx.x >= a.first[l + 1]) ++l;
  const int b = blockIdx.x - a.first[l], nbx = a.N[l] / 16;
  dw_tile(a.dZ[l], a.X[l], a.dW[l], a.B, a.K[l], a.N[l], a.dZf[l], a.db[l], b % nbx, b / nbx);
  if (a.tstep && blockIdx.x == 0 && threadIdx.x == 0) a.tstep[0] = a.tstep[0] + 1;
}

// the previous layer's dZ = (dZ W) * (1 - tanh(Zp)^2) (dZ: B x N bf16, W: N x K bf16, Zp: its
// pre-activations, B x K f32) -> dZo as f32 and bf16
__global__ void __launch_bounds__(64) k_dx(const uint16_t* __restrict__ dZ, const uint16_t* __restrict__ W,
                                           const float* __restrict__ Zp, float* __restrict__ dZo,
                                           uint16_t* __restrict__ dZbo, int K, int N) {
  const int l = threadIdx.x;
  const int b0 = blockIdx.x * 16, k0 = blockIdx.y * 16;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const bf4* za = reinterpret_cast<const bf4*>(dZ + (size_t)(b0 + l % 16) * N + 4 * (l / 16));
  for (int n0 = 0; n0 < N; n0 += 16) {
    const bf4 a = za[n0 / 4];                            // A[i = row][k = neuron], contiguous
    bf4 b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int n = n0 + 4 * (l / 16) + q;
      b[q] = (short)W[(size_t)n * K + k0 + l % 16];      // B[k = neuron][j = feature]
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const size_t o = (size_t)(b0 + 4 * (l / 16) + q) * K + k0 + l % 16;
    const float th = tanh_act(Zp[o]);
    const float gq = acc[q] * (1.f - th * th);
    dZo[o] = gq;
    dZbo[o] = f2bf(gq);
  }
}

// squared error of the `nout` real outputs of B rows: dZ of the output layer,
// dY = 2 (Y - T) / (B nout), as f32 and bf16, and one partial sum of (Y - T)^2 per 256-element
// block (fixed-order tree); with tstep, the first thread advances the step counter (every
// block of the step's gather has finished)
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             float* __restrict__ dZ, uint16_t* __restrict__ dZb,
                                             float* __restrict__ part, int* __restrict__ tstep, int B, int NP,
                                             int nout) {
  __shared__ float red[256];
  const int e = blockIdx.x * 256 + threadIdx.x;
  const float sc = 2.f / (float)(B * nout);
  float s = 0.f;
  if (e < B * NP) {
    const int r = e / NP, n = e % NP;
    float d = 0.f;
    if (n < nout) {
      d = Y[e] - T[(size_t)r * nout + n];
      s = d * d;
    }
    dZ[e] = sc * d;
    dZb[e] = f2bf(sc * d);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = red[0];
    if (tstep && blockIdx.x == 0) tstep[0] = tstep[0] + 1;
  }
}

// One training pass of 16 batch rows per block for the closure network's shape: K0 inputs,
// NHH + 1 hidden layers of NH (tanh), 16 padded outputs; NH / 16 waves, wave w computing the
// output columns 16 w .. 16 w + 15 of every hidden layer.  It runs the forward pass, the loss
// terms and the input gradients with the operands, instructions and order of k_fwd, k_mse
// and k_dx for every 16x16 tile, so its results are bit-identical to theirs:
//   forward: acc over k in 32-wide steps, z = acc + b, x = bf16(tanh z);
//   loss (wave 0): dZ = 2 (Y - T) / (B nout) and the tile's sum of (Y - T)^2 by k_mse's
//   256-element tree (one part per 16-row tile: 256 = 16 rows x 16 outputs);
//   input gradients: acc over n in 16-wide steps, dZ = acc (1 - tanh(z)^2).
// A lane keeps the pre-activations it produced (rows 4 (l / 16) + q, column l % 16 of its
// wave's tile) in registers: the input-gradient epilogue of the same tile reads them in the
// same lanes.  Layer inputs and dZ pass between layers through ping-pong LDS tiles (one
// barrier per layer) and go to global memory for the weight gradients (k_dw_all).
struct TrainNet {
  const uint16_t* w[kMaxL];   // bf16 weights of each layer (N x K)
  const float* b[kMaxL];      // FP32 biases
  uint16_t* X[kMaxL];         // layer inputs (bf16, B x K): X[0] read, the others written
  float* dZf[kMaxL];          // dZ of each layer (B x N), FP32 and bf16
  uint16_t* dZb[kMaxL];
  const float* T;             // targets (B x nout)
  float* part;                // loss partial sums, one per 16-row tile
  int B, nout;
};
template <int K0, int NH, int NHH>
__global__ void __launch_bounds__(4 * NH) k_train_s(TrainNet tn) {
  static_assert(K0 % 32 == 0 && NH % 32 == 0, "32-wide forward steps without a tail");
  constexpr int L = NHH + 2, NO = 16;
  __shared__ __align__(16) uint16_t xt[2][16 * NH];    // layer inputs (ping-pong)
  __shared__ __align__(16) uint16_t dzt[2][16 * NH];   // dZ tiles (ping-pong, row stride N)
  __shared__ float red[256];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, li = l & 15, lg = l >> 4;
  const int b0 = blockIdx.x * 16, col = w * 16 + li;
  f4 zr[NHH + 1];                                      // this wave's hidden pre-activations
  {
    const uint16_t* xr = tn.X[0] + (size_t)(b0 + li) * K0;
    const uint16_t* wr = tn.w[0] + (size_t)col * K0;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K0; k0 += 32) acc = mfma32(ld8(xr + k0 + 8 * lg), ld8(wr + k0 + 8 * lg), acc);
    zr[0] = acc;
  }
#pragma unroll
  for (int h = 0; h <= NHH; ++h) {
    uint16_t* xo = xt[(h + 1) & 1];
    const float bb = tn.b[h][col];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float z = zr[h][q] + bb;
      zr[h][q] = z;
      const uint16_t x = f2bf(tanh_act(z));
      xo[(4 * lg + q) * NH + col] = x;
      tn.X[h + 1][(size_t)(b0 + 4 * lg + q) * NH + col] = x;
    }
    __syncthreads();
    bf8 a[NH / 32];
#pragma unroll
    for (int c = 0; c < NH / 32; ++c) a[c] = ld8(xo + li * NH + c * 32 + 8 * lg);
    if (h < NHH) {
      const uint16_t* wr = tn.w[h + 1] + (size_t)col * NH;
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < NH / 32; ++c) acc = mfma32(a[c], ld8(wr + c * 32 + 8 * lg), acc);
      zr[h + 1] = acc;
    } else if (w == 0) {
      // output layer and the loss terms
      const uint16_t* wr = tn.w[L - 1] + (size_t)li * NH;
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < NH / 32; ++c) acc = mfma32(a[c], ld8(wr + c * 32 + 8 * lg), acc);
      const float bo = tn.b[L - 1][li];
      const float sc = 2.f / (float)(tn.B * tn.nout);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = b0 + 4 * lg + q;
        const float y = acc[q] + bo;
        float d = 0.f, s = 0.f;
        if (li < tn.nout) {
          d = y - tn.T[(size_t)r * tn.nout + li];
          s = d * d;
        }
        const size_t o = (size_t)r * NO + li;
        tn.dZf[L - 1][o] = sc * d;
        const uint16_t db = f2bf(sc * d);
        tn.dZb[L - 1][o] = db;
        dzt[0][(4 * lg + q) * NO + li] = db;
        red[(4 * lg + q) * 16 + li] = s;
      }
    }
  }
  __syncthreads();
  if (w == 0) {
    // the tile's loss partial: k_mse's tree (red[i] += red[i + w], w = 128 .. 1)
    red[l] += red[l + 128];
    red[l + 64] += red[l + 192];
#pragma unroll
    for (int s = 64; s > 0; s >>= 1)
      if (l < s) red[l] += red[l + s];
    if (l == 0) tn.part[blockIdx.x] = red[0];
  }
  // input gradients, layer L - 1 down to 1: dZ(l - 1) = (dZ(l) W(l)) (1 - tanh(z)^2)
#pragma unroll
  for (int ly = L - 1; ly >= 1; --ly) {
    constexpr int K = NH;
    const int N = ly == L - 1 ? NO : NH;
    const uint16_t* src = dzt[(L - 1 - ly) & 1];
    uint16_t* dst = dzt[(L - ly) & 1];
    bf4 a[NH / 16];
#pragma unroll
    for (int n0 = 0; n0 < NH; n0 += 16)
      if (n0 < N) a[n0 / 16] = *reinterpret_cast<const bf4*>(src + li * N + n0 + 4 * lg);
    const uint16_t* W = tn.w[ly];
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int n0 = 0; n0 < NH; n0 += 16) {
      if (n0 < N) {
        bf4 b;
#pragma unroll
        for (int q = 0; q < 4; ++q) b[q] = (short)W[(size_t)(n0 + 4 * lg + q) * K + col];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[n0 / 16], b, acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float th = tanh_act(zr[ly - 1][q]);
      const float gq = acc[q] * (1.f - th * th);
      const size_t o = (size_t)(b0 + 4 * lg + q) * K + col;
      tn.dZf[ly - 1][o] = gq;
      const uint16_t gb = f2bf(gq);
      tn.dZb[ly - 1][o] = gb;
      dst[(4 * lg + q) * K + col] = gb;
    }
    if (ly > 1) __syncthreads();
  }
}

// the loss from the partial sums, in a fixed order (one thread)
__device__ inline void sum_loss(const float* __restrict__ part, int nparts, float den, float* __restrict__ loss) {
  float s = 0.f;
  for (int i = 0; i < nparts; ++i) s += part[i];
  loss[0] = s / den;
}
__global__ void k_loss(const float* __restrict__ part, int nparts, float den, float* __restrict__ loss) {
  if (threadIdx.x == 0) sum_loss(part, nparts, den, loss);
}

// batch of B rows drawn with replacement from the data set (x: n x nin f32, y: n x nout f32)
// -> X (B x K0 bf16, zero pad) and T (B x nout f32)
__global__ void k_gather(const float* __restrict__ x, const float* __restrict__ y, long long n, int nin, int nout,
                         int K0, int B, uint64_t seed, const int* __restrict__ tstep, uint16_t* __restrict__ X,
                         float* __restrict__ T) {
  const int r = blockIdx.x;
  if (r >= B) return;
  const uint64_t step = (uint64_t)tstep[0];
  const long long src = (long long)(mix64(seed ^ mix64(step * 0x100000001b3ull + (uint64_t)r)) % (uint64_t)n);
  for (int f = threadIdx.x; f < K0; f += blockDim.x)
    X[(size_t)r * K0 + f] = f < nin ? f2bf(x[(size_t)src * nin + f]) : (uint16_t)0;
  for (int f = threadIdx.x; f < nout; f += blockDim.x) T[(size_t)r * nout + f] = y[(size_t)src * nout + f];
}

// rows of host-given inputs -> X (bf16, zero pad)
__global__ void k_pack_x(const float* __restrict__ x, int nrow, int nin, int K0, int B, uint16_t* __restrict__ X) {
  const int r = blockIdx.x;
  for (int f = threadIdx.x; f < K0; f += blockDim.x)
    X[(size_t)r * K0 + f] = (r < nrow && f < nin) ? f2bf(x[(size_t)r * nin + f]) : (uint16_t)0;
}

// inputs of grid points r0 .. r0 + nrow - 1 (point q = (j nz + k) nx + i): the nch fields g on
// the 3x3x3 neighbourhood, feature ((dj + 1) 9 + (dk + 1) 3 + (di + 1)) nch + c, j clamped to
// 0..ny-1, k and i periodic; then the extra field ex at the point itself when nin = 27 nch + 1.
// Each input is standardised, (v - mean[f]) * rscale[f] with rscale = 1 / scale rounded to
// float, and rounded to bf16 (zero pad).
__global__ void k_pack_grid(const float* __restrict__ g, const float* __restrict__ ex,
                            const float* __restrict__ mean, const float* __restrict__ rscale, int nch, int ny,
                            int nz, int nx, long long r0, int nrow, int nin, int K0, uint16_t* __restrict__ X) {
  const int r = blockIdx.x;
  const long long q = r0 + r;
  const bool live = r < nrow;
  int j = 0, k = 0, i = 0;
  if (live) {
    i = (int)(q % nx);
    k = (int)((q / nx) % nz);
    j = (int)(q / ((long long)nx * nz));
  }
  const size_t plane = (size_t)nz * nx, field = plane * ny;
  for (int f = threadIdx.x; f < K0; f += blockDim.x) {
    uint16_t out = 0;
    if (live && f < nin) {
      float v;
      if (f < 27 * nch) {
        const int c = f % nch, s = f / nch;
        const int dj = s / 9 - 1, dk = (s / 3) % 3 - 1, di = s % 3 - 1;
        const int jj = min(max(j + dj, 0), ny - 1), kk = (k + dk + nz) % nz, ii = (i + di + nx) % nx;
        v = g[c * field + (size_t)jj * plane + (size_t)kk * nx + ii];
      } else {
        v = ex[(size_t)j * plane + (size_t)k * nx + i];
      }
      out = f2bf((v - mean[f]) * rscale[f]);
    }
    X[(size_t)r * K0 + f] = out;
  }
}

// The whole network on 16-point tiles of a grid, one wave per tile: the weights (bf16) sit
// in LDS for the block; the first layer's operands are gathered and standardised straight
// into the matrix-core registers (the inputs k_pack_grid would write); each hidden layer's
// activations go through a 16 x 64 bf16 tile of the wave's own LDS.  Same operands, k order
// and instructions as k_pack_grid + k_fwd, so the outputs are the same bits (every input
// width a multiple of 32, so that k_fwd takes 32-wide steps only).
constexpr int kFuseMaxL = 6;
constexpr int kFuseWaves = 8;     // waves per block sharing the staged weights
constexpr int kFuseMaxK0 = 256;   // first-layer width held in registers
struct FusedNet {
  int L, np, nin, nout;
  int Kp[kFuseMaxL], Np[kFuseMaxL];
  long long offW[kFuseMaxL], offB[kFuseMaxL];
};
__global__ void __launch_bounds__(64 * kFuseWaves) k_mlp_grid(FusedNet fn, const uint16_t* __restrict__ wb,
                                                              const float* __restrict__ p, const float* __restrict__ g,
                                                              const float* __restrict__ ex,
                                                              const float* __restrict__ mean,
                                                              const float* __restrict__ rscale, int nch, int ny, int nz,
                                                              int nx, long long n, float* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t* Wl = reinterpret_cast<uint16_t*>(smem);
  const int wbytes = (2 * fn.np + 15) & ~15;
  const int K0 = fn.Kp[0];
  float* ms = reinterpret_cast<float*>(smem + wbytes);
  int* fdec = reinterpret_cast<int*>(ms + 2 * K0);   // per input: channel, stencil offsets (or ex / pad)
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  uint16_t* act = reinterpret_cast<uint16_t*>(fdec + K0) + wave * 16 * 64;
  for (int e = threadIdx.x; e < fn.np; e += blockDim.x) Wl[e] = wb[e];
  for (int e = threadIdx.x; e < K0; e += blockDim.x) {
    ms[e] = e < fn.nin ? mean[e] : 0.f;
    ms[K0 + e] = e < fn.nin ? rscale[e] : 1.f;
    int d = -1;                                    // padding
    if (e < 27 * nch) {
      const int c = e % nch, s = e / nch;
      d = c | ((s / 9) << 8) | (((s / 3) % 3) << 10) | ((s % 3) << 12);
    } else if (e < fn.nin) {
      d = -2;                                      // the extra field at the point
    }
    fdec[e] = d;
  }
  __syncthreads();
  const size_t plane = (size_t)nz * nx, field = plane * ny;
  const long long ntile = (n + 15) / 16;
  for (long long tile = (long long)blockIdx.x * kFuseWaves + wave; tile < ntile;
       tile += (long long)gridDim.x * kFuseWaves) {
    const long long r0 = tile * 16;
    const long long q = r0 + l % 16;               // this lane's point (rows of A)
    const bool live = q < n;
    int i = 0, k = 0, j = 0;
    if (live) {
      i = (int)(q % nx);
      k = (int)((q / nx) % nz);
      j = (int)(q / ((long long)nx * nz));
    }
    // every first-layer operand of the lane first (the loads overlap), then the products
    float v[kFuseMaxK0 / 32][8];
#pragma unroll
    for (int c = 0; c < kFuseMaxK0 / 32; ++c)
#pragma unroll
      for (int qq = 0; qq < 8; ++qq) {
        const int f = c * 32 + 8 * (l / 16) + qq;
        float x = 0.f;
        if (c * 32 < K0 && live) {
          const int d = fdec[f];
          if (d >= 0) {
            int jj = j + ((d >> 8) & 3) - 1, kk = k + ((d >> 10) & 3) - 1, ii = i + ((d >> 12) & 3) - 1;
            jj = jj < 0 ? 0 : (jj >= ny ? ny - 1 : jj);
            kk = kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk);
            ii = ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii);
            x = g[(size_t)(d & 255) * field + (size_t)jj * plane + (size_t)kk * nx + ii];
          } else if (d == -2) {
            x = ex[(size_t)j * plane + (size_t)k * nx + i];
          }
        }
        v[c][qq] = x;
      }
    f4 acc[4];
    {
      const int nnt = fn.Np[0] / 16;
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < kFuseMaxK0 / 32; ++c)
        if (c * 32 < K0) {
          s8 a;
#pragma unroll
          for (int qq = 0; qq < 8; ++qq) {
            const int f = c * 32 + 8 * (l / 16) + qq;
            a[qq] = (short)((live && f < fn.nin) ? f2bf((v[c][qq] - ms[f]) * ms[K0 + f]) : (uint16_t)0);
          }
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (t < nnt)
              acc[t] = mfma32(__builtin_bit_cast(bf8, a),
                              ld8(Wl + fn.offW[0] + (size_t)(t * 16 + l % 16) * K0 + c * 32 + 8 * (l / 16)), acc[t]);
        }
    }
    for (int layer = 0; layer < fn.L; ++layer) {
      const int nnt = fn.Np[layer] / 16;
      const float* bias = p + fn.offB[layer];
      if (layer == fn.L - 1) {                     // output rows (pre-activations)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (t < nnt) {
            const int nn = t * 16 + l % 16;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
              const long long r = r0 + 4 * (l / 16) + qq;
              if (r < n && nn < fn.nout) y[(size_t)r * fn.nout + nn] = acc[t][qq] + bias[nn];
            }
          }
        break;
      }
      // activations of this layer -> the wave's tile (D layout in, row-major out)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t < nnt) {
          const int nn = t * 16 + l % 16;
#pragma unroll
          for (int qq = 0; qq < 4; ++qq)
            act[(4 * (l / 16) + qq) * 64 + nn] = f2bf(tanh_act(acc[t][qq] + bias[nn]));
        }
      // next layer: A from the tile (row l % 16, eight contiguous k per 32-wide chunk)
      const int K = fn.Kp[layer + 1], nnx = fn.Np[layer + 1] / 16;
      bf8 a[2];
#pragma unroll
      for (int c = 0; c < 2; ++c)
        if (c * 32 < K) a[c] = ld8(act + (l % 16) * 64 + c * 32 + 8 * (l / 16));
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 2; ++c)
        if (c * 32 < K) {
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (t < nnx)
              acc[t] = mfma32(a[c], ld8(Wl + fn.offW[layer + 1] + (size_t)(t * 16 + l % 16) * K + c * 32 + 8 * (l / 16)),
                              acc[t]);
        }
    }
  }
}

// k_mlp_grid with the closure network's widths known at compile time (first layer K0 wide,
// NHH + 1 hidden layers NH wide, 16 padded outputs): every LDS operand address is a per-lane
// base plus a constant; the gather offsets come from an LDS table four at a time, and tiles
// whose 16 points (one x row, nx a multiple of 16) need no clamp or wrap skip that logic.
// Same operands, k order and instructions as k_mlp_grid.
constexpr int kOffEx = -2147483647;      // input offset sentinels: the extra field, padding
constexpr int kOffPad = -2147483647 - 1;
// waves per block of k_mlp_grid_s: one block per CU (the weights take 59 KB of LDS per block),
// four waves per SIMD (the compiler fits 128 VGPRs with 36 bytes of spill; 1.05x faster per
// coupling than three waves at 138 VGPRs, 1.15x than two)
constexpr int kSpecWaves = 16;
template <int K0, int NH, int NHH>
__global__ void __launch_bounds__(64 * kSpecWaves) k_mlp_grid_s(FusedNet fn, const uint16_t* __restrict__ wb,
                                                                const float* __restrict__ p,
                                                                const float* __restrict__ g,
                                                                const float* __restrict__ ex,
                                                                const float* __restrict__ mean,
                                                                const float* __restrict__ rscale, int nch, int ny,
                                                                int nz, int nx, long long n, float* __restrict__ y) {
  static_assert(K0 % 32 == 0 && NH % 32 == 0, "32-wide matrix-core steps");
  constexpr int NO = 16;
  constexpr int WH = NH * K0, WO = WH + NHH * NH * NH, WN = WO + NO * NH;   // LDS weight offsets
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t* Wl = reinterpret_cast<uint16_t*>(smem);
  float* ms = reinterpret_cast<float*>(smem + ((2 * WN + 15) & ~15));   // mean[K0], rscale[K0]
  // interior offset per input, three tables: as is, and with the x neighbour wrapped for the
  // first (i = 0: dx = -1 reads x = nx - 1) and the last point of a row (dx = +1 reads x = 0)
  int* offs = reinterpret_cast<int*>(ms + 2 * K0);
  int* fdec = offs + 3 * K0;                                            // channel / stencil code
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, li = l & 15, lg = l >> 4;
  uint16_t* act = reinterpret_cast<uint16_t*>(fdec + K0) + wave * 16 * NH;
  const int plane = nz * nx, field = plane * ny;
  for (int e = threadIdx.x; e < WN; e += blockDim.x) {
    long long src;
    if (e < WH) src = fn.offW[0] + e;
    else if (e < WO) src = fn.offW[1 + (e - WH) / (NH * NH)] + (e - WH) % (NH * NH);
    else src = fn.offW[1 + NHH] + (e - WO);
    Wl[e] = wb[src];
  }
  for (int e = threadIdx.x; e < K0; e += blockDim.x) {
    ms[e] = e < fn.nin ? mean[e] : 0.f;
    ms[K0 + e] = e < fn.nin ? rscale[e] : 1.f;
    int d = -1, o = kOffPad, ol = kOffPad, orr = kOffPad;
    if (e < 27 * nch) {
      const int c = e % nch, s = e / nch, dx = s % 3 - 1;
      d = c | ((s / 9) << 8) | (((s / 3) % 3) << 10) | ((s % 3) << 12);
      o = c * field + (s / 9 - 1) * plane + ((s / 3) % 3 - 1) * nx + dx;
      ol = o + (dx == -1 ? nx : 0);
      orr = o - (dx == 1 ? nx : 0);
    } else if (e < fn.nin) {
      d = -2;
      o = ol = orr = kOffEx;
    }
    fdec[e] = d;
    offs[e] = o;
    offs[K0 + e] = ol;
    offs[2 * K0 + e] = orr;
  }
  __syncthreads();
  const uint16_t* w0 = Wl + li * K0 + 8 * lg;        // B operands: lane base + constants
  const uint16_t* wh = Wl + WH + li * NH + 8 * lg;
  const uint16_t* wo = Wl + WO + li * NH + 8 * lg;
  const long long ntile = (n + 15) / 16;
  for (long long tile = (long long)blockIdx.x * kSpecWaves + wave; tile < ntile;
       tile += (long long)gridDim.x * kSpecWaves) {
    const long long r0 = tile * 16;
    const int i0 = (int)(r0 % nx), k = (int)((r0 / nx) % nz), j = (int)(r0 / ((long long)nx * nz));
    const int i = i0 + li;
    const bool live = r0 + li < n;
    // full tiles off the y end planes and the z end rows (a row of nx % 16 == 0 points holds
    // whole tiles; its end points take the wrapped tables)
    const bool inner = r0 + 16 <= n && k > 0 && k + 1 < nz && j > 0 && j + 1 < ny;
    const int pidx = j * plane + k * nx + i;
    float v[K0 / 32][8];
    if (inner) {
      // no branches: every input loads (a sentinel loads the point itself) and selects
      const float exv = ex ? ex[pidx] : 0.f;
      const int* ot = offs + (i == 0 ? K0 : (i == nx - 1 ? 2 * K0 : 0));
#pragma unroll
      for (int c = 0; c < K0 / 32; ++c) {
        const int4 oa = *reinterpret_cast<const int4*>(ot + c * 32 + 8 * lg);
        const int4 ob = *reinterpret_cast<const int4*>(ot + c * 32 + 8 * lg + 4);
        const int oo[8] = {oa.x, oa.y, oa.z, oa.w, ob.x, ob.y, ob.z, ob.w};
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) {
          const float gv = g[pidx + (oo[qq] > kOffEx ? oo[qq] : 0)];
          v[c][qq] = oo[qq] > kOffEx ? gv : (oo[qq] == kOffEx ? exv : 0.f);
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < K0 / 32; ++c)
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) {
          const int d = fdec[c * 32 + 8 * lg + qq];
          float x = 0.f;
          if (live && d >= 0) {
            int jj = j + ((d >> 8) & 3) - 1, kk = k + ((d >> 10) & 3) - 1, ii = i + ((d >> 12) & 3) - 1;
            jj = jj < 0 ? 0 : (jj >= ny ? ny - 1 : jj);
            kk = kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk);
            ii = ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii);
            x = g[(size_t)(d & 255) * field + (size_t)jj * plane + (size_t)kk * nx + ii];
          } else if (live && d == -2) {
            x = ex[pidx];
          }
          v[c][qq] = x;
        }
    }
    f4 acc[NH / 16];
#pragma unroll
    for (int t = 0; t < NH / 16; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < K0 / 32; ++c) {
      const float4 m0 = *reinterpret_cast<const float4*>(ms + c * 32 + 8 * lg);
      const float4 m1 = *reinterpret_cast<const float4*>(ms + c * 32 + 8 * lg + 4);
      const float4 s0 = *reinterpret_cast<const float4*>(ms + K0 + c * 32 + 8 * lg);
      const float4 s1 = *reinterpret_cast<const float4*>(ms + K0 + c * 32 + 8 * lg + 4);
      const float mv[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
      const float sv[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
      s8 a;
#pragma unroll
      for (int qq = 0; qq < 8; ++qq) {
        const int f = c * 32 + 8 * lg + qq;
        a[qq] = (short)((live && f < fn.nin) ? f2bf((v[c][qq] - mv[qq]) * sv[qq]) : (uint16_t)0);
      }
#pragma unroll
      for (int t = 0; t < NH / 16; ++t) acc[t] = mfma32(__builtin_bit_cast(bf8, a), ld8(w0 + t * 16 * K0 + c * 32), acc[t]);
    }
    // hidden layers: activations through the wave's tile, then the next product
#pragma unroll
    for (int h = 0; h <= NHH; ++h) {
      const float* bias = p + fn.offB[h];
#pragma unroll
      for (int t = 0; t < NH / 16; ++t) {
        const float bb = bias[t * 16 + li];
#pragma unroll
        for (int qq = 0; qq < 4; ++qq)
          act[(4 * lg + qq) * NH + t * 16 + li] = f2bf(tanh_act(acc[t][qq] + bb));
      }
      bf8 a[NH / 32];
#pragma unroll
      for (int c = 0; c < NH / 32; ++c) a[c] = ld8(act + li * NH + c * 32 + 8 * lg);
      if (h < NHH) {
#pragma unroll
        for (int t = 0; t < NH / 16; ++t) {
          acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int c = 0; c < NH / 32; ++c) acc[t] = mfma32(a[c], ld8(wh + h * NH * NH + t * 16 * NH + c * 32), acc[t]);
        }
      } else {
        f4 ao = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < NH / 32; ++c) ao = mfma32(a[c], ld8(wo + c * 32), ao);
        const float bo = p[fn.offB[NHH + 1] + li];
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const long long r = r0 + 4 * lg + qq;
          if (r < n && li < fn.nout) y[(size_t)r * fn.nout + li] = ao[qq] + bo;
        }
      }
    }
  }
}

// real outputs of nrow rows (Z: B x NP f32) -> y (nrow x nout)
__global__ void k_unpack_out(const float* __restrict__ Z, int NP, int nout, int nrow, float* __restrict__ y) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nrow * nout) y[e] = Z[(size_t)(e / nout) * NP + e % nout];
}

// Adam on the FP32 master parameters (step t = tstep, already advanced by k_mse; bias
// corrections 1 - beta^t); refreshes the bf16 copy of the weights; the first thread also sums
// the step's loss
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, long long n, float lr, float b1, float b2,
                       float eps, const int* __restrict__ tstep, const float* __restrict__ part, int nparts,
                       float den, float* __restrict__ loss) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0) {                            // the step's loss: fixed-order tree
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = red[0] / den;
  }
  if (e >= n) return;
  const float t = (float)tstep[0];
  const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
  const float ge = g[e];
  const float me = b1 * m[e] + (1.f - b1) * ge;
  const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  const float pe = p[e] - lr * (me / c1) / (sqrtf(ve / c2) + eps);
  p[e] = pe;
  if (wb) wb[e] = f2bf(pe);
}

__global__ void k_to_bf(const float* __restrict__ p, uint16_t* __restrict__ wb, long long n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) wb[e] = f2bf(p[e]);
}

}  // namespace

// ---- handle ------------------------------------------------------------------------------
struct lh_mlp {
  int nin = 0, nhid = 0, nlay = 0, nout = 0, B = 0;
  int L = 0;                                   // weight layers = nlay + 1
  std::vector<int> Kp, Np, K, N;               // padded / real input and output widths
  std::vector<long long> offW, offB;           // offsets in the padded parameter vector
  long long np = 0;                            // padded parameter count
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // FP32, padded
  uint16_t* wb = nullptr;                      // bf16 copy of p (weights and biases)
  std::vector<uint16_t*> X;                    // layer inputs (bf16), X[0] = batch input
  std::vector<float*> Z;                       // pre-activations (f32)
  float *dZ = nullptr, *dZ2 = nullptr, *T = nullptr, *loss = nullptr, *part = nullptr;
  uint16_t *dZb = nullptr, *dZb2 = nullptr;     // dZ of layers of even / odd index (f32, bf16)
  int nparts = 0;                               // loss partial sums per batch
  // the closure network's shape (256 inputs, four hidden layers of 64, 16 padded outputs)
  // trains through k_train_s and k_dw_all, with one dZ buffer per layer
  bool spec = false;
  bool fused = true;                           // lh_mlp_set_fused
  std::vector<float*> dZl;
  std::vector<uint16_t*> dZbl;
  float *dx = nullptr, *dy = nullptr;          // data set on the device
  long long ndata = 0;
  int t = 0;                                   // Adam step count (host copy)
  int* tdev = nullptr;                         // Adam step count on the device
  hipGraphExec_t gstep = nullptr;              // one captured training step
  hipGraphExec_t gmulti = nullptr;             // kTrainSteps steps in one graph
  // inference workspace of Be rows (allocated on first use by the grid evaluation)
  int Be = 0;
  int ncu = 0;                                 // compute units (fused grid evaluation)
  int ncu_s = 0;                               // the same, for the compile-time-shape kernel
  std::vector<uint16_t*> Xe;
  std::vector<float*> Ze;
  float g_lr = 0.f;
  unsigned long long g_seed = 0;
  hipStream_t st = nullptr;
  std::string err;
};

// training steps per captured graph
constexpr int kTrainSteps = 16;

static void drop_train_graphs(lh_mlp* h) {
  if (h->gstep) (void)hipGraphExecDestroy(h->gstep);
  if (h->gmulti) (void)hipGraphExecDestroy(h->gmulti);
  h->gstep = h->gmulti = nullptr;
}

static int mfail(lh_mlp* h, int code, const std::string& m) {
  if (h) h->err = m;
  return code;
}
#define MCHK(h, x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return mfail(h, LH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static hipError_t zalloc(T** p, size_t n) {
  hipError_t e = hipMalloc((void**)p, n * sizeof(T) + 16);
  if (e == hipSuccess) e = hipMemset(*p, 0, n * sizeof(T) + 16);
  return e;
}

// forward pass of `rows` rows (a multiple of 16) in X[0]; the output pre-activations land in
// Z[L-1]
static void forward_rows(lh_mlp* h, int rows, uint16_t* const* X, float* const* Z, hipStream_t st) {
  for (int l = 0; l < h->L; ++l) {
    const bool last = l == h->L - 1;
    hipLaunchKernelGGL(k_fwd, dim3(rows / 16, h->Np[l] / 16), dim3(64), 0, st, X[l], h->wb + h->offW[l],
                       h->p + h->offB[l], Z[l], last ? nullptr : X[l + 1], h->Kp[l], h->Np[l], last ? 0 : 1);
  }
}

// forward pass of the batch in X[0] (B rows)
static void mlp_forward(lh_mlp* h) { forward_rows(h, h->B, h->X.data(), h->Z.data(), h->st); }

static float* dz_f(lh_mlp* h, int l) { return l % 2 ? h->dZ2 : h->dZ; }
static uint16_t* dz_b(lh_mlp* h, int l) { return l % 2 ? h->dZb2 : h->dZb; }

// loss terms of the batch (output rows Z[L-1] against T): dZ of the output layer and the loss
// partial sums; with tick, the step counter advances
static void mlp_loss(lh_mlp* h, bool tick) {
  const int l = h->L - 1;
  hipLaunchKernelGGL(k_mse, dim3(h->nparts), dim3(256), 0, h->st, h->Z[l], h->T, dz_f(h, l), dz_b(h, l), h->part,
                     tick ? h->tdev : nullptr, h->B, h->Np[l], h->nout);
}

// backward pass from the output layer's dZ; fills g (weights and biases)
static void mlp_backward(lh_mlp* h) {
  for (int l = h->L - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_dw, dim3(h->Np[l] / 16, h->Kp[l] / 16 + 1), dim3(64 * kDwWaves), 0, h->st, dz_b(h, l),
                       h->X[l], h->g + h->offW[l], h->B, h->Kp[l], h->Np[l], dz_f(h, l), h->g + h->offB[l]);
    if (l > 0)
      hipLaunchKernelGGL(k_dx, dim3(h->B / 16, h->Kp[l] / 16), dim3(64), 0, h->st, dz_b(h, l), h->wb + h->offW[l],
                         h->Z[l - 1], dz_f(h, l - 1), dz_b(h, l - 1), h->Kp[l], h->Np[l]);
  }
}

// the closure network's shape: forward, loss terms and input gradients of the batch in X[0]
// (k_train_s), then every layer's weight gradients (k_dw_all); with tick, the step counter
// advances.  The same results as mlp_forward + mlp_loss + mlp_backward, in two launches.
static void spec_pass(lh_mlp* h, bool tick) {
  TrainNet tn{};
  DwAll a{};
  int nb = 0;
  for (int l = 0; l < h->L; ++l) {
    tn.w[l] = h->wb + h->offW[l];
    tn.b[l] = h->p + h->offB[l];
    tn.X[l] = h->X[l];
    tn.dZf[l] = h->dZl[l];
    tn.dZb[l] = h->dZbl[l];
    a.dZ[l] = h->dZbl[l];
    a.X[l] = h->X[l];
    a.dW[l] = h->g + h->offW[l];
    a.dZf[l] = h->dZl[l];
    a.db[l] = h->g + h->offB[l];
    a.K[l] = h->Kp[l];
    a.N[l] = h->Np[l];
    a.first[l] = nb;
    nb += (h->Np[l] / 16) * (h->Kp[l] / 16 + 1);
  }
  a.first[h->L] = nb;
  a.L = h->L;
  a.B = h->B;
  a.tstep = tick ? h->tdev : nullptr;
  tn.T = h->T;
  tn.part = h->part;
  tn.B = h->B;
  tn.nout = h->nout;
  hipLaunchKernelGGL((k_train_s<256, 64, 3>), dim3(h->B / 16), dim3(4 * 64), 0, h->st, tn);
  hipLaunchKernelGGL(k_dw_all, dim3(nb), dim3(64 * kDwWaves), 0, h->st, a);
}

extern "C" {

int lh_mlp_create(int nin, int nhid, int nlayers, int nout, int batch, unsigned long long seed, lh_mlp** out) {
  if (!out) return LH_ERR_ARG;
  *out = nullptr;
  lh_mlp* h = new lh_mlp();
  *out = h;
  if (nin < 1 || nhid < 1 || nlayers < 1 || nout < 1 || batch < 16 || batch % 16)
    return mfail(h, LH_ERR_ARG, "lh_mlp_create: nin, nhid, nlayers, nout >= 1; batch a multiple of 16");
  h->nin = nin; h->nhid = nhid; h->nlay = nlayers; h->nout = nout; h->B = batch;
  h->L = nlayers + 1;
  long long off = 0;
  for (int l = 0; l < h->L; ++l) {
    const int k = l == 0 ? nin : nhid, n = l == h->L - 1 ? nout : nhid;
    h->K.push_back(k); h->N.push_back(n);
    h->Kp.push_back(pad16(k)); h->Np.push_back(pad16(n));
    h->offW.push_back(off); off += (long long)pad16(n) * pad16(k);
    h->offB.push_back(off); off += pad16(n);
  }
  h->np = off;
  MCHK(h, hipStreamCreate(&h->st));
  MCHK(h, zalloc(&h->p, h->np)); MCHK(h, zalloc(&h->g, h->np));
  MCHK(h, zalloc(&h->m, h->np)); MCHK(h, zalloc(&h->v, h->np));
  MCHK(h, zalloc(&h->wb, h->np));
  int wmax = 0;
  for (int l = 0; l < h->L; ++l) {
    uint16_t* x = nullptr;
    float* z = nullptr;
    MCHK(h, zalloc(&x, (size_t)batch * h->Kp[l]));
    MCHK(h, zalloc(&z, (size_t)batch * h->Np[l]));
    h->X.push_back(x); h->Z.push_back(z);
    wmax = std::max(wmax, std::max(h->Kp[l], h->Np[l]));
  }
  MCHK(h, zalloc(&h->dZ, (size_t)batch * wmax)); MCHK(h, zalloc(&h->dZ2, (size_t)batch * wmax));
  MCHK(h, zalloc(&h->dZb, (size_t)batch * wmax)); MCHK(h, zalloc(&h->dZb2, (size_t)batch * wmax));
  MCHK(h, zalloc(&h->T, (size_t)batch * nout));
  h->spec = h->L == 5 && h->Kp[0] == 256 && h->Np[h->L - 1] == 16;
  for (int l = 0; l + 1 < h->L; ++l) h->spec = h->spec && h->Np[l] == 64;
  if (h->spec) {
    for (int l = 0; l < h->L; ++l) {
      float* f = nullptr;
      uint16_t* b = nullptr;
      MCHK(h, zalloc(&f, (size_t)batch * h->Np[l]));
      h->dZl.push_back(f);
      MCHK(h, zalloc(&b, (size_t)batch * h->Np[l]));
      h->dZbl.push_back(b);
    }
  }
  h->nparts = (batch * h->Np[h->L - 1] + 255) / 256;
  MCHK(h, zalloc(&h->part, (size_t)h->nparts));
  MCHK(h, zalloc(&h->loss, 16));
  MCHK(h, zalloc(&h->tdev, 16));
  // Glorot-uniform weights from a counter-based generator, zero biases
  std::vector<float> hp(h->np, 0.f);
  for (int l = 0; l < h->L; ++l) {
    const float lim = std::sqrt(6.f / (float)(h->K[l] + h->N[l]));
    for (int n = 0; n < h->N[l]; ++n)
      for (int k = 0; k < h->K[l]; ++k) {
        const uint64_t r = mix64(seed ^ mix64(((uint64_t)l << 40) ^ ((uint64_t)n << 20) ^ (uint64_t)k));
        const float u = (float)((r >> 11) * (1.0 / 9007199254740992.0));
        hp[h->offW[l] + (long long)n * h->Kp[l] + k] = lim * (2.f * u - 1.f);
      }
  }
  MCHK(h, hipMemcpy(h->p, hp.data(), h->np * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((h->np + 255) / 256)), dim3(256), 0, h->st, h->p, h->wb, h->np);
  MCHK(h, hipStreamSynchronize(h->st));
  return LH_OK;
}

int lh_mlp_destroy(lh_mlp* h) {
  if (!h) return LH_OK;
  if (h->st) (void)hipStreamSynchronize(h->st);
  if (h->gstep) (void)hipGraphExecDestroy(h->gstep);
  if (h->gmulti) (void)hipGraphExecDestroy(h->gmulti);
  void* ptrs[] = {h->p,  h->g,    h->m,    h->v,  h->wb,   h->dZ,  h->dZ2, h->dZb,
                  h->dZb2, h->T, h->loss, h->part, h->dx, h->dy, h->tdev};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (auto q : h->X) if (q) (void)hipFree(q);
  for (auto q : h->Z) if (q) (void)hipFree(q);
  for (auto q : h->Xe) if (q) (void)hipFree(q);
  for (auto q : h->Ze) if (q) (void)hipFree(q);
  for (auto q : h->dZl) if (q) (void)hipFree(q);
  for (auto q : h->dZbl) if (q) (void)hipFree(q);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return LH_OK;
}

const char* lh_mlp_last_error(lh_mlp* h) { return h ? h->err.c_str() : "null handle"; }

long long lh_mlp_param_count(lh_mlp* h) {
  if (!h) return 0;
  long long n = 0;
  for (int l = 0; l < h->L; ++l) n += (long long)h->N[l] * h->K[l] + h->N[l];
  return n;
}

// parameters (FP32, unpadded): per layer W (N x K, row-major) then b (N)
static void pack_params(const lh_mlp* h, const float* padded, float* flat) {
  long long o = 0;
  for (int l = 0; l < h->L; ++l) {
    for (int n = 0; n < h->N[l]; ++n)
      for (int k = 0; k < h->K[l]; ++k) flat[o++] = padded[h->offW[l] + (long long)n * h->Kp[l] + k];
    for (int n = 0; n < h->N[l]; ++n) flat[o++] = padded[h->offB[l] + n];
  }
}

int lh_mlp_get_params(lh_mlp* h, float* params) {
  if (!h || !params) return mfail(h, LH_ERR_ARG, "lh_mlp_get_params: null argument");
  std::vector<float> hp(h->np);
  MCHK(h, hipStreamSynchronize(h->st));
  MCHK(h, hipMemcpy(hp.data(), h->p, h->np * sizeof(float), hipMemcpyDeviceToHost));
  pack_params(h, hp.data(), params);
  return LH_OK;
}

int lh_mlp_set_params(lh_mlp* h, const float* params) {
  if (!h || !params) return mfail(h, LH_ERR_ARG, "lh_mlp_set_params: null argument");
  std::vector<float> hp(h->np, 0.f);
  long long o = 0;
  for (int l = 0; l < h->L; ++l) {
    for (int n = 0; n < h->N[l]; ++n)
      for (int k = 0; k < h->K[l]; ++k) hp[h->offW[l] + (long long)n * h->Kp[l] + k] = params[o++];
    for (int n = 0; n < h->N[l]; ++n) hp[h->offB[l] + n] = params[o++];
  }
  MCHK(h, hipStreamSynchronize(h->st));
  MCHK(h, hipMemcpy(h->p, hp.data(), h->np * sizeof(float), hipMemcpyHostToDevice));
  MCHK(h, hipMemset(h->m, 0, h->np * sizeof(float)));
  MCHK(h, hipMemset(h->v, 0, h->np * sizeof(float)));
  MCHK(h, hipMemset(h->tdev, 0, sizeof(int)));
  h->t = 0;
  hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((h->np + 255) / 256)), dim3(256), 0, h->st, h->p, h->wb, h->np);
  MCHK(h, hipStreamSynchronize(h->st));
  return LH_OK;
}

int lh_mlp_set_data(lh_mlp* h, const float* x, const float* y, long long n) {
  if (!h || !x || !y || n < 1) return mfail(h, LH_ERR_ARG, "lh_mlp_set_data: null argument or n < 1");
  if (h->dx) (void)hipFree(h->dx);
  if (h->dy) (void)hipFree(h->dy);
  h->dx = h->dy = nullptr;
  drop_train_graphs(h);                        // they hold dx, dy
  MCHK(h, hipMalloc((void**)&h->dx, (size_t)n * h->nin * sizeof(float)));
  MCHK(h, hipMalloc((void**)&h->dy, (size_t)n * h->nout * sizeof(float)));
  MCHK(h, hipMemcpy(h->dx, x, (size_t)n * h->nin * sizeof(float), hipMemcpyHostToDevice));
  MCHK(h, hipMemcpy(h->dy, y, (size_t)n * h->nout * sizeof(float), hipMemcpyHostToDevice));
  h->ndata = n;
  return LH_OK;
}

// loss and FP32 gradients of one batch of `batch` host rows (x: batch x nin, y: batch x nout)
int lh_mlp_loss_grad(lh_mlp* h, const float* x, const float* y, float* loss, float* grad) {
  if (!h || !x || !y) return mfail(h, LH_ERR_ARG, "lh_mlp_loss_grad: null argument");
  float* dx = nullptr;
  MCHK(h, hipMalloc((void**)&dx, (size_t)h->B * h->nin * sizeof(float)));
  MCHK(h, hipMemcpy(dx, x, (size_t)h->B * h->nin * sizeof(float), hipMemcpyHostToDevice));
  MCHK(h, hipMemcpy(h->T, y, (size_t)h->B * h->nout * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_pack_x, dim3(h->B), dim3(64), 0, h->st, dx, h->B, h->nin, h->Kp[0], h->B, h->X[0]);
  if (h->spec && h->fused) {
    spec_pass(h, false);
  } else {
    mlp_forward(h);
    mlp_loss(h, false);
    mlp_backward(h);
  }
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(64), 0, h->st, h->part, h->nparts, (float)(h->B * h->nout), h->loss);
  MCHK(h, hipStreamSynchronize(h->st));
  (void)hipFree(dx);
  if (loss) MCHK(h, hipMemcpy(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost));
  if (grad) {
    std::vector<float> hg(h->np);
    MCHK(h, hipMemcpy(hg.data(), h->g, h->np * sizeof(float), hipMemcpyDeviceToHost));
    pack_params(h, hg.data(), grad);
  }
  return LH_OK;
}

// one Adam step on a batch drawn from the data set; the step count lives on the device, so
// the same launches (captured once as a graph) serve every step
static void issue_train_step(lh_mlp* h, float lr, unsigned long long seed) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  hipLaunchKernelGGL(k_gather, dim3(h->B), dim3(64), 0, h->st, h->dx, h->dy, h->ndata, h->nin, h->nout,
                     h->Kp[0], h->B, (uint64_t)seed, (const int*)h->tdev, h->X[0], h->T);
  if (h->spec && h->fused) {
    spec_pass(h, true);
  } else {
    mlp_forward(h);
    mlp_loss(h, true);
    mlp_backward(h);
  }
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((h->np + 255) / 256)), dim3(256), 0, h->st, h->p, h->g, h->m,
                     h->v, h->wb, h->np, lr, b1, b2, eps, (const int*)h->tdev, h->part, h->nparts,
                     (float)(h->B * h->nout), h->loss);
}

// captures `n` consecutive training steps as one graph
static int capture_steps(lh_mlp* h, int n, float lr, unsigned long long seed, hipGraphExec_t* out) {
  hipGraph_t g;
  MCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < n; ++s) issue_train_step(h, lr, seed);
  MCHK(h, hipStreamEndCapture(h->st, &g));
  const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MCHK(h, e);
  return LH_OK;
}

// `steps` Adam steps on batches drawn from the data set; loss_last = loss of the last batch.
// The steps replay a graph of kTrainSteps steps, the remainder a graph of one step: each graph
// launch has a fixed cost that one step alone does not cover.
int lh_mlp_train(lh_mlp* h, int steps, float lr, unsigned long long seed, float* loss_last) {
  if (!h || steps < 0) return mfail(h, LH_ERR_ARG, "lh_mlp_train: bad argument");
  if (!h->dx) return mfail(h, LH_ERR_ARG, "lh_mlp_train: no data set (lh_mlp_set_data)");
  MCHK(h, hipStreamSynchronize(h->st));
  if ((h->gstep || h->gmulti) && (h->g_lr != lr || h->g_seed != seed)) drop_train_graphs(h);
  h->g_lr = lr;
  h->g_seed = seed;
  int r;
  if (steps >= kTrainSteps && !h->gmulti && (r = capture_steps(h, kTrainSteps, lr, seed, &h->gmulti))) return r;
  if (steps % kTrainSteps && !h->gstep && (r = capture_steps(h, 1, lr, seed, &h->gstep))) return r;
  for (int s = 0; s + kTrainSteps <= steps; s += kTrainSteps) MCHK(h, hipGraphLaunch(h->gmulti, h->st));
  for (int s = 0; s < steps % kTrainSteps; ++s) MCHK(h, hipGraphLaunch(h->gstep, h->st));
  h->t += steps;
  MCHK(h, hipGetLastError());
  MCHK(h, hipStreamSynchronize(h->st));
  if (loss_last) MCHK(h, hipMemcpy(loss_last, h->loss, sizeof(float), hipMemcpyDeviceToHost));
  return LH_OK;
}

int lh_mlp_set_fused(lh_mlp* h, int on) {
  if (!h) return LH_ERR_ARG;
  MCHK(h, hipStreamSynchronize(h->st));
  if (h->fused != (on != 0)) drop_train_graphs(h);   // the captured steps hold the other path
  h->fused = on != 0;
  return LH_OK;
}

// outputs of n host rows (x: n x nin) -> y (n x nout), in batches of B
int lh_mlp_predict(lh_mlp* h, const float* x, long long n, float* y) {
  if (!h || !x || !y || n < 0) return mfail(h, LH_ERR_ARG, "lh_mlp_predict: bad argument");
  float* dx = nullptr;
  MCHK(h, hipMalloc((void**)&dx, (size_t)h->B * h->nin * sizeof(float)));
  std::vector<float> z((size_t)h->B * h->Np[h->L - 1]);
  for (long long r0 = 0; r0 < n; r0 += h->B) {
    const int nr = (int)std::min<long long>(h->B, n - r0);
    MCHK(h, hipMemcpy(dx, x + (size_t)r0 * h->nin, (size_t)nr * h->nin * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pack_x, dim3(h->B), dim3(64), 0, h->st, dx, nr, h->nin, h->Kp[0], h->B, h->X[0]);
    forward_rows(h, h->B, h->X.data(), h->Z.data(), h->st);
    MCHK(h, hipStreamSynchronize(h->st));
    MCHK(h, hipMemcpy(z.data(), h->Z[h->L - 1], z.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int r = 0; r < nr; ++r)
      for (int o = 0; o < h->nout; ++o) y[(size_t)(r0 + r) * h->nout + o] = z[(size_t)r * h->Np[h->L - 1] + o];
  }
  (void)hipFree(dx);
  return LH_OK;
}

// outputs at every point of an ny x nz x nx grid from stencil inputs gathered on the device
// (k_pack_grid): g holds nch fields (nch x ny x nz x nx), ex the extra field (ny x nz x nx) or
// null, mean / scale the nin standardisation constants (input f becomes (v - mean[f]) *
// (1 / scale[f]), the reciprocal rounded to float); y receives ny nz nx rows of nout
int lh_mlp_predict_grid(lh_mlp* h, const float* g, int nch, const float* ex, int ny, int nz, int nx,
                        const float* mean, const float* scale, float* y) {
  if (!h || !g || !mean || !scale || !y || nch < 1 || ny < 1 || nz < 1 || nx < 1)
    return mfail(h, LH_ERR_ARG, "lh_mlp_predict_grid: bad argument");
  if (h->nin != 27 * nch + (ex ? 1 : 0))
    return mfail(h, LH_ERR_ARG, "lh_mlp_predict_grid: nin must be 27 nch (+1 with the extra field)");
  const long long n = (long long)ny * nz * nx;
  const size_t field = (size_t)n;
  float *dg = nullptr, *dex = nullptr, *dms = nullptr, *dy = nullptr;
  auto cleanup = [&]() {
    for (float* q : {dg, dex, dms, dy})
      if (q) (void)hipFree(q);
  };
  hipError_t e = hipMalloc((void**)&dg, field * nch * sizeof(float));
  if (e == hipSuccess && ex) e = hipMalloc((void**)&dex, field * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&dms, 2 * (size_t)h->nin * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&dy, (size_t)n * h->nout * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dg, g, field * nch * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess && ex) e = hipMemcpy(dex, ex, field * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dms, mean, h->nin * sizeof(float), hipMemcpyHostToDevice);
  std::vector<float> rs(h->nin);
  for (int f = 0; f < h->nin; ++f) rs[f] = 1.0f / scale[f];
  if (e == hipSuccess) e = hipMemcpy(dms + h->nin, rs.data(), h->nin * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    cleanup();
    return mfail(h, LH_ERR_HIP, std::string("lh_mlp_predict_grid: ") + hipGetErrorString(e));
  }
  int r = lh_mlp_eval_grid_dev(h, dg, nch, dex, ny, nz, nx, dms, dms + h->nin, dy, h->st);
  if (r == LH_OK) {
    e = hipStreamSynchronize(h->st);
    if (e == hipSuccess) e = hipMemcpy(y, dy, (size_t)n * h->nout * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = mfail(h, LH_ERR_HIP, std::string("lh_mlp_predict_grid: ") + hipGetErrorString(e));
  }
  cleanup();
  return r;
}

}  // extern "C"

// ---- views for the closure coupling (lh_internal.hpp) ------------------------------------
constexpr int kEvalRows = 65536;   // rows per forward pass of the grid evaluation

int lh_mlp_dims(const lh_mlp* h, int* nin, int* nout) {
  if (!h) return LH_ERR_ARG;
  if (nin) *nin = h->nin;
  if (nout) *nout = h->nout;
  return LH_OK;
}

int lh_mlp_eval_grid_dev(lh_mlp* h, const float* g, int nch, const float* ex, int ny, int nz, int nx,
                         const float* mean, const float* rscale, float* y, hipStream_t st) {
  if (!h || !g || !mean || !rscale || !y || nch < 1 || ny < 1 || nz < 1 || nx < 1)
    return mfail(h, LH_ERR_ARG, "lh_mlp_eval_grid: bad argument");
  if (h->nin != 27 * nch + (ex ? 1 : 0))
    return mfail(h, LH_ERR_ARG, "lh_mlp_eval_grid: nin must be 27 nch (+1 with the extra field)");
  // the closure network's shape: widths known at compile time
  bool spec = h->L == 5 && h->Kp[0] == 256 && h->Np[h->L - 1] == 16 && nx % 16 == 0 && nch < 256 &&
              (long long)ny * nz * nx * nch < (1LL << 31);
  for (int l = 0; l + 1 < h->L; ++l) spec = spec && h->Np[l] == 64;
  if (spec) {
    FusedNet fn{};
    fn.L = h->L;
    fn.np = (int)h->np;
    fn.nin = h->nin;
    fn.nout = h->nout;
    for (int l = 0; l < h->L; ++l) {
      fn.Kp[l] = h->Kp[l];
      fn.Np[l] = h->Np[l];
      fn.offW[l] = h->offW[l];
      fn.offB[l] = h->offB[l];
    }
    constexpr int WN = 64 * 256 + 3 * 64 * 64 + 16 * 64;
    const size_t lds_s = (size_t)((2 * WN + 15) & ~15) + 6 * 256 * sizeof(int) + kSpecWaves * 16 * 64 * 2;
    if (!h->ncu_s) {
      int dev = 0;
      MCHK(h, hipGetDevice(&dev));
      MCHK(h, hipDeviceGetAttribute(&h->ncu_s, hipDeviceAttributeMultiprocessorCount, dev));
      MCHK(h, hipFuncSetAttribute((const void*)k_mlp_grid_s<256, 64, 3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024));
    }
    const long long n = (long long)ny * nz * nx;
    const long long tiles = (n + 15) / 16;
    const int blocks = (int)std::min<long long>((tiles + kSpecWaves - 1) / kSpecWaves, (long long)h->ncu_s);
    hipLaunchKernelGGL((k_mlp_grid_s<256, 64, 3>), dim3(blocks), dim3(64 * kSpecWaves), lds_s, st, fn, h->wb, h->p,
                       g, ex, mean, rscale, nch, ny, nz, nx, n, y);
    MCHK(h, hipGetLastError());
    return LH_OK;
  }
  // fused path: every width <= 64, weights and tiles within the LDS of a block
  bool fuse = h->L <= kFuseMaxL && h->Kp[0] <= kFuseMaxK0 && nch < 256;
  for (int l = 0; l < h->L; ++l) fuse = fuse && h->Np[l] <= 64 && (l == 0 || h->Kp[l] <= 64) && h->Kp[l] % 32 == 0;
  const size_t lds = (size_t)((2 * h->np + 15) & ~15LL) + 3 * (size_t)h->Kp[0] * sizeof(float) +
                     kFuseWaves * 16 * 64 * 2;
  fuse = fuse && lds <= 160 * 1024;
  if (fuse) {
    FusedNet fn{};
    fn.L = h->L;
    fn.np = (int)h->np;
    fn.nin = h->nin;
    fn.nout = h->nout;
    for (int l = 0; l < h->L; ++l) {
      fn.Kp[l] = h->Kp[l];
      fn.Np[l] = h->Np[l];
      fn.offW[l] = h->offW[l];
      fn.offB[l] = h->offB[l];
    }
    if (!h->ncu) {
      int dev = 0;
      MCHK(h, hipGetDevice(&dev));
      MCHK(h, hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, dev));
      MCHK(h, hipFuncSetAttribute((const void*)k_mlp_grid, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024));
    }
    const long long n = (long long)ny * nz * nx;
    const long long tiles = (n + 15) / 16;
    const int blocks = (int)std::min<long long>((tiles + kFuseWaves - 1) / kFuseWaves, 2LL * h->ncu);
    hipLaunchKernelGGL(k_mlp_grid, dim3(blocks), dim3(64 * kFuseWaves), lds, st, fn, h->wb, h->p, g, ex, mean, rscale, nch,
                       ny, nz, nx, n, y);
    MCHK(h, hipGetLastError());
    return LH_OK;
  }
  if (!h->Be) {
    for (int l = 0; l < h->L; ++l) {
      uint16_t* x = nullptr;
      float* z = nullptr;
      MCHK(h, zalloc(&x, (size_t)kEvalRows * h->Kp[l]));
      MCHK(h, zalloc(&z, (size_t)kEvalRows * h->Np[l]));
      h->Xe.push_back(x);
      h->Ze.push_back(z);
    }
    h->Be = kEvalRows;
  }
  const long long n = (long long)ny * nz * nx;
  for (long long r0 = 0; r0 < n; r0 += h->Be) {
    const int nr = (int)std::min<long long>(h->Be, n - r0);
    const int rows = (nr + 15) & ~15;
    hipLaunchKernelGGL(k_pack_grid, dim3(rows), dim3(64), 0, st, g, ex, mean, rscale, nch, ny, nz, nx, r0, nr,
                       h->nin, h->Kp[0], h->Xe[0]);
    forward_rows(h, rows, h->Xe.data(), h->Ze.data(), st);
    hipLaunchKernelGGL(k_unpack_out, dim3((nr * h->nout + 255) / 256), dim3(256), 0, st, h->Ze[h->L - 1],
                       h->Np[h->L - 1], h->nout, nr, y + (size_t)r0 * h->nout);
  }
  MCHK(h, hipGetLastError());
  return LH_OK;
}
