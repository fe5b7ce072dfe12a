// lh_closure.hip -- the learned subfilter closure inside the coarse channel, on the GPU.
//
// The reference has no counterpart (the closure is project-defined).  Once per RK3 step,
// before the step, lh_closure_apply turns the coarse state into the forcing d_j tau_ij that
// every substep adds to Nu_new, with every stage on the channel's stream:
//   k_cl_sample  spectral u, v (mean of the two faces of a centre), w of each interior centre
//                plane -> the nz x nx coarse points shifted to the box-filter centres
//                (modes times exp(i (kx dx + kz dz))): an inverse DFT in z then in x, staged
//                in LDS (FP32, one block per plane and field);
//   k_cl_grad    the nine velocity gradients by central differences (x, z periodic; y
//                between the neighbouring centres, one-sided at the end planes) -- the
//                definitions the training inputs were built with;
//   network      3x3x3 stencil inputs gathered on the device, standardised, BF16 matrix-core
//                forward pass (lh_mlp_eval_grid_dev) -- or, for the Smagorinsky closure,
//                k_cl_smag from the gradients at the point;
//   k_cl_tau     de-standardised stresses -> spectral columns: a forward DFT in x then in z
//                scaled by 1/(nz nx), Nyquist row and column dropped, the shift undone;
//   (k_cl_sample_fft and k_cl_tau_fft replace the two DFT kernels when nx = nz is a length
//   with a Stockham plan in lh_fft.hpp: 32, 48, 64 or 96)
//   k_cl_div     f1 = i kx T_uu + i kz T_uw + dy T_uv and f3 = i kx T_uw + i kz T_ww + dy T_vw
//                at the centres; f2 = the x and z derivatives of T_uv, T_vw averaged to the
//                face plus (T_vv(m+1) - T_vv(m)) / (y(m+1) - y(m)) at the interior faces;
//                written into the channel's forcing buffers (FP64); with mean_only, only the
//                plane-mean column (the stresses' effect on the mean profile) is kept.
#include <hip/hip_runtime.h>
#include <math.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/littleharsh_amd.h"
#include "lh_fft.hpp"
#include "lh_internal.hpp"

namespace {

constexpr int kNStress = 6;   // uu, vv, ww, uv, uw, vw
constexpr int kNGrad = 9;     // d/dx, d/dy, d/dz of u, v, w

struct CL {
  int ny, nz, nx, mx, ncol;   // centre planes, coarse points (nx = 2 (mx - 1)), columns
  int mean_only;              // 1: only the plane-mean column (kx = kz = 0) of the forcing
  const float2* shift;        // exp(i (kx dx + kz dz)) per column
  const float2* twz;          // exp(2 pi i n / nz)
  const float2* twx;          // exp(2 pi i n / nx)
};

__device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float2 cmulc(float2 a, float2 b) {   // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// block (m, q): centre plane m, field q (0 u, 1 v, 2 w) -> phys[q][m] (nz x nx)
__global__ void __launch_bounds__(256) k_cl_sample(LhSolverView sv, CL c, float* __restrict__ phys) {
  extern __shared__ float2 sm[];
  const int nm = c.nz * c.mx;
  float2* B = sm;            // shifted modes, [kz row][kx]
  float2* T = B + nm;        // after the z pass, [z][kx]
  float2* tz = T + nm;
  float2* tx = tz + c.nz;
  const int m = blockIdx.x, q = blockIdx.y;
  for (int e = threadIdx.x; e < c.nz; e += blockDim.x) tz[e] = c.twz[e];
  for (int e = threadIdx.x; e < c.nx; e += blockDim.x) tx[e] = c.twx[e];
  for (int e = threadIdx.x; e < nm; e += blockDim.x) {
    double2 a;
    if (q == 0) {
      a = sv.u[(size_t)(m + 1) * c.ncol + e];
    } else if (q == 2) {
      a = sv.w[(size_t)(m + 1) * c.ncol + e];
    } else {
      const double2 p = sv.v[(size_t)m * c.ncol + e], r = sv.v[(size_t)(m + 1) * c.ncol + e];
      a = make_double2(0.5 * (p.x + r.x), 0.5 * (p.y + r.y));
    }
    B[e] = cmul(make_float2((float)a.x, (float)a.y), c.shift[e]);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nm; e += blockDim.x) {
    const int k = e / c.mx, cx = e % c.mx;
    float2 s = make_float2(0.f, 0.f);
    int idx = 0;                                       // (r k) mod nz
    for (int r = 0; r < c.nz; ++r) {
      const float2 t = cmul(B[r * c.mx + cx], tz[idx]);
      s.x += t.x;
      s.y += t.y;
      idx += k;
      if (idx >= c.nz) idx -= c.nz;
    }
    T[e] = s;
  }
  __syncthreads();
  const int np = c.nz * c.nx;
  float* out = phys + ((size_t)q * c.ny + m) * np;
  for (int e = threadIdx.x; e < np; e += blockDim.x) {
    const int k = e / c.nx, i = e % c.nx;
    float s = 0.f;
    int idx = 0;                                       // (kx i) mod nx
    for (int cx = 0; cx < c.mx; ++cx) {
      const float2 a = T[k * c.mx + cx], w = tx[idx];
      const float re = a.x * w.x - a.y * w.y;
      s += (cx == 0 || 2 * cx == c.nx) ? re : 2.f * re;
      idx += i;
      if (idx >= c.nx) idx -= c.nx;
    }
    out[e] = s;
  }
}

// one thread per coarse point: grad[3 q + d] for field q, direction d (x, y, z)
__global__ void k_cl_grad(CL c, const double* __restrict__ yc, float ix2, float iz2, const float* __restrict__ phys,
                          float* __restrict__ grad) {
  const long long n = (long long)c.ny * c.nz * c.nx;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int i = (int)(e % c.nx), k = (int)((e / c.nx) % c.nz), m = (int)(e / ((long long)c.nx * c.nz));
  const int ip = i + 1 == c.nx ? 0 : i + 1, im = i == 0 ? c.nx - 1 : i - 1;
  const int kp = k + 1 == c.nz ? 0 : k + 1, km = k == 0 ? c.nz - 1 : k - 1;
  const int mp = m + 1 < c.ny ? m + 1 : c.ny - 1, mm = m > 0 ? m - 1 : 0;
  const float iy = (float)(1.0 / (yc[mp] - yc[mm]));
  const size_t plane = (size_t)c.nz * c.nx;
  const size_t at = (size_t)m * plane + (size_t)k * c.nx;
  for (int q = 0; q < 3; ++q) {
    const float* f = phys + (size_t)q * n;
    grad[(size_t)(3 * q + 0) * n + e] = (f[at + ip] - f[at + im]) * ix2;
    grad[(size_t)(3 * q + 1) * n + e] = (f[(size_t)mp * plane + (size_t)k * c.nx + i] -
                                         f[(size_t)mm * plane + (size_t)k * c.nx + i]) * iy;
    grad[(size_t)(3 * q + 2) * n + e] = (f[(size_t)m * plane + (size_t)kp * c.nx + i] -
                                         f[(size_t)m * plane + (size_t)km * c.nx + i]) * iz2;
  }
}

// Smagorinsky stresses cs2 (-2 Delta^2 |S| S_ij) per point from the nine gradients
// (G[3 i + j] = d u_i / d x_j), rows of 6 in the order uu, vv, ww, uv, uw, vw; Delta^3 = dxc dzc dy
// with dy the centred spacing of the centres (one-sided at the end planes)
__global__ void k_cl_smag(CL c, const double* __restrict__ yc, double dxc, double dzc, double cs2,
                          const float* __restrict__ grad, float* __restrict__ y) {
  const long long n = (long long)c.ny * c.nz * c.nx;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int m = (int)(e / ((long long)c.nx * c.nz));
  const double dy = m == 0 ? yc[1] - yc[0] : (m == c.ny - 1 ? yc[m] - yc[m - 1] : 0.5 * (yc[m + 1] - yc[m - 1]));
  const double delta = cbrt(dxc * dzc * dy);
  float G[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) G[q] = grad[(size_t)q * n + e];
  float S[3][3], s2 = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      S[i][j] = 0.5f * (G[3 * i + j] + G[3 * j + i]);
      s2 += S[i][j] * S[i][j];
    }
  const float base = (float)(-2.0 * cs2 * delta * delta) * sqrtf(2.f * s2);
  float* o = y + (size_t)e * kNStress;
  o[0] = base * S[0][0];
  o[1] = base * S[1][1];
  o[2] = base * S[2][2];
  o[3] = base * S[0][1];
  o[4] = base * S[0][2];
  o[5] = base * S[1][2];
}

// block (m, q): stress q of centre plane m (network rows y, n x 6) -> T[q][m] (spectral columns)
__global__ void __launch_bounds__(256) k_cl_tau(CL c, const float* __restrict__ y, const float* __restrict__ ym,
                                                const float* __restrict__ ysd, float2* __restrict__ T) {
  extern __shared__ float2 sm[];
  const int nm = c.nz * c.mx, np = c.nz * c.nx;
  float2* S = sm;            // after the x pass, [z][kx]
  float2* tz = S + nm;
  float2* tx = tz + c.nz;
  float* f = reinterpret_cast<float*>(tx + c.nx);
  const int m = blockIdx.x, q = blockIdx.y;
  for (int e = threadIdx.x; e < c.nz; e += blockDim.x) tz[e] = c.twz[e];
  for (int e = threadIdx.x; e < c.nx; e += blockDim.x) tx[e] = c.twx[e];
  const float a = ysd[q], b = ym[q];
  for (int e = threadIdx.x; e < np; e += blockDim.x)
    f[e] = y[((size_t)m * np + e) * kNStress + q] * a + b;
  __syncthreads();
  for (int e = threadIdx.x; e < nm; e += blockDim.x) {
    const int k = e / c.mx, cx = e % c.mx;
    float2 s = make_float2(0.f, 0.f);
    int idx = 0;
    for (int i = 0; i < c.nx; ++i) {
      const float v = f[k * c.nx + i];
      const float2 w = tx[idx];
      s.x += v * w.x;
      s.y -= v * w.y;
      idx += cx;
      if (idx >= c.nx) idx -= c.nx;
    }
    S[e] = s;
  }
  __syncthreads();
  const float sc = 1.f / (float)(c.nz * c.nx);
  float2* out = T + ((size_t)q * c.ny + m) * c.ncol;
  for (int e = threadIdx.x; e < nm; e += blockDim.x) {
    const int r = e / c.mx, cx = e % c.mx;
    float2 s = make_float2(0.f, 0.f);
    if (2 * r != c.nz && 2 * cx != c.nx) {
      int idx = 0;
      for (int k = 0; k < c.nz; ++k) {
        const float2 t = cmulc(S[k * c.mx + cx], tz[idx]);
        s.x += t.x;
        s.y += t.y;
        idx += r;
        if (idx >= c.nz) idx -= c.nz;
      }
      s = cmulc(make_float2(s.x * sc, s.y * sc), c.shift[e]);
    }
    out[e] = s;
  }
}

// FFT forms of k_cl_sample and k_cl_tau for nx = nz = N (O(N log N) per line instead of
// O(N^2)).  Two sets of MX = N / 2 + 1 LDS lines of stride N + 1:
//   Z: line kx holds one spectral column along z (the z transforms);
//   X: line p holds the rows z = 2p and 2p + 1 as one complex line a + i b (the x
//      transforms of two real rows at once); its last line is zero and only keeps the
//      MX-th thread group in step with the others at the transforms' barriers.
// Block: MX groups of Plan<N>::T threads, one line per group.  tw: exp(-2 pi i m / N).
template <int N> constexpr int fft_threads() { return (N / 2 + 1) * lh::Plan<N>::T; }

template <int N>
__global__ void __launch_bounds__(fft_threads<N>()) k_cl_sample_fft(LhSolverView sv, CL c,
                                                                    const float2* __restrict__ tw,
                                                                    float* __restrict__ phys) {
  constexpr int MX = N / 2 + 1, LS = N + 1, T = lh::Plan<N>::T;
  extern __shared__ float2 sm[];
  float2* Z = sm;
  float2* X = Z + MX * LS;
  const int m = blockIdx.x, q = blockIdx.y;
  const int ln = threadIdx.x / T, t = threadIdx.x % T;
  typename lh::Plan<N>::template Tw<float2> w;
  w.load(t, tw);
  for (int e = threadIdx.x; e < N * MX; e += blockDim.x) {   // e = kz row * MX + kx
    double2 a;
    if (q == 0) {
      a = sv.u[(size_t)(m + 1) * c.ncol + e];
    } else if (q == 2) {
      a = sv.w[(size_t)(m + 1) * c.ncol + e];
    } else {
      const double2 p = sv.v[(size_t)m * c.ncol + e], r = sv.v[(size_t)(m + 1) * c.ncol + e];
      a = make_double2(0.5 * (p.x + r.x), 0.5 * (p.y + r.y));
    }
    Z[(e % MX) * LS + e / MX] = cmul(make_float2((float)a.x, (float)a.y), c.shift[e]);
  }
  __syncthreads();
  lh::Plan<N>::template run<true>(Z + ln * LS, t, w);        // inverse z transforms
  // Hermitian x lines of rows 2p (real part) and 2p + 1 (imaginary part); the zero and
  // Nyquist modes contribute their real parts, as in the one-sided sum of k_cl_sample
  for (int e = threadIdx.x; e < MX * N; e += blockDim.x) {
    const int p = e / N, n = e % N;
    float2 v = make_float2(0.f, 0.f);
    if (p < N / 2) {
      if (n == 0 || 2 * n == N) {
        v = make_float2(Z[n * LS + 2 * p].x, Z[n * LS + 2 * p + 1].x);
      } else if (2 * n < N) {
        const float2 a = Z[n * LS + 2 * p], b = Z[n * LS + 2 * p + 1];
        v = make_float2(a.x - b.y, a.y + b.x);
      } else {
        const float2 a = Z[(N - n) * LS + 2 * p], b = Z[(N - n) * LS + 2 * p + 1];
        v = make_float2(a.x + b.y, b.x - a.y);
      }
    }
    X[p * LS + n] = v;
  }
  __syncthreads();
  lh::Plan<N>::template run<true>(X + ln * LS, t, w);        // inverse x transforms
  float* out = phys + ((size_t)q * c.ny + m) * (N * N);
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const float2 v = X[(e / N / 2) * LS + e % N];
    out[e] = ((e / N) & 1) ? v.y : v.x;
  }
}

template <int N>
__global__ void __launch_bounds__(fft_threads<N>()) k_cl_tau_fft(CL c, const float* __restrict__ y,
                                                                 const float* __restrict__ ym,
                                                                 const float* __restrict__ ysd,
                                                                 const float2* __restrict__ tw,
                                                                 float2* __restrict__ T) {
  constexpr int MX = N / 2 + 1, LS = N + 1, TL = lh::Plan<N>::T;
  extern __shared__ float2 sm[];
  float2* X = sm;
  float2* Z = X + MX * LS;
  const int m = blockIdx.x, q = blockIdx.y;
  const int ln = threadIdx.x / TL, t = threadIdx.x % TL;
  typename lh::Plan<N>::template Tw<float2> w;
  w.load(t, tw);
  const float a = ysd[q], b = ym[q];
  const float* src = y + (size_t)m * (N * N) * kNStress + q;
  for (int e = threadIdx.x; e < MX * N; e += blockDim.x) {
    const int p = e / N, n = e % N;
    float2 v = make_float2(0.f, 0.f);
    if (p < N / 2)
      v = make_float2(src[(size_t)(2 * p * N + n) * kNStress] * a + b,
                      src[(size_t)((2 * p + 1) * N + n) * kNStress] * a + b);
    X[p * LS + n] = v;
  }
  __syncthreads();
  lh::Plan<N>::template run<false>(X + ln * LS, t, w);       // forward x transforms
  // the half spectra of the two rows of each line, into the z lines
  for (int e = threadIdx.x; e < (N / 2) * MX; e += blockDim.x) {
    const int p = e / MX, k = e % MX;
    const float2 zk = X[p * LS + k], zn = X[p * LS + (k == 0 ? 0 : N - k)];
    Z[k * LS + 2 * p] = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
    Z[k * LS + 2 * p + 1] = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
  }
  __syncthreads();
  lh::Plan<N>::template run<false>(Z + ln * LS, t, w);       // forward z transforms
  const float sc = 1.f / (float)(N * N);
  float2* out = T + ((size_t)q * c.ny + m) * c.ncol;
  for (int e = threadIdx.x; e < N * MX; e += blockDim.x) {   // e = kz row * MX + kx
    const int r = e / MX, cx = e % MX;
    float2 s = make_float2(0.f, 0.f);
    if (2 * r != N && 2 * cx != N) {
      const float2 v = Z[cx * LS + r];
      s = cmulc(make_float2(v.x * sc, v.y * sc), c.shift[e]);
    }
    out[e] = s;
  }
}

template <int N> constexpr size_t fft_lds() { return 2 * (size_t)(N / 2 + 1) * (N + 1) * sizeof(float2); }

// one thread per (column, centre plane): the divergence into the forcing buffers
__global__ void k_cl_div(LhSolverView sv, CL c, const double* __restrict__ yc, const float2* __restrict__ T) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)c.ncol * c.ny) return;
  const int e = (int)(g % c.ncol), m = (int)(g / c.ncol);
  if (c.mean_only && e != 0) {
    sv.f1[(size_t)(m + 1) * c.ncol + e] = make_double2(0.0, 0.0);
    sv.f3[(size_t)(m + 1) * c.ncol + e] = make_double2(0.0, 0.0);
    if (m + 1 < c.ny) sv.f2[(size_t)(m + 1) * c.ncol + e] = make_double2(0.0, 0.0);
    return;
  }
  const double kx = sv.k1x[e % c.mx], kz = sv.k1z[e / c.mx];
  auto at = [&](int q, int mm) {
    const float2 v = T[((size_t)q * c.ny + mm) * c.ncol + e];
    return make_double2(v.x, v.y);
  };
  const int mp = m + 1 < c.ny ? m + 1 : c.ny - 1, mm = m > 0 ? m - 1 : 0;
  const double iy = 1.0 / (yc[mp] - yc[mm]);
  const double2 uu = at(0, m), ww = at(2, m), uw = at(4, m);
  const double2 uvp = at(3, mp), uvm = at(3, mm), vwp = at(5, mp), vwm = at(5, mm);
  // i k a = (-k a.y, k a.x)
  const double2 f1 = make_double2(-kx * uu.y - kz * uw.y + (uvp.x - uvm.x) * iy,
                                  kx * uu.x + kz * uw.x + (uvp.y - uvm.y) * iy);
  const double2 f3 = make_double2(-kx * uw.y - kz * ww.y + (vwp.x - vwm.x) * iy,
                                  kx * uw.x + kz * ww.x + (vwp.y - vwm.y) * iy);
  sv.f1[(size_t)(m + 1) * c.ncol + e] = f1;
  sv.f3[(size_t)(m + 1) * c.ncol + e] = f3;
  if (m + 1 < c.ny) {
    const double2 uv0 = at(3, m), uv1 = at(3, m + 1), vw0 = at(5, m), vw1 = at(5, m + 1);
    const double2 vv0 = at(1, m), vv1 = at(1, m + 1);
    const double iyf = 1.0 / (yc[m + 1] - yc[m]);
    const double2 sx = make_double2(uv0.x + uv1.x, uv0.y + uv1.y), sz = make_double2(vw0.x + vw1.x, vw0.y + vw1.y);
    const double2 f2 = make_double2(0.5 * (-kx * sx.y - kz * sz.y) + (vv1.x - vv0.x) * iyf,
                                    0.5 * (kx * sx.x + kz * sz.x) + (vv1.y - vv0.y) * iyf);
    sv.f2[(size_t)(m + 1) * c.ncol + e] = f2;
  }
}

}  // namespace

struct lh_closure {
  lh_handle* h = nullptr;
  lh_mlp* net = nullptr;          // null: Smagorinsky with constant cs2
  double cs2 = 0.0;
  int nin = 0;
  LhSolverView sv{};
  CL c{};
  double dxc = 0.0, dzc = 0.0;
  double* yc = nullptr;
  float2 *shift = nullptr, *tw = nullptr;
  float *sf = nullptr, *stdz = nullptr, *phys = nullptr, *grad = nullptr, *y = nullptr;
  float2* T = nullptr;
  size_t lds_sample = 0, lds_tau = 0;
  int fft_n = 0;                // N of the FFT kernels (nx = nz = N), 0: the DFT kernels
  bool ready = false;           // set when creation completed; apply refuses a handle without it
  float2* twf = nullptr;        // exp(-2 pi i m / N) for the FFT kernels
  std::string err;
};

static int cfail(lh_closure* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}
#define CCHK(c, x)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return cfail(c, LH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" {

}  // extern "C"

template <int N> static hipError_t fft_attr() {
  hipError_t e = hipFuncSetAttribute((const void*)k_cl_sample_fft<N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)fft_lds<N>());
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_cl_tau_fft<N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)fft_lds<N>());
}

// shared set-up: channel view, shapes, shifts, twiddles, work buffers and the output scaling
// (ymean, yscale: 6 each); the network inputs' standardisation (nin each) comes first in stdz
static int closure_setup(lh_closure* c, lh_handle* h, const double* yc, double dxc, double dzc, double shift_x,
                         double shift_z, int nin, const float* xmean, const float* xscale, const float* ymean,
                         const float* yscale) {
  if (lh_solver_view(h, &c->sv) != LH_OK) return cfail(c, LH_ERR_HIP, "lh_closure_create: channel view failed");
  const LhSolverView& sv = c->sv;
  c->h = h;
  c->dxc = dxc;
  c->dzc = dzc;
  c->nin = nin;
  CL& cl = c->c;
  cl.ny = sv.nyu - 2;
  cl.nz = sv.nz;
  cl.mx = sv.mx;
  cl.nx = 2 * (sv.mx - 1);
  cl.ncol = sv.ncol;
  if (cl.ny < 2 || sv.nyv != sv.nyu - 1 || cl.nz % 2 || cl.nx < 2)
    return cfail(c, LH_ERR_ARG, "lh_closure_create: unsupported channel shape");
  // column shifts and twiddles (host, double) -> device (float)
  std::vector<double> k1x(cl.mx), k1z(cl.nz);
  CCHK(c, hipMemcpy(k1x.data(), sv.k1x, cl.mx * sizeof(double), hipMemcpyDeviceToHost));
  CCHK(c, hipMemcpy(k1z.data(), sv.k1z, cl.nz * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<float2> hs(cl.ncol), ht(cl.nz + cl.nx);
  for (int e = 0; e < cl.ncol; ++e) {
    const double ph = k1x[e % cl.mx] * shift_x + k1z[e / cl.mx] * shift_z;
    hs[e] = make_float2((float)std::cos(ph), (float)std::sin(ph));
  }
  for (int n = 0; n < cl.nz; ++n) {
    const double ph = 2.0 * M_PI * n / cl.nz;
    ht[n] = make_float2((float)std::cos(ph), (float)std::sin(ph));
  }
  for (int n = 0; n < cl.nx; ++n) {
    const double ph = 2.0 * M_PI * n / cl.nx;
    ht[cl.nz + n] = make_float2((float)std::cos(ph), (float)std::sin(ph));
  }
  const size_t npt = (size_t)cl.ny * cl.nz * cl.nx;
  CCHK(c, hipMalloc((void**)&c->yc, cl.ny * sizeof(double)));
  CCHK(c, hipMalloc((void**)&c->shift, cl.ncol * sizeof(float2)));
  CCHK(c, hipMalloc((void**)&c->tw, ht.size() * sizeof(float2)));
  CCHK(c, hipMalloc((void**)&c->stdz, (2 * (size_t)nin + 2 * kNStress) * sizeof(float)));
  CCHK(c, hipMalloc((void**)&c->phys, 3 * npt * sizeof(float)));
  CCHK(c, hipMalloc((void**)&c->grad, kNGrad * npt * sizeof(float)));
  CCHK(c, hipMalloc((void**)&c->y, kNStress * npt * sizeof(float)));
  CCHK(c, hipMalloc((void**)&c->T, (size_t)kNStress * cl.ny * cl.ncol * sizeof(float2)));
  CCHK(c, hipMemcpy(c->yc, yc, cl.ny * sizeof(double), hipMemcpyHostToDevice));
  CCHK(c, hipMemcpy(c->shift, hs.data(), hs.size() * sizeof(float2), hipMemcpyHostToDevice));
  CCHK(c, hipMemcpy(c->tw, ht.data(), ht.size() * sizeof(float2), hipMemcpyHostToDevice));
  if (nin > 0) {
    CCHK(c, hipMemcpy(c->stdz, xmean, nin * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> rs(nin);                     // the network kernels take 1 / scale
    for (int f = 0; f < nin; ++f) rs[f] = 1.0f / xscale[f];
    CCHK(c, hipMemcpy(c->stdz + nin, rs.data(), nin * sizeof(float), hipMemcpyHostToDevice));
  }
  CCHK(c, hipMemcpy(c->stdz + 2 * nin, ymean, kNStress * sizeof(float), hipMemcpyHostToDevice));
  CCHK(c, hipMemcpy(c->stdz + 2 * nin + kNStress, yscale, kNStress * sizeof(float), hipMemcpyHostToDevice));
  cl.shift = c->shift;
  cl.twz = c->tw;
  cl.twx = c->tw + cl.nz;
  const size_t nm = (size_t)cl.nz * cl.mx;
  c->lds_sample = (2 * nm + cl.nz + cl.nx) * sizeof(float2);
  c->lds_tau = (nm + cl.nz + cl.nx) * sizeof(float2) + (size_t)cl.nz * cl.nx * sizeof(float);
  if (c->lds_sample > 160 * 1024 || c->lds_tau > 160 * 1024)
    return cfail(c, LH_ERR_ARG, "lh_closure_create: coarse planes too large for LDS");
  CCHK(c, hipFuncSetAttribute((const void*)k_cl_sample, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)c->lds_sample));
  CCHK(c, hipFuncSetAttribute((const void*)k_cl_tau, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)c->lds_tau));
  if (cl.nx == cl.nz && (cl.nx == 32 || cl.nx == 48 || cl.nx == 64 || cl.nx == 96)) {
    const int n = cl.nx;
    std::vector<float2> hf(n);
    for (int k = 0; k < n; ++k) {
      const double ph = 2.0 * M_PI * k / n;
      hf[k] = make_float2((float)std::cos(ph), (float)-std::sin(ph));
    }
    CCHK(c, hipMalloc((void**)&c->twf, n * sizeof(float2)));
    CCHK(c, hipMemcpy(c->twf, hf.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    switch (n) {
      case 32: CCHK(c, fft_attr<32>()); break;
      case 48: CCHK(c, fft_attr<48>()); break;
      case 64: CCHK(c, fft_attr<64>()); break;
      default: CCHK(c, fft_attr<96>()); break;
    }
    c->fft_n = n;
  }
  return LH_OK;
}

// the closure's two spectral stages with the FFT kernels of length N
template <int N> static void launch_fft_sample(lh_closure* c, hipStream_t st) {
  hipLaunchKernelGGL(k_cl_sample_fft<N>, dim3(c->c.ny, 3), dim3(fft_threads<N>()), fft_lds<N>(), st, c->sv, c->c,
                     c->twf, c->phys);
}
template <int N> static void launch_fft_tau(lh_closure* c, hipStream_t st, const float* ym, const float* ysd) {
  hipLaunchKernelGGL(k_cl_tau_fft<N>, dim3(c->c.ny, kNStress), dim3(fft_threads<N>()), fft_lds<N>(), st, c->c, c->y,
                     ym, ysd, c->twf, c->T);
}

extern "C" {

int lh_closure_create(lh_handle* h, lh_mlp* net, const double* yc, double dxc, double dzc, double shift_x,
                      double shift_z, const float* sf, const float* xmean, const float* xscale,
                      const float* ymean, const float* yscale, lh_closure** out) {
  if (!out) return LH_ERR_ARG;
  *out = nullptr;
  lh_closure* c = new lh_closure();
  *out = c;
  if (!h || !net || !yc || !sf || !xmean || !xscale || !ymean || !yscale || !(dxc > 0.0) || !(dzc > 0.0))
    return cfail(c, LH_ERR_ARG, "lh_closure_create: null argument or non-positive spacing");
  int nin = 0, nout = 0;
  lh_mlp_dims(net, &nin, &nout);
  if (nin != 27 * kNGrad + 1 || nout != kNStress)
    return cfail(c, LH_ERR_ARG, "lh_closure_create: the network must map 244 inputs to 6 stresses");
  c->net = net;
  const int r = closure_setup(c, h, yc, dxc, dzc, shift_x, shift_z, nin, xmean, xscale, ymean, yscale);
  if (r != LH_OK) return r;
  const size_t npt = (size_t)c->c.ny * c->c.nz * c->c.nx;
  CCHK(c, hipMalloc((void**)&c->sf, npt * sizeof(float)));
  CCHK(c, hipMemcpy(c->sf, sf, npt * sizeof(float), hipMemcpyHostToDevice));
  c->ready = true;
  return LH_OK;
}

int lh_closure_create_smagorinsky(lh_handle* h, const double* yc, double dxc, double dzc, double shift_x,
                                  double shift_z, double cs2, lh_closure** out) {
  if (!out) return LH_ERR_ARG;
  *out = nullptr;
  lh_closure* c = new lh_closure();
  *out = c;
  if (!h || !yc || !(dxc > 0.0) || !(dzc > 0.0))
    return cfail(c, LH_ERR_ARG, "lh_closure_create_smagorinsky: null argument or non-positive spacing");
  c->cs2 = cs2;
  const float zero[kNStress] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, one[kNStress] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
  const int r = closure_setup(c, h, yc, dxc, dzc, shift_x, shift_z, 0, nullptr, nullptr, zero, one);
  c->ready = r == LH_OK;
  return r;
}

int lh_closure_apply(lh_closure* c) {
  if (!c) return LH_ERR_ARG;
  if (!c->ready || !c->h) return cfail(c, LH_ERR_ARG, "lh_closure_apply: the closure was not created successfully");
  if (lh_solver_view(c->h, &c->sv) != LH_OK) return cfail(c, LH_ERR_HIP, "lh_closure_apply: channel view failed");
  const CL& cl = c->c;
  const hipStream_t st = c->sv.st;
  const long long npt = (long long)cl.ny * cl.nz * cl.nx;
  const int nin = c->nin;
  switch (c->fft_n) {
    case 32: launch_fft_sample<32>(c, st); break;
    case 48: launch_fft_sample<48>(c, st); break;
    case 64: launch_fft_sample<64>(c, st); break;
    case 96: launch_fft_sample<96>(c, st); break;
    default: hipLaunchKernelGGL(k_cl_sample, dim3(cl.ny, 3), dim3(256), c->lds_sample, st, c->sv, cl, c->phys);
  }
  hipLaunchKernelGGL(k_cl_grad, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st, cl, c->yc,
                     (float)(0.5 / c->dxc), (float)(0.5 / c->dzc), c->phys, c->grad);
  CCHK(c, hipGetLastError());
  if (c->net) {
    if (lh_mlp_eval_grid_dev(c->net, c->grad, kNGrad, c->sf, cl.ny, cl.nz, cl.nx, c->stdz, c->stdz + nin, c->y,
                             st) != LH_OK)
      return cfail(c, LH_ERR_HIP, std::string("lh_closure_apply: network: ") + lh_mlp_last_error(c->net));
  } else {
    hipLaunchKernelGGL(k_cl_smag, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st, cl, c->yc, c->dxc, c->dzc,
                       c->cs2, c->grad, c->y);
  }
  const float *ym = c->stdz + 2 * nin, *ysd = c->stdz + 2 * nin + kNStress;
  switch (c->fft_n) {
    case 32: launch_fft_tau<32>(c, st, ym, ysd); break;
    case 48: launch_fft_tau<48>(c, st, ym, ysd); break;
    case 64: launch_fft_tau<64>(c, st, ym, ysd); break;
    case 96: launch_fft_tau<96>(c, st, ym, ysd); break;
    default: hipLaunchKernelGGL(k_cl_tau, dim3(cl.ny, kNStress), dim3(256), c->lds_tau, st, cl, c->y, ym, ysd, c->T);
  }
  hipLaunchKernelGGL(k_cl_div, dim3((unsigned)(((long long)cl.ncol * cl.ny + 255) / 256)), dim3(256), 0, st,
                     c->sv, cl, c->yc, c->T);
  CCHK(c, hipGetLastError());
  return LH_OK;
}

int lh_closure_set_mean_only(lh_closure* c, int on) {
  if (!c) return LH_ERR_ARG;
  c->c.mean_only = on ? 1 : 0;
  return LH_OK;
}

int lh_closure_destroy(lh_closure* c) {
  if (!c) return LH_OK;
  if (c->sv.st) (void)hipStreamSynchronize(c->sv.st);
  void* ptrs[] = {c->yc, c->shift, c->tw, c->twf, c->sf, c->stdz, c->phys, c->grad, c->y, c->T};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete c;
  return LH_OK;
}

const char* lh_closure_last_error(const lh_closure* c) { return c ? c->err.c_str() : "null handle"; }

}  // extern "C"
