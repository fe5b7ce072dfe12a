// lh_stats.hpp -- statistics and spectra of the spectral state (stats.f90, spectra.f90).
//
// Statistics are sums over the physical collocation points of each y-plane.  The planes
// are made with the same z/x transforms as the nonlinear stage, in FP64:
//   u grid : u, w, v_itp  -> U, U2, W, W2, UV (u*v_itp), UW
//   v grid : v, w_itp, wx -> V, V2, VW (v*w_itp), wx, wx2
//   p grid : p            -> P, P2
// v_itp = (v(j-1) + v(j))/2 at the u centres (0 on the two ghost planes), w_itp =
// (w(j) + w(j+1))/2 at the v faces and wx = ddthetavi*dthdyv(j)*(w(j+1) - w(j)) - i kz v(j)
// are built in Fourier space (k_stats_prep).  Each block of k_stats owns one plane and one
// row class k of the dnx x dnz canopy tile: it transforms the ntilez rows z = k + dnz*iz
// and folds them onto the tile row k (conddown1/2, stats.f90:110-178), so every partial
// sum has one writer and the result does not depend on scheduling.  k_stats_fin adds the
// partials into the accumulators in a fixed order, the mirrored top band with the sign of
// v flipped (condup1/2, stats.f90:180-253) on top of the bottom band (recvXC,
// stats.f90:474-508).
#pragma once
#include "lh_kernels.hpp"

namespace lh {

enum { ST_U = 0, ST_V = 1, ST_P = 2 };

// v_itp (u planes) -> a, w_itp (v planes) -> b, wx (v planes) -> c; one thread per
// (column, plane s), planes on grid y
__global__ void __launch_bounds__(256) k_stats_prep(const double2* __restrict__ u2,
                                                    const double2* __restrict__ u3,
                                                    const double* __restrict__ dthdyv,
                                                    const double* __restrict__ k1z, double2* a,
                                                    double2* b, double2* c, Dims dm) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (col >= dm.ncol) return;
  const size_t n = dm.ncol;
  double2 vi = make_double2(0.0, 0.0);
  if (s >= 1 && s <= dm.nyu - 2) {
    const double2 v0 = u2[(s - 1) * n + col], v1 = u2[s * n + col];
    vi = make_double2(0.5 * (v0.x + v1.x), 0.5 * (v0.y + v1.y));
  }
  a[s * n + col] = vi;
  if (s < dm.nyv) {
    const double kz = k1z[col / dm.mx];
    const double2 w0 = u3[s * n + col], w1 = u3[(s + 1) * n + col], v = u2[s * n + col];
    b[s * n + col] = make_double2(0.5 * (w0.x + w1.x), 0.5 * (w0.y + w1.y));
    const double m = dm.ddthetavi * dthdyv[s];
    // dw/dy - i kz v
    c[s * n + col] = make_double2(m * (w1.x - w0.x) + kz * v.y, m * (w1.y - w0.y) - kz * v.x);
  }
}

struct STATS {
  const double2* z;   // Z slots: field f of plane p at slot f*n + p
  int nf, n, type;    // fields loaded, planes, ST_U / ST_V / ST_P
  int ntx, ntz, dnx, dnz;
  double* psum;       // [n][dnz][6]
  double* pfold;      // [n][dnz][6][dnx]
};

// local quantities of one point: q[0..5], zero beyond the type's count (6, 5 or 2), so
// that every loop over them has a fixed length and the arrays stay in registers
__device__ __forceinline__ void stats_q(int type, double a, double b, double c, double (&q)[6]) {
  q[0] = a;
  q[1] = a * a;
  if (type == ST_U) {
    q[2] = b; q[3] = b * b; q[4] = a * c; q[5] = a * b;
  } else if (type == ST_V) {
    q[2] = a * b; q[3] = c; q[4] = c * c; q[5] = 0.0;
  } else {
    q[2] = 0.0; q[3] = 0.0; q[4] = 0.0; q[5] = 0.0;
  }
}

// LDS lines of k_stats: one pad slot after every 2^LH_STATS_PAD values, so that the strided
// stage writes of the transforms spread over the banks
#ifndef LH_STATS_PAD
#define LH_STATS_PAD 5
#endif
template <int PX> constexpr int stats_ls() { return PadLen<PX, LH_STATS_PAD>::value + 1; }
// twiddles of the statistics transforms: one LDS table per block
template <int PX>
using StatsTw = TwRowsPlan<PX, Plan<PX>::T, Plan<PX>::RA, Plan<PX>::RB, Plan<PX>::RC, cx<double>>;
// parts of the canopy-tile fold: each of the 3 T threads owns one (tile column, part) slot
// when dnx <= 3 T, the parts taking every nparts-th tile
__host__ __device__ constexpr int stats_parts(int nt, int dnx, int ntx) {
  return nt / dnx < 1 ? 1 : (nt / dnx < ntx ? nt / dnx : ntx);
}
template <int PX> constexpr size_t stats_lds_bytes(int dnx, int ntx) {
  return (size_t)(3 * stats_ls<PX>() + StatsTw<PX>::SIZE) * sizeof(cx<double>) +
         (size_t)6 * dnx * (1 + stats_parts(3 * Plan<PX>::T, dnx, ntx)) * sizeof(double);
}

// One block per (plane, tile row class kk): the rows z = kk + dnz iz, two per pass as one
// complex line per field.  Per pass: the half spectra of the three fields are loaded by
// (field, k) item, all before any is used; the lines are transformed to physical space; then
// every thread takes its fold slots (tile column i, part) and adds the quantities of the
// points x = i + dnx ix of its part to its own partial of the tile fold (kept in LDS across
// the passes, so no barrier guards it) and to its plain sums.  The tiles cover the line
// (dnx ntilex = px), so the fold visits every x once.  At the end the parts are added in
// part order and the plain sums by a shuffle tree per wavefront, then in wave order.
template <int PX, typename TR>
__global__ void __launch_bounds__(3 * Plan<PX>::T)
    k_stats(STATS a, Dims dm, const cx<double>* __restrict__ tw) {
  using C = cx<double>;
  using P = Plan<PX>;
  using TWP = StatsTw<PX>;
  constexpr int T = P::T;
  constexpr int PS = LH_STATS_PAD;
  constexpr int LS = stats_ls<PX>();
  constexpr int NT = 3 * T;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  C* const twt = lds + 3 * LS;
  double* const fold = reinterpret_cast<double*>(twt + TWP::SIZE);   // [6][dnx]
  double* const foldp = fold + 6 * a.dnx;                             // [part][6][dnx]
  constexpr int NW = (NT + 63) / 64;
  __shared__ double red[6][NW];
  const int tid = threadIdx.x;
  const int kk = blockIdx.x, p = blockIdx.y;
  const int mx = dm.mx, dnx = a.dnx;
  const size_t slot = (size_t)dm.pz * mx;
  const int nparts = stats_parts(NT, dnx, a.ntx);
  for (int e = tid; e < nparts * 6 * dnx; e += NT) foldp[e] = 0.0;
  double s[6] = {0, 0, 0, 0, 0, 0};
  const int ln = tid / T, t = tid % T;
  TWP w;
  TWP::fill(twt, tw, tid, NT);
  w.init(twt, t);
  const C* const zbase = reinterpret_cast<const C*>(a.z);
  // the host holds px = 3 nx / 2, so mx = px / 3 + 1
  constexpr int NI = 3 * (PX / 3 + 1);
  constexpr int MI = (NI + NT - 1) / NT;
  const int nzr = PX - 2 * mx + 1;                    // zero positions mx .. px - mx
  for (int iz = 0; iz < a.ntz; iz += 2) {
    const int zA = kk + a.dnz * iz;
    const bool hasB = iz + 1 < a.ntz;
    const int zB = hasB ? zA + a.dnz : zA;
    {
      C ca[MI], cb[MI];
#pragma unroll
      for (int q = 0; q < MI; ++q) {
        const int it = min(tid + q * NT, 3 * mx - 1);
        const int f = min(it / mx, a.nf - 1), k = it % mx;
        const C* base = zbase + (size_t)(f * a.n + p) * slot;
        ca[q] = base[(size_t)zA * mx + k];
        cb[q] = base[(size_t)zB * mx + k];
      }
#pragma unroll
      for (int q = 0; q < MI; ++q) {
        const int it = tid + q * NT;
        if (it < 3 * mx) {
          const int f = it / mx, k = it % mx;
          const bool on = f < a.nf;
          const C x = on ? ca[q] : mk<C>(0.0, 0.0);
          const C y = on && hasB ? cb[q] : mk<C>(0.0, 0.0);
          C* line = lds + f * LS;
          if (k == 0) {
            line[0] = mk<C>(x.x, y.x);
          } else {
            line[lpad<PS>(k)] = mk<C>(x.x - y.y, x.y + y.x);
            line[lpad<PS>(PX - k)] = mk<C>(x.x + y.y, -x.y + y.x);
          }
        }
      }
      for (int e = tid; e < 3 * nzr; e += NT) lds[(e / nzr) * LS + lpad<PS>(mx + e % nzr)] = mk<C>(0.0, 0.0);
    }
    lds_bar();
    P::template run<true, PS>(lds + ln * LS, t, w);
    // fold slots of this thread; plain sums of wx (type V, quantities 3 and 4) only for
    // x < nx, z < nz
    for (int e = tid; e < nparts * dnx; e += NT) {
      const int i = e % dnx, part = e / dnx;
      double acc[6] = {0, 0, 0, 0, 0, 0};
      for (int ix = part; ix < a.ntx; ix += nparts) {
        const int x = i + dnx * ix;
        const int xp = lpad<PS>(x);
        const C f0 = lds[xp], f1 = lds[LS + xp], f2 = lds[2 * LS + xp];
        double q[6];
        stats_q(a.type, f0.x, f1.x, f2.x, q);
        const bool wxa = x < dm.nx && zA < dm.nz, wxb = x < dm.nx && zB < dm.nz;
#pragma unroll
        for (int l = 0; l < 6; ++l) {
          acc[l] += q[l];
          if (a.type != ST_V || l < 3 || wxa) s[l] += q[l];
        }
        if (hasB) {
          stats_q(a.type, f0.y, f1.y, f2.y, q);
#pragma unroll
          for (int l = 0; l < 6; ++l) {
            acc[l] += q[l];
            if (a.type != ST_V || l < 3 || wxb) s[l] += q[l];
          }
        }
      }
#pragma unroll
      for (int l = 0; l < 6; ++l) foldp[(part * 6 + l) * dnx + i] += acc[l];
    }
    lds_bar();
  }
  // the fold: parts added in part order
  for (int e = tid; e < 6 * dnx; e += NT) {
    double v = 0.0;
    for (int part = 0; part < nparts; ++part) v += foldp[part * 6 * dnx + e];
    fold[e] = v;
  }
  // block reduction of the plain sums: a shuffle tree inside each wavefront (lane 0 ends
  // with the sum of the wave's active lanes), then the waves' sums in wave order; the order
  // is fixed, so the result does not depend on scheduling
  {
    const int wv = tid / 64, lane = tid % 64;
    const int act = min(64, NT - wv * 64);
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      double v = s[l];
      for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(v, off, 64);
        if (lane + off < act) v += o;
      }
      if (lane == 0) red[l][wv] = v;
    }
  }
  __syncthreads();
  const size_t pk = (size_t)p * a.dnz + kk;
  if (tid < 6) {
    double v = 0.0;
    for (int q = 0; q < NW; ++q) v += red[tid][q];
    a.psum[pk * 6 + tid] = v;
  }
  for (int e = tid; e < 6 * dnx; e += NT) a.pfold[pk * 6 * dnx + e] = fold[e];
}

struct STFIN {
  const double* psum;
  const double* pfold;
  double* st;          // [13][nyu]
  double* cst;         // conditional accumulators
  long long qoff[13];  // offsets of the 13 conditional arrays
  int qmap[6];         // local -> global row
  double topsign[6];
  int nloc, n, type, dnx, dnz, nyu;
  int nb, toplo;       // planes per band on this grid, first top-band slot
};

__global__ void __launch_bounds__(256) k_stats_fin_plain(STFIN a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n * a.nloc) return;
  const int p = e / a.nloc, l = e % a.nloc;
  double v = 0.0;
  for (int k = 0; k < a.dnz; ++k) v += a.psum[((size_t)p * a.dnz + k) * 6 + l];
  a.st[(size_t)a.qmap[l] * a.nyu + p] += v;
}

__global__ void __launch_bounds__(256) k_stats_fin_cond(STFIN a) {
  const long long tile = (long long)a.dnx * a.dnz;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)a.nloc * a.nb * tile) return;
  const int l = (int)(e / (a.nb * tile));
  const long long r = e % (a.nb * tile);
  const int j1 = (int)(r / tile);
  const int kk = (int)((r % tile) / a.dnx), i = (int)(r % a.dnx);
  auto at = [&](int p) { return a.pfold[(((size_t)p * a.dnz + kk) * 6 + l) * a.dnx + i]; };
  double v = at(j1);
  const int ts = a.n - 1 - j1;
  if (ts >= a.toplo && ts > j1) v += a.topsign[l] * at(ts);
  a.cst[a.qoff[a.qmap[l]] + r] += v;
}

// spectra (spectra.f90:18-56): [5][nyu][ncol] accumulators U, V, W, UV, P
__global__ void __launch_bounds__(256) k_spectra(const double2* __restrict__ u1,
                                                 const double2* __restrict__ u2,
                                                 const double2* __restrict__ u3,
                                                 const double2* __restrict__ p, double* sp,
                                                 Dims dm) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (col >= dm.ncol) return;
  const size_t n = dm.ncol, pl = (size_t)dm.nyu * n, o = (size_t)s * n + col;
  const double2 a = u1[o], w = u3[o];
  sp[o] += a.x * a.x + a.y * a.y;
  sp[2 * pl + o] += w.x * w.x + w.y * w.y;
  double2 vc = make_double2(0.0, 0.0);
  if (s >= 1 && s <= dm.nyu - 2) {
    const double2 b0 = u2[o - n], b1 = u2[o];
    vc = make_double2(0.5 * (b0.x + b1.x), 0.5 * (b0.y + b1.y));
  }
  sp[3 * pl + o] += a.x * vc.x + a.y * vc.y;
  if (s < dm.nyv) {
    const double2 v = u2[o];
    sp[pl + o] += v.x * v.x + v.y * v.y;
  }
  if (s < dm.nyp) {
    const double2 q = p[o];
    sp[4 * pl + o] += q.x * q.x + q.y * q.y;
  }
}

}  // namespace lh
