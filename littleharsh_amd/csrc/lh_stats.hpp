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

// v_itp (u planes) -> a, w_itp (v planes) -> b, wx (v planes) -> c
__global__ void __launch_bounds__(256) k_stats_prep(const double2* __restrict__ u2,
                                                    const double2* __restrict__ u3,
                                                    const double* __restrict__ dthdyv,
                                                    const double* __restrict__ k1z, double2* a,
                                                    double2* b, double2* c, Dims dm) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= dm.ncol) return;
  const size_t n = dm.ncol;
  const double kz = k1z[col / dm.mx];
  for (int s = 0; s < dm.nyu; ++s) {
    double2 vi = make_double2(0.0, 0.0);
    if (s >= 1 && s <= dm.nyu - 2) {
      const double2 v0 = u2[(s - 1) * n + col], v1 = u2[s * n + col];
      vi = make_double2(0.5 * (v0.x + v1.x), 0.5 * (v0.y + v1.y));
    }
    a[s * n + col] = vi;
  }
  for (int s = 0; s < dm.nyv; ++s) {
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

// local quantities of one point; returns how many
__device__ __forceinline__ int stats_q(int type, double a, double b, double c, double* q) {
  if (type == ST_U) {
    q[0] = a; q[1] = a * a; q[2] = b; q[3] = b * b; q[4] = a * c; q[5] = a * b;
    return 6;
  }
  if (type == ST_V) {
    q[0] = a; q[1] = a * a; q[2] = a * b; q[3] = c; q[4] = c * c;
    return 5;
  }
  q[0] = a; q[1] = a * a;
  return 2;
}

template <int PX, typename TR>
__global__ void __launch_bounds__(3 * Plan<PX>::T)
    k_stats(STATS a, Dims dm, const cx<double>* __restrict__ tw) {
  using C = cx<double>;
  constexpr int T = Plan<PX>::T;
  constexpr int LS = PX + 1;
  constexpr int NT = 3 * T;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  double* fold = reinterpret_cast<double*>(lds + 3 * LS);   // [6][dnx]
  double* foldp = fold + 6 * a.dnx;                          // [part][6][dnx], NT / dnx parts
  constexpr int NW = (NT + 63) / 64;
  __shared__ double red[6][NW];
  const int kk = blockIdx.x, p = blockIdx.y;
  const int mx = dm.mx, dnx = a.dnx;
  const size_t slot = (size_t)dm.pz * mx;
  for (int e = threadIdx.x; e < 6 * dnx; e += NT) fold[e] = 0.0;
  double s[6] = {0, 0, 0, 0, 0, 0};
  const int ln = threadIdx.x / T, t = threadIdx.x % T;
  typename Plan<PX>::template Tw<C> w;
  w.load(t, tw);
  for (int iz = 0; iz < a.ntz; iz += 2) {
    const int zA = kk + a.dnz * iz;
    const bool hasB = iz + 1 < a.ntz;
    const int zB = zA + a.dnz;
    // every load issued unconditionally (clamped to a valid element and masked afterwards),
    // so that the two rows wait for memory once
    constexpr int NE = (3 * PX + NT - 1) / NT;
    C ca[NE], cb[NE];
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = min((int)threadIdx.x + j * NT, 3 * PX - 1);
      const int f = min(e / PX, a.nf - 1), nn = e % PX;
      const int m = nn < mx ? nn : (nn > PX - mx ? PX - nn : 0);
      const C* base = reinterpret_cast<const C*>(a.z) + (size_t)(f * a.n + p) * slot;
      ca[j] = base[(size_t)zA * mx + m];
      cb[j] = base[(size_t)(hasB ? zB : zA) * mx + m];
    }
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * NT;
      if (NE * NT == 3 * PX || e < 3 * PX) {
        const int f = e / PX, nn = e % PX;
        const bool lo = nn < mx, hi = !lo && nn > PX - mx;
        const C b = hasB ? cb[j] : mk<C>(0.0, 0.0);
        C v = mk<C>(0.0, 0.0);
        if (f < a.nf && nn == 0) v = mk<C>(ca[j].x, b.x);
        else if (f < a.nf && lo) v = mk<C>(ca[j].x - b.y, ca[j].y + b.x);
        else if (f < a.nf && hi) v = mk<C>(ca[j].x + b.y, -ca[j].y + b.x);
        lds[f * LS + nn] = v;
      }
    }
    lds_bar();
    Plan<PX>::template run<true>(lds + ln * LS, t, w);
    lds_bar();
    // plain sums (wx rows: x < nx, z < nz only)
    for (int x = threadIdx.x; x < PX; x += NT) {
      const C f0 = lds[x], f1 = lds[LS + x], f2 = lds[2 * LS + x];
      for (int r = 0; r < 2; ++r) {
        if (r == 1 && !hasB) break;
        const int zr = r ? zB : zA;
        double q[6];
        const int nq = stats_q(a.type, r ? f0.y : f0.x, r ? f1.y : f1.x, r ? f2.y : f2.x, q);
        const bool wxin = x < dm.nx && zr < dm.nz;
        for (int l = 0; l < nq; ++l)
          if (!(a.type == ST_V && l >= 3 && !wxin)) s[l] += q[l];
      }
    }
    // tile fold: tile column i = e mod dnx of part e / dnx, the parts taking every nparts-th
    // tile; the parts' sums are added in part order
    const int nparts = max(1, min(NT / dnx, a.ntx));
    for (int e = threadIdx.x; e < nparts * dnx; e += NT) {
      const int i = e % dnx, part = e / dnx;
      double acc[6] = {0, 0, 0, 0, 0, 0};
      for (int ix = part; ix < a.ntx; ix += nparts) {
        const int x = i + dnx * ix;
        const C f0 = lds[x], f1 = lds[LS + x], f2 = lds[2 * LS + x];
        for (int r = 0; r < 2; ++r) {
          if (r == 1 && !hasB) break;
          double q[6];
          const int nq = stats_q(a.type, r ? f0.y : f0.x, r ? f1.y : f1.x, r ? f2.y : f2.x, q);
          for (int l = 0; l < nq; ++l) acc[l] += q[l];
        }
      }
      for (int l = 0; l < 6; ++l) foldp[(part * 6 + l) * dnx + i] = acc[l];
    }
    lds_bar();
    for (int e = threadIdx.x; e < 6 * dnx; e += NT) {
      double v = fold[e];
      for (int part = 0; part < nparts; ++part) v += foldp[part * 6 * dnx + e];
      fold[e] = v;
    }
    lds_bar();
  }
  // block reduction of the plain sums: a shuffle tree inside each wavefront (lane 0 ends
  // with the sum of the wave's active lanes), then the waves' sums in wave order; the order
  // is fixed, so the result does not depend on scheduling
  {
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
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
  if (threadIdx.x < 6) {
    double v = 0.0;
    for (int q = 0; q < NW; ++q) v += red[threadIdx.x][q];
    a.psum[pk * 6 + threadIdx.x] = v;
  }
  for (int e = threadIdx.x; e < 6 * dnx; e += NT) a.pfold[pk * 6 * dnx + e] = fold[e];
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
