// lh_kernels.hpp -- HIP kernels of one RK3 substep for gfx950.
//
// Device layout: every spectral field is [plane][column] complex FP64 with the column
// index c = k*mx + i (i = 0..nx/2 fastest).  A y-plane is contiguous (the x/z transforms
// work on it) and a column is a stride-ncol sequence whose consecutive columns sit on
// consecutive lanes (the y sweeps are coalesced).  The reference's column <-> plane
// transposes (modes_to_planes_*/planes_to_modes_*) therefore do not exist here.
//
// Intermediate "Z" buffers hold z-transformed data [slot][z][i], z = 0..pz-1, i < mx:
// the inverse z pass writes them, the x pass reads/writes rows, the forward z pass reads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "lh_fft.hpp"

namespace lh {

struct Dims {
  int nx, nz, px, pz, mx, ncol, ny, d, nyu, nyv, nyp;
  double dxi, dzi;      // 1/dx, 1/dz on the collocation grid
  double re, ddthetavi;
};

// device scalar slots
enum { S_DT = 0, S_MPGX = 1, S_QX = 2, S_CFLMAX = 3, S_DGX = 4, S_CFL = 5, S_MAXDIV = 6,
       S_QXT = 7, S_NSCAL = 8 };

struct RK {
  double a[3], b[3], g[3], c[3], d[3];
};

// ------------------------------------------------------------------------------------
// z transforms
// ------------------------------------------------------------------------------------
struct ZJob {
  const double2* src;   // spectral [plane][ncol] (inverse) or Z slots (forward)
  void* dst;            // Z slots (inverse) or spectral [plane][ncol] (forward)
  const int* planes;    // plane list (nullptr: first + p)
  int first;            // first plane when planes == nullptr
  int n;                // number of planes
  int slot0;            // first Z slot
};
struct ZJobs {
  ZJob j[4];
};

// inverse: spectral plane (nz x mx), zero-padded to pz in k, inverse cft along z
template <int PZ, int L, typename TR>
__global__ void __launch_bounds__(L* Plan<PZ>::T)
    k_zinv(ZJobs jobs, Dims dm, const cx<TR>* __restrict__ tw) {
  using C = cx<TR>;
  constexpr int LS = PZ + 1;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const ZJob jb = jobs.j[blockIdx.z];
  const int p = blockIdx.y;
  if (p >= jb.n) return;
  const int plane = jb.planes ? jb.planes[p] : jb.first + p;
  const int i0 = blockIdx.x * L;
  const double2* src = jb.src + (size_t)plane * dm.ncol;
  const int nz = dm.nz, mx = dm.mx, h = nz / 2;
  for (int e = threadIdx.x; e < L * LS; e += blockDim.x) lds[e] = mk<C>(0, 0);
  __syncthreads();
  for (int e = threadIdx.x; e < L * nz; e += blockDim.x) {
    const int l = e % L, k = e / L, i = i0 + l;
    if (i < mx) {
      const double2 v = src[(size_t)k * mx + i];
      const int kk = k < h ? k : k + (PZ - nz);
      lds[l * LS + kk] = mk<C>((TR)v.x, (TR)v.y);
    }
  }
  __syncthreads();
  const int l = threadIdx.x / Plan<PZ>::T, t = threadIdx.x % Plan<PZ>::T;
  Plan<PZ>::template run<true>(lds + l * LS, t, tw);
  C* dst = reinterpret_cast<C*>(jb.dst) + (size_t)(jb.slot0 + p) * PZ * mx;
  for (int e = threadIdx.x; e < L * PZ; e += blockDim.x) {
    const int l2 = e % L, z = e / L, i = i0 + l2;
    if (i < mx) dst[(size_t)z * mx + i] = lds[l2 * LS + z];
  }
}

// forward: Z slot -> forward cft along z, x 1/pz, truncation to nz modes -> spectral FP64
template <int PZ, int L, typename TR>
__global__ void __launch_bounds__(L* Plan<PZ>::T)
    k_zfwd(ZJobs jobs, Dims dm, const cx<TR>* __restrict__ tw) {
  using C = cx<TR>;
  constexpr int LS = PZ + 1;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const ZJob jb = jobs.j[blockIdx.z];
  const int p = blockIdx.y;
  if (p >= jb.n) return;
  const int plane = jb.planes ? jb.planes[p] : jb.first + p;
  const int i0 = blockIdx.x * L;
  const int nz = dm.nz, mx = dm.mx, h = nz / 2;
  const C* src = reinterpret_cast<const C*>(jb.src) + (size_t)(jb.slot0 + p) * PZ * mx;
  for (int e = threadIdx.x; e < L * PZ; e += blockDim.x) {
    const int l = e % L, z = e / L, i = i0 + l;
    lds[l * LS + z] = i < mx ? src[(size_t)z * mx + i] : mk<C>(0, 0);
  }
  __syncthreads();
  const int l = threadIdx.x / Plan<PZ>::T, t = threadIdx.x % Plan<PZ>::T;
  Plan<PZ>::template run<false>(lds + l * LS, t, tw);
  double2* dst = reinterpret_cast<double2*>(jb.dst) + (size_t)plane * dm.ncol;
  const double sz = 1.0 / (double)PZ;
  for (int e = threadIdx.x; e < L * nz; e += blockDim.x) {
    const int l2 = e % L, k = e / L, i = i0 + l2;
    if (i < mx) {
      const int kk = k < h ? k : k + (PZ - nz);
      const C v = lds[l2 * LS + kk];
      double2 o;
      o.x = sz * (double)v.x;
      o.y = sz * (double)v.y;
      dst[(size_t)k * mx + i] = o;
    }
  }
}

// ------------------------------------------------------------------------------------
// x pass helpers: rows z0, z0+1 of a Z slot packed into one complex line
//   line = A + iB (A, B Hermitian half spectra) -> inverse -> Re = row z0, Im = row z0+1
// ------------------------------------------------------------------------------------
template <int PX, typename C>
__device__ __forceinline__ C pack_rows(const C* __restrict__ A, const C* __restrict__ B, int n, int mx) {
  if (n == 0) return mk<C>(A[0].x, B[0].x);
  if (n < mx) {
    const C a = A[n], b = B[n];
    return mk<C>(a.x - b.y, a.y + b.x);
  }
  if (n > PX - mx) {
    const int m = PX - n;
    const C a = A[m], b = B[m];
    return mk<C>(a.x + b.y, -a.y + b.x);
  }
  return mk<C>(0, 0);
}

// forward unpack: Z_k of a+ib -> half spectra of a and b, x 1/px (rft_buff.f:63-73)
template <int PX, typename C>
__device__ __forceinline__ void unpack_rows(const C* line, int k, C& a, C& b) {
  using T = decltype(C::x);
  const C zk = line[k], zn = line[k == 0 ? 0 : PX - k];
  const T sx = (T)1 / (T)PX;
  a = mk<C>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y));
  b = mk<C>((T)0.5 * (zk.y + zn.y), -(T)0.5 * (zk.x - zn.x));
  a = mk<C>(sx * a.x, sx * a.y);
  b = mk<C>(sx * b.x, sx * b.y);
  if (k == 0) { a.y = 0; b.y = 0; }
}

__device__ __forceinline__ void atomic_max_pos(double* addr, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// Nonlinear x pass: one row pair and one level j (centre j and face j) per block.
//   lines 0..5 <- U(j), U(j+1), W(j), W(j+1), V(j-1), V(j)
//   products   -> uu_c(j), uw_c(j), ww_c(j), vv_c(j), uv_f(j), vw_f(j)
struct XNL {
  const void* zu;
  const void* zw;
  const void* zv;
  void* zp;                  // [6][level][pz][mx]
  const double* dthdyu;      // per u plane
  double* scal;
  int want_cfl;
};

template <int PX, typename TR>
__global__ void __launch_bounds__(6 * Plan<PX>::T)
    k_xpass_nl(XNL a, Dims dm, const cx<TR>* __restrict__ tw) {
  using C = cx<TR>;
  constexpr int T = Plan<PX>::T;
  constexpr int LS = PX + 1;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  __shared__ double red[16];
  const int z0 = 2 * blockIdx.x;
  const int level = blockIdx.y;            // centre / face j = level - d + 1
  const int su = level + 1;                // u slot of centre j
  const int mx = dm.mx, pz = dm.pz;
  const size_t slot = (size_t)pz * mx;
  const C* zu = reinterpret_cast<const C*>(a.zu);
  const C* zw = reinterpret_cast<const C*>(a.zw);
  const C* zv = reinterpret_cast<const C*>(a.zv);
  const C* src[6] = {zu + su * slot, zu + (su + 1) * slot, zw + su * slot, zw + (su + 1) * slot,
                     zv + (su - 1) * slot, zv + su * slot};
  for (int e = threadIdx.x; e < 6 * PX; e += blockDim.x) {
    const int ln = e / PX, n = e % PX;
    const C* A = src[ln] + (size_t)z0 * mx;
    lds[ln * LS + n] = pack_rows<PX>(A, A + mx, n, mx);
  }
  __syncthreads();
  const int ln = threadIdx.x / T, t = threadIdx.x % T;
  Plan<PX>::template run<true>(lds + ln * LS, t, tw);
  // products (each thread owns its x positions in all six lines)
  const TR half = (TR)0.5;
  const double dyi = a.dthdyu[su];
  double cmax = 0.0;
  for (int x = threadIdx.x; x < PX; x += blockDim.x) {
    const C U = lds[0 * LS + x], Ua = lds[1 * LS + x], W = lds[2 * LS + x], Wa = lds[3 * LS + x];
    const C Vb = lds[4 * LS + x], Vt = lds[5 * LS + x];
    C pr[6];
    {
      const TR vc0 = half * (Vb.x + Vt.x), vc1 = half * (Vb.y + Vt.y);
      const TR uf0 = half * (U.x + Ua.x), uf1 = half * (U.y + Ua.y);
      const TR wf0 = half * (W.x + Wa.x), wf1 = half * (W.y + Wa.y);
      pr[0] = mk<C>(U.x * U.x, U.y * U.y);
      pr[1] = mk<C>(U.x * W.x, U.y * W.y);
      pr[2] = mk<C>(W.x * W.x, W.y * W.y);
      pr[3] = mk<C>(vc0 * vc0, vc1 * vc1);
      pr[4] = mk<C>(uf0 * Vt.x, uf1 * Vt.y);
      pr[5] = mk<C>(Vt.x * wf0, Vt.y * wf1);
      if (a.want_cfl) {
        const double s0 = fabs((double)U.x) * dm.dxi + fabs((double)vc0) * dyi + fabs((double)W.x) * dm.dzi;
        const double s1 = fabs((double)U.y) * dm.dxi + fabs((double)vc1) * dyi + fabs((double)W.y) * dm.dzi;
        cmax = fmax(cmax, fmax(s0, s1));
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) lds[q * LS + x] = pr[q];
  }
  __syncthreads();
  Plan<PX>::template run<false>(lds + ln * LS, t, tw);
  C* zp = reinterpret_cast<C*>(a.zp);
  const size_t nlev = dm.nyu - 2;
  for (int e = threadIdx.x; e < 6 * mx; e += blockDim.x) {
    const int q = e / mx, k = e % mx;
    C va, vb;
    unpack_rows<PX>(lds + q * LS, k, va, vb);
    C* dst = zp + ((size_t)q * nlev + level) * slot + (size_t)z0 * mx;
    dst[k] = va;
    dst[mx + k] = vb;
  }
  if (a.want_cfl) {
    // block max (wave64 shuffles, then LDS across waves)
    for (int off = 32; off > 0; off >>= 1) cmax = fmax(cmax, __shfl_down(cmax, off, 64));
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) red[wv] = cmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = 0.0;
      for (int w = 0; w < (int)((blockDim.x + 63) / 64); ++w) m = fmax(m, red[w]);
      atomic_max_pos(a.scal + S_CFLMAX, m);
    }
  }
}

// Immersed-boundary x pass: rhs rows -> physical, rhs = 0 on canopy points, -beta*Lu_y on
// canopy points of tip planes, -> Fourier (tridLU_3D_IB.f90:224-245, 306-325)
struct IBEntry {
  int slot;      // Z slot of rhs (overwritten in place)
  int lslot;     // Z slot of Lu_y (tip planes) or -1
  int mask;      // mask index
  int pad;
};
struct XIB {
  void* z;
  const IBEntry* ent;
  const unsigned char* masks;   // [nmask][pz][px]
  const double* scal;
  double bre;                   // bRK(kRK)/Re
};

template <int PX, typename TR>
__global__ void __launch_bounds__(2 * Plan<PX>::T)
    k_xpass_ib(XIB a, Dims dm, const cx<TR>* __restrict__ tw) {
  using C = cx<TR>;
  constexpr int T = Plan<PX>::T;
  constexpr int LS = PX + 1;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int z0 = 2 * blockIdx.x;
  const IBEntry en = a.ent[blockIdx.y];
  const int mx = dm.mx, pz = dm.pz;
  const size_t slot = (size_t)pz * mx;
  C* z = reinterpret_cast<C*>(a.z);
  const C* A = z + en.slot * slot + (size_t)z0 * mx;
  const C* B = en.lslot >= 0 ? z + en.lslot * slot + (size_t)z0 * mx : A;
  const int nl = en.lslot >= 0 ? 2 : 1;
  for (int e = threadIdx.x; e < 2 * PX; e += blockDim.x) {
    const int ln = e / PX, n = e % PX;
    lds[ln * LS + n] = pack_rows<PX>(ln ? B : A, (ln ? B : A) + mx, n, mx);
  }
  __syncthreads();
  const int ln = threadIdx.x / T, t = threadIdx.x % T;
  Plan<PX>::template run<true>(lds + ln * LS, t, tw);
  const double beta = a.bre * a.scal[S_DT];
  const unsigned char* m0 = a.masks + ((size_t)en.mask * pz + z0) * dm.px;
  const unsigned char* m1 = m0 + dm.px;
  for (int x = threadIdx.x; x < PX; x += blockDim.x) {
    C v = lds[x];
    if (nl == 2) {
      const C l = lds[LS + x];
      if (m0[x]) v.x = (TR)(-beta * (double)l.x);
      if (m1[x]) v.y = (TR)(-beta * (double)l.y);
    } else {
      if (m0[x]) v.x = 0;
      if (m1[x]) v.y = 0;
    }
    lds[x] = v;
  }
  __syncthreads();
  Plan<PX>::template run<false>(lds + ln * LS, t, tw);
  C* dst = z + en.slot * slot + (size_t)z0 * mx;
  for (int k = threadIdx.x; k < mx; k += blockDim.x) {
    C va, vb;
    unpack_rows<PX>(lds, k, va, vb);
    dst[k] = va;
    dst[mx + k] = vb;
  }
}

// ------------------------------------------------------------------------------------
// RHS assembly (FP64): RHS0_u1/u2/u3 + nonlinear term + the rhs of solveU/solveV
// (littleharsh.f90:614-731, 733-824), divided by axz (tridLU_3D_IB.f90:90-98).
// One thread per (plane slot, column).
// ------------------------------------------------------------------------------------
struct ASM {
  const double2 *u1, *u2, *u3, *p;
  double2 *nu1, *nu2, *nu3;        // old in / new out, full plane ranges
  const double2* ph;               // [6][level][ncol] product spectra
  double2 *r1, *r2, *r3;           // rhs out
  double2* tipl;                   // [6][ncol] Lu_y on tip planes
  const double *dyu2i, *dyv2i, *dthdyu, *dthdyv;
  const double *k1x, *k2x, *k1z, *k2z;
  const double* scal;
  int tip_u0, tip_u1, tip_v0, tip_v1;   // slots
  double C1, C2, dR, g, bre;
};

__device__ __forceinline__ double2 d2(double a, double b) { double2 r; r.x = a; r.y = b; return r; }
__device__ __forceinline__ double2 dadd(double2 a, double2 b) { return d2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dsub(double2 a, double2 b) { return d2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 drm(double s, double2 a) { return d2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 dim(double k, double2 a) { return d2(-k * a.y, k * a.x); }

__global__ void __launch_bounds__(256) k_assemble(ASM a, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;          // plane slot (u grid 0..nyu-1; v grid 0..nyv-1)
  const int ncol = dm.ncol;
  if (c >= ncol) return;
  const int i = c % dm.mx, k = c / dm.mx;
  const double kx = a.k1x[i], kz = a.k1z[k], k2 = a.k2x[i] + a.k2z[k];
  const double dt = a.scal[S_DT];
  const double beta = a.bre * dt;
  const double axz = 1.0 - beta * (a.k2x[i] + a.k2z[k]);
  const size_t o = (size_t)s * ncol + c;
  const size_t nlev = dm.nyu - 2;
  const double2 zero = d2(0.0, 0.0);
  // ---------------- u and w at centres ----------------
  if (s == 0 || s == dm.nyu - 1) {
    a.r1[o] = zero;
    a.r3[o] = zero;
  } else {
    const int lev = s - 1;
    const double* w = a.dyu2i + 3 * s;
    const double2 p0 = a.p[(size_t)(s - 1) * ncol + c];
    const double m = dm.ddthetavi * a.dthdyu[s];
    const double2 UU = a.ph[(0 * nlev + lev) * (size_t)ncol + c];
    const double2 UW = a.ph[(1 * nlev + lev) * (size_t)ncol + c];
    const double2 WW = a.ph[(2 * nlev + lev) * (size_t)ncol + c];
    const double2 UV = a.ph[(4 * nlev + lev) * (size_t)ncol + c];
    const double2 VW = a.ph[(5 * nlev + lev) * (size_t)ncol + c];
    const double2 UVm = lev >= 1 ? a.ph[(4 * nlev + lev - 1) * (size_t)ncol + c] : zero;
    const double2 VWm = lev >= 1 ? a.ph[(5 * nlev + lev - 1) * (size_t)ncol + c] : zero;
    {
      const double2 um = a.u1[o - ncol], u0 = a.u1[o], up = a.u1[o + ncol];
      const double2 Lu = dadd(dadd(drm(w[0], um), drm(k2 + w[1], u0)), drm(w[2], up));
      double2 du = dadd(dadd(drm(a.C1, Lu), drm(a.C2, a.nu1[o])), dim(-a.dR * kx, p0));
      if (c == 0) du.x += -a.dR * a.scal[S_MPGX];
      const double2 n1 = dadd(dadd(dim(kx, UU), dim(kz, UW)), drm(m, dsub(UV, UVm)));
      a.nu1[o] = n1;
      du = dsub(du, drm(a.g, n1));
      double2 r = dadd(u0, drm(dt, du));
      r.x /= axz;
      r.y /= axz;
      a.r1[o] = r;
      if (s == a.tip_u0 || s == a.tip_u1) {
        const double2 Luy = dadd(dadd(drm(w[0], um), drm(w[1], u0)), drm(w[2], up));
        a.tipl[(size_t)(s == a.tip_u0 ? 0 : 1) * ncol + c] = Luy;
      }
    }
    {
      const double2 um = a.u3[o - ncol], u0 = a.u3[o], up = a.u3[o + ncol];
      const double2 Lu = dadd(dadd(drm(w[0], um), drm(k2 + w[1], u0)), drm(w[2], up));
      double2 du = dadd(dadd(drm(a.C1, Lu), drm(a.C2, a.nu3[o])), dim(-a.dR * kz, p0));
      const double2 n3 = dadd(dadd(dim(kx, UW), dim(kz, WW)), drm(m, dsub(VW, VWm)));
      a.nu3[o] = n3;
      du = dsub(du, drm(a.g, n3));
      double2 r = dadd(u0, drm(dt, du));
      r.x /= axz;
      r.y /= axz;
      a.r3[o] = r;
      if (s == a.tip_u0 || s == a.tip_u1) {
        const double2 Luy = dadd(dadd(drm(w[0], um), drm(w[1], u0)), drm(w[2], up));
        a.tipl[(size_t)(s == a.tip_u0 ? 2 : 3) * ncol + c] = Luy;
      }
    }
  }
  // ---------------- v at faces ----------------
  if (s < dm.nyv) {
    if (s == 0 || s == dm.nyv - 1) {
      a.r2[o] = zero;
    } else {
      const double* w = a.dyv2i + 3 * s;
      const double2 vm = a.u2[o - ncol], v0 = a.u2[o], vp = a.u2[o + ncol];
      const double2 Lv = dadd(dadd(drm(w[0], vm), drm(k2 + w[1], v0)), drm(w[2], vp));
      const double C3 = -a.dR * dm.ddthetavi * a.dthdyv[s];
      const double2 pd = dsub(a.p[(size_t)s * ncol + c], a.p[(size_t)(s - 1) * ncol + c]);
      double2 du = dadd(dadd(drm(a.C1, Lv), drm(a.C2, a.nu2[o])), drm(C3, pd));
      const int lev = s - 1;   // face j = s - d has level s-1; centres j, j+1 have levels s-1, s
      const double2 UV = a.ph[(4 * nlev + lev) * (size_t)ncol + c];
      const double2 VW = a.ph[(5 * nlev + lev) * (size_t)ncol + c];
      const double2 VVa = a.ph[(3 * nlev + lev + 1) * (size_t)ncol + c];
      const double2 VVb = a.ph[(3 * nlev + lev) * (size_t)ncol + c];
      const double m = dm.ddthetavi * a.dthdyv[s];
      const double2 n2 = dadd(dadd(dim(kx, UV), dim(kz, VW)), drm(m, dsub(VVa, VVb)));
      a.nu2[o] = n2;
      du = dsub(du, drm(a.g, n2));
      double2 r = dadd(v0, drm(dt, du));
      r.x /= axz;
      r.y /= axz;
      a.r2[o] = r;
      if (s == a.tip_v0 || s == a.tip_v1) {
        const double2 Lvy = dadd(dadd(drm(w[0], vm), drm(w[1], v0)), drm(w[2], vp));
        a.tipl[(size_t)(s == a.tip_v0 ? 4 : 5) * ncol + c] = Lvy;
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// Column stage
// ------------------------------------------------------------------------------------
// LU_build + LU_dec of the shared velocity matrices (tridLU_3D_IB.f90:427-493, 601-640)
struct PREP {
  const double *dyu2i, *dyv2i, *scal;
  double *ayu, *ayv;     // [rows][3]
  double gw0, gw1, bre;
  int nyu, nyv;
};

__global__ void k_prep(PREP a) {
  const int t = threadIdx.x;
  if (t > 1) return;
  const int nj = t == 0 ? a.nyu : a.nyv;
  const double* dd = t == 0 ? a.dyu2i : a.dyv2i;
  double* ay = t == 0 ? a.ayu : a.ayv;
  const double beta = a.bre * a.scal[S_DT];
  ay[0] = 0.0;
  ay[1] = 1.0;
  ay[2] = t == 0 ? a.gw0 : 0.0;
  for (int r = 1; r < nj - 1; ++r) {
    ay[3 * r + 0] = -beta * dd[3 * r + 0];
    ay[3 * r + 1] = 1.0 - beta * dd[3 * r + 1];
    ay[3 * r + 2] = -beta * dd[3 * r + 2];
  }
  const int r = nj - 1;
  ay[3 * r + 0] = t == 0 ? a.gw1 : 0.0;
  ay[3 * r + 1] = 1.0;
  ay[3 * r + 2] = 0.0;
  ay[1] = 1.0 / ay[1];
  for (int q = 1; q < nj; ++q) {
    ay[3 * (q - 1) + 2] = ay[3 * (q - 1) + 2] * ay[3 * (q - 1) + 1];
    ay[3 * q + 1] = 1.0 / (ay[3 * q + 1] - ay[3 * q + 0] * ay[3 * (q - 1) + 2]);
  }
}

// Thomas sweeps of LUsolU/LUsolV (tridLU_3D_IB.f90:106-116), one column per thread.
struct VEL {
  double2 *r1, *r2, *r3;   // rhs in (forward sweep results kept in place)
  double2 *u1, *u2, *u3;   // solution out
  const double *ayu, *ayv;
};

__device__ __forceinline__ void thomas_col(double2* x, double2* out, const double* ay, int nj,
                                           int ncol, int c) {
  double2 prev = drm(ay[1], x[c]);
  x[c] = prev;
  for (int r = 1; r < nj; ++r) {
    const size_t o = (size_t)r * ncol + c;
    prev = drm(ay[3 * r + 1], dsub(x[o], drm(ay[3 * r], prev)));
    x[o] = prev;
  }
  out[(size_t)(nj - 1) * ncol + c] = prev;
  for (int r = nj - 2; r >= 0; --r) {
    const size_t o = (size_t)r * ncol + c;
    prev = dsub(x[o], drm(ay[3 * r + 2], prev));
    out[o] = prev;
  }
}

__global__ void __launch_bounds__(256) k_vel(VEL a, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dm.ncol) return;
  const int f = blockIdx.y;
  if (f == 0) thomas_col(a.r1, a.u1, a.ayu, dm.nyu, dm.ncol, c);
  else if (f == 1) thomas_col(a.r3, a.u3, a.ayu, dm.nyu, dm.ncol, c);
  else thomas_col(a.r2, a.u2, a.ayv, dm.nyv, dm.ncol, c);
}

// flowrate_corr + LUsol0 (littleharsh.f90:453-485; tridLU_3D_IB.f90:398-424, 573-599,
// 686-703) on the (0,0) column; one thread.
struct CTU {
  double2* u1;
  double* scal;
  double* work;            // 4*(ny+1)
  const double *dyu2i, *yv;
  double gw0, gw1, bre, dR;
  int ny, d, ncol;
};

__global__ void k_ctU(CTU a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int n = a.ny;
  const double dt = a.scal[S_DT];
  double* u11 = a.work;
  double* m = a.work + (n + 1);
  const double bdt = -a.dR * dt;
  u11[0] = 0.0;
  for (int j = 1; j <= n - 1; ++j) u11[j] = bdt;
  u11[n] = 0.0;
  const double beta = a.bre * dt;
  m[0] = 0.0; m[1] = 1.0; m[2] = a.gw0;
  for (int j = 1; j <= n - 1; ++j) {
    const double* w = a.dyu2i + 3 * (j + a.d);
    m[3 * j] = -beta * w[0];
    m[3 * j + 1] = 1.0 - beta * w[1];
    m[3 * j + 2] = -beta * w[2];
  }
  m[3 * n] = a.gw1; m[3 * n + 1] = 1.0; m[3 * n + 2] = 0.0;
  m[1] = 1.0 / m[1];
  for (int j = 1; j <= n; ++j) {
    m[3 * (j - 1) + 2] = m[3 * (j - 1) + 2] * m[3 * (j - 1) + 1];
    m[3 * j + 1] = 1.0 / (m[3 * j + 1] - m[3 * j] * m[3 * (j - 1) + 2]);
  }
  u11[0] = u11[0] * m[1];
  for (int j = 1; j <= n; ++j) u11[j] = (u11[j] - m[3 * j] * u11[j - 1]) * m[3 * j + 1];
  for (int j = n - 1; j >= 0; --j) u11[j] = u11[j] - m[3 * j + 2] * u11[j + 1];
  double qx1 = 0.0, qx0 = 0.0;
  for (int j = 1; j <= n - 1; ++j) {
    const double dy = a.yv[j + a.d] - a.yv[j - 1 + a.d];
    qx1 += dy * u11[j];
    qx0 += dy * a.u1[(size_t)(j + a.d) * a.ncol].x;
  }
  const double g = (a.scal[S_QXT] - qx0) / qx1;
  a.scal[S_MPGX] = a.scal[S_MPGX] + g;
  a.scal[S_DGX] = g;
  for (int j = 0; j <= n; ++j) a.u1[(size_t)(j + a.d) * a.ncol].x += g * u11[j];
}

// LU_buildP + LU_decP once at start (tridLU_3D_IB.f90:497-571, 643-684)
struct BUILDP {
  double *dg1, *dg2, *dg3;      // [nyp][ncol]
  const double *dthdyu, *dthdyv, *yu, *k2x, *k2z;
};

__global__ void __launch_bounds__(256) k_buildP(BUILDP a, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dm.ncol) return;
  const int i = c % dm.mx, k = c / dm.mx;
  const double k2x = a.k2x[i], k2z = a.k2z[k];
  const int np = dm.nyp;
  const double dv = dm.ddthetavi;
  const int ncol = dm.ncol;
  // p row r <-> plane j = r - d + 1, u/v slot r + 1
#define DB(r) (-dv * a.dthdyu[(r) + 2])   /* D_vec_b(j) = -dv*dthdyu(j+1) */
#define GB(r) (-dv * a.dthdyv[(r) + 1])   /* G_vec_b(j) */
#define DT(r) (dv * a.dthdyu[(r) + 1])    /* D_vec_t(j) */
#define GT(r) (dv * a.dthdyv[(r) + 0])    /* G_vec_t(j) = dv*dthdyv(j-1) */
  double a1, a2, a3p;   // current row: lower, inverse pivot; a3p = modified upper of previous row
  // row 0
  double d0, u0;
  if (c == 0) {
    const double dy = a.yu[2 + dm.d] - a.yu[1 + dm.d];
    d0 = 1.0 / (dy * dy);
    u0 = 0.0;
  } else {
    d0 = (DT(0) * GB(0)) + k2x + k2z;
    u0 = (DT(0) * GT(1));
  }
  a2 = 1.0 / d0;
  a.dg1[c] = 0.0;
  a.dg2[c] = a2;
  a3p = u0 * a2;
  a.dg3[c] = a3p;
  for (int r = 1; r < np; ++r) {
    double diag, up;
    a1 = (DB(r - 1) * GB(r - 1));
    if (r < np - 1) {
      diag = (DB(r - 1) * GT(r)) + (DT(r) * GB(r)) + k2x + k2z;
      up = (DT(r) * GT(r + 1));
    } else {
      diag = (DB(r - 1) * GT(r)) + k2x + k2z;
      up = 0.0;
    }
    a2 = 1.0 / (diag - a1 * a3p);
    a3p = up * a2;
    a.dg1[(size_t)r * ncol + c] = a1;
    a.dg2[(size_t)r * ncol + c] = a2;
    a.dg3[(size_t)r * ncol + c] = a3p;
  }
#undef DB
#undef GB
#undef DT
#undef GT
}

// solveP + LUsolP + v_corr (littleharsh.f90:557-612, 826-929; tridLU_3D_IB.f90:362-396),
// one column per thread.
struct PRES {
  double2 *u1, *u2, *u3, *p, *psi;
  const double *dg1, *dg2, *dg3, *dthdyu, *dthdyv, *k1x, *k1z, *scal;
  double ab, gw0, gw1;       // aRK+bRK
};

__global__ void __launch_bounds__(256) k_pressure(PRES a, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = dm.ncol;
  if (c >= ncol) return;
  const int i = c % dm.mx, k = c / dm.mx;
  const double kx = a.k1x[i], kz = a.k1z[k];
  const double dt = a.scal[S_DT];
  const double dtirk = (1.0 / dt) / a.ab;
  const int np = dm.nyp;
  const double dv = dm.ddthetavi;
  // divergence, psi = dtirk*div, forward elimination
  double2 prev = d2(0.0, 0.0);
  for (int r = 0; r < np; ++r) {
    const size_t ou = (size_t)(r + 1) * ncol + c;
    const double2 dvg = dadd(dadd(dim(kx, a.u1[ou]), dim(kz, a.u3[ou])),
                             drm(dv * a.dthdyu[r + 1], dsub(a.u2[ou], a.u2[ou - ncol])));
    double2 ps = drm(dtirk, dvg);
    if (c == 0 && r == 0) ps = d2(0.0, 0.0);
    const size_t o = (size_t)r * ncol + c;
    if (r == 0) prev = drm(a.dg2[o], ps);
    else prev = drm(a.dg2[o], dsub(ps, drm(a.dg1[o], prev)));
    a.psi[o] = prev;
  }
  // back substitution fused with p += psi and v_corr
  const double dtrk = dt * a.ab;
  const double kxr = dtrk * kx, kzr = dtrk * kz, dtrk_u2 = dtrk * dv;
  double2 pnext = prev;   // psi(np-1)
  {
    const size_t o = (size_t)(np - 1) * ncol + c;
    a.p[o] = dadd(a.p[o], pnext);
    const size_t ou = (size_t)np * ncol + c;   // u slot np = last interior centre
    a.u1[ou] = dsub(a.u1[ou], dim(kxr, pnext));
    a.u3[ou] = dsub(a.u3[ou], dim(kzr, pnext));
  }
  for (int r = np - 2; r >= 0; --r) {
    const size_t o = (size_t)r * ncol + c;
    const double2 ps = dsub(a.psi[o], drm(a.dg3[o], pnext));
    a.p[o] = dadd(a.p[o], ps);
    const size_t ou = (size_t)(r + 1) * ncol + c;
    a.u1[ou] = dsub(a.u1[ou], dim(kxr, ps));
    a.u3[ou] = dsub(a.u3[ou], dim(kzr, ps));
    // v face slot r+1 lies between p rows r and r+1
    a.u2[ou] = dsub(a.u2[ou], drm(dtrk_u2 * a.dthdyv[r + 1], dsub(pnext, ps)));
    pnext = ps;
  }
  // ghost u/w and v walls (littleharsh.f90:866-878, 894-895, 917-929)
  const size_t top = (size_t)(dm.nyu - 1) * ncol + c;
  a.u1[c] = drm(-a.gw0, a.u1[c + ncol]);
  a.u1[top] = drm(-a.gw1, a.u1[top - ncol]);
  a.u3[c] = drm(-a.gw0, a.u3[c + ncol]);
  a.u3[top] = drm(-a.gw1, a.u3[top - ncol]);
  a.u2[c] = d2(0.0, 0.0);
  a.u2[(size_t)(dm.nyv - 1) * ncol + c] = d2(0.0, 0.0);
}

// meanflow_ctP (littleharsh.f90:505-521): Qx of the (0,0) column
__global__ void k_flowrate(const double2* u1, const double* yv, double* scal, int ny, int d, int ncol) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double q = 0.0;
  for (int j = 1; j <= ny - 1; ++j) q += (yv[j + d] - yv[j - 1 + d]) * u1[(size_t)(j + d) * ncol].x;
  scal[S_QX] = q;
}

// error (littleharsh.f90:323-352): max |div| over the p grid
__global__ void __launch_bounds__(256) k_maxdiv(const double2* u1, const double2* u2, const double2* u3,
                                                const double* dthdyu, const double* k1x,
                                                const double* k1z, double* scal, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (c < dm.ncol) {
    const int i = c % dm.mx, k = c / dm.mx;
    const double kx = k1x[i], kz = k1z[k];
    for (int r = 0; r < dm.nyp; ++r) {
      const size_t ou = (size_t)(r + 1) * dm.ncol + c;
      const double2 dvg = dadd(dadd(dim(kx, u1[ou]), dim(kz, u3[ou])),
                               drm(dm.ddthetavi * dthdyu[r + 1], dsub(u2[ou], u2[ou - dm.ncol])));
      e = fmax(e, hypot(dvg.x, dvg.y));
    }
  }
  for (int off = 32; off > 0; off >>= 1) e = fmax(e, __shfl_down(e, off, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_pos(scal + S_MAXDIV, e);
}

// small scalar kernels (graph-capturable replacements for host writes)
__global__ void k_set_scalar(double* scal, int idx, double v) {
  if (threadIdx.x == 0) scal[idx] = v;
}
__global__ void k_cfl_dt(double* scal) {
  if (threadIdx.x == 0) scal[S_DT] = scal[S_CFL] / scal[S_CFLMAX];
}

}  // namespace lh
