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
#include "lh_div.hpp"
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
  int maxn;             // planes of the largest job (the grid covers njobs x maxn planes)
};

#ifndef LH_Z_PAD
#define LH_Z_PAD 0
#endif
#ifndef LH_Z_LSP
#define LH_Z_LSP 2
#endif
#ifndef LH_Z_WPE
#define LH_Z_WPE 3
#endif
// FP32 z transforms of the long lengths (pz >= 384): 6 waves per SIMD, two 768-thread
// blocks per CU (as the FP32 z assembly)
#ifndef LH_Z_WPE32
#define LH_Z_WPE32 6
#endif
// LDS line stride of the z transforms (lines padded by one slot per 2^LH_Z_PAD elements);
// LH_Z_LSP = 2 puts the 8 columns of a 16-lane group on distinct banks.
template <int PZ> constexpr int zls() { return PadLen<PZ, LH_Z_PAD>::value + LH_Z_LSP; }
#ifndef LH_Z_XCD
#define LH_Z_XCD 1
#endif
// Block -> (column group, plane, job).  The grid is linear: with LH_Z_XCD the linear id is
// remapped so that consecutive column groups of a plane run on one XCD (blocks are dealt
// round-robin over the 8 XCDs).
struct ZBlock { int cg, p, job; };
__device__ __forceinline__ ZBlock z_block(int ncg, int maxn, int bid, int nblk) {
  if (LH_Z_XCD) {
    const int per = nblk / 8;
    if (bid < per * 8) bid = (bid % 8) * per + bid / 8;
  }
  ZBlock r;
  r.cg = bid % ncg;
  r.p = (bid / ncg) % maxn;
  r.job = bid / (ncg * maxn);
  return r;
}

// z transforms of L kx columns per block.  Threads run column-fastest (l = tid % L), so
// every global access of a wave covers L consecutive columns of a z row.  The first stage
// takes its inputs straight from global memory into registers and the last stage stores
// its outputs straight to global memory; only the middle stage (three-stage plans) reads
// and writes the LDS line in place.
template <int PZ, int V = 0> struct ZShape {
  using P = Plan<PZ, V>;
  static constexpr int T = P::T, R1 = P::RA, R2 = P::RB, R3 = P::RC;
  static constexpr bool THREE = R3 > 1;
  static constexpr int RL = THREE ? R3 : R2, NSL = PZ / RL;   // last stage
  static constexpr int NB1 = PZ / R1, PER1 = (NB1 + T - 1) / T;
  static constexpr int NB2 = PZ / R2, PER2 = (NB2 + T - 1) / T;
  static constexpr int PERL = (NSL + T - 1) / T;
};

// middle stage of a three-stage plan, in place on the line
template <int PZ, int V, bool INV, int PS, typename C, typename W>
__device__ __forceinline__ void z_middle(C* line, int t, const W& w) {
  using Z = ZShape<PZ, V>;
  C v[Z::PER2][Z::R2];
#pragma unroll
  for (int q = 0; q < Z::PER2; ++q) {
    const int b = t + q * Z::T;
    if (Z::PER2 * Z::T == Z::NB2 || b < Z::NB2) {
#pragma unroll
      for (int r = 0; r < Z::R2; ++r) v[q][r] = line[lpad<PS>(b + r * Z::NB2)];
    }
  }
  lds_bar();
#pragma unroll
  for (int q = 0; q < Z::PER2; ++q) {
    const int b = t + q * Z::T;
    if (Z::PER2 * Z::T == Z::NB2 || b < Z::NB2)
      stage_regs_store<PZ, Z::R2, Z::R1, INV, PS>(v[q], b, q, w.s2, line);
  }
  lds_bar();
}
// last stage: butterfly q of this thread, outputs at positions b + r * NSL
template <int PZ, int V, bool INV, int PS, typename C, typename W>
__device__ __forceinline__ void z_last(const C* line, int b, int q, const W& w, C (&v)[ZShape<PZ, V>::RL]) {
  using Z = ZShape<PZ, V>;
  if constexpr (Z::THREE) stage_load_regs<PZ, Z::RL, Z::R1 * Z::R2, INV, PS>(line, b, q, w.s3, v);
  else stage_load_regs<PZ, Z::RL, Z::R1, INV, PS>(line, b, q, w.s2, v);
}

// inverse: spectral plane (nz x mx), zero-padded to pz in k, inverse cft along z
template <int PZ, int L, typename TR, int V>
__device__ __forceinline__ void zinv_body(const ZJobs& jobs, const Dims& dm, const cx<TR>* __restrict__ tw,
                                          int bid, int nblk) {
  using C = cx<TR>;
  using Z = ZShape<PZ, V>;
  constexpr int LS = zls<PZ>();
  constexpr int PS = LH_Z_PAD;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const ZBlock zb = z_block((dm.mx + L - 1) / L, jobs.maxn, bid, nblk);
  const ZJob jb = jobs.j[zb.job];
  const int p = zb.p;
  if (p >= jb.n) return;
  const int plane = jb.planes ? gp(jb.planes)[p] : jb.first + p;
  const int nz = dm.nz, mx = dm.mx, h = nz / 2;
  const int l = threadIdx.x % L, t = threadIdx.x / L;
  const int i = zb.cg * L + l;
  const bool ok = i < mx;
  const double2* src = jb.src + (size_t)plane * dm.ncol + i;
  C* line = lds + l * LS;
  typename Plan<PZ, V>::template Tw<C> w;
  w.load(t, tw);
  // position n of the padded line holds mode n (n < h), the zero padding, or mode n-(pz-nz)
#pragma unroll
  for (int q = 0; q < Z::PER1; ++q) {
    const int b = t + q * Z::T;
    if (Z::PER1 * Z::T == Z::NB1 || b < Z::NB1) {
      C v[Z::R1];
#pragma unroll
      for (int r = 0; r < Z::R1; ++r) {
        const int n = b + r * Z::NB1;
        const int k = n < h ? n : n - (PZ - nz);
        v[r] = mk<C>(0, 0);
        if (ok && (n < h || k >= h)) {
          const double2 s = gld(src, (size_t)k * mx);
          v[r] = mk<C>((TR)s.x, (TR)s.y);
        }
      }
      stage_regs_store<PZ, Z::R1, 1, true, PS>(v, b, q, NoTw{}, line);
    }
  }
  lds_bar();
  if constexpr (Z::THREE) z_middle<PZ, V, true, PS>(line, t, w);
  C* dst = reinterpret_cast<C*>(jb.dst) + (size_t)(jb.slot0 + p) * PZ * mx + i;
#pragma unroll
  for (int q = 0; q < Z::PERL; ++q) {
    const int b = t + q * Z::T;
    if (Z::PERL * Z::T == Z::NSL || b < Z::NSL) {
      C v[Z::RL];
      z_last<PZ, V, true, PS>(line, b, q, w, v);
      if (ok) {
#pragma unroll
        for (int r = 0; r < Z::RL; ++r) gst(dst, (size_t)(b + r * Z::NSL) * mx, v[r]);
      }
    }
  }
}

// forward: Z slot -> forward cft along z, x 1/pz, truncation to nz modes -> spectral FP64
template <int PZ, int L, typename TR, int V>
__device__ __forceinline__ void zfwd_body(const ZJobs& jobs, const Dims& dm, const cx<TR>* __restrict__ tw,
                                          int bid, int nblk) {
  using C = cx<TR>;
  using Z = ZShape<PZ, V>;
  constexpr int LS = zls<PZ>();
  constexpr int PS = LH_Z_PAD;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const ZBlock zb = z_block((dm.mx + L - 1) / L, jobs.maxn, bid, nblk);
  const ZJob jb = jobs.j[zb.job];
  const int p = zb.p;
  if (p >= jb.n) return;
  const int plane = jb.planes ? gp(jb.planes)[p] : jb.first + p;
  const int nz = dm.nz, mx = dm.mx, h = nz / 2;
  const int l = threadIdx.x % L, t = threadIdx.x / L;
  const int i = zb.cg * L + l;
  const bool ok = i < mx;
  const C* src = reinterpret_cast<const C*>(jb.src) + (size_t)(jb.slot0 + p) * PZ * mx + i;
  C* line = lds + l * LS;
  typename Plan<PZ, V>::template Tw<C> w;
  w.load(t, tw);
#pragma unroll
  for (int q = 0; q < Z::PER1; ++q) {
    const int b = t + q * Z::T;
    if (Z::PER1 * Z::T == Z::NB1 || b < Z::NB1) {
      C v[Z::R1];
#pragma unroll
      for (int r = 0; r < Z::R1; ++r) v[r] = ok ? gld(src, (size_t)(b + r * Z::NB1) * mx) : mk<C>(0, 0);
      stage_regs_store<PZ, Z::R1, 1, false, PS>(v, b, q, NoTw{}, line);
    }
  }
  lds_bar();
  if constexpr (Z::THREE) z_middle<PZ, V, false, PS>(line, t, w);
  double2* dst = reinterpret_cast<double2*>(jb.dst) + (size_t)plane * dm.ncol + i;
  const double sz = 1.0 / (double)PZ;
#pragma unroll
  for (int q = 0; q < Z::PERL; ++q) {
    const int b = t + q * Z::T;
    if (Z::PERL * Z::T == Z::NSL || b < Z::NSL) {
      C v[Z::RL];
      z_last<PZ, V, false, PS>(line, b, q, w, v);
      if (ok) {
#pragma unroll
        for (int r = 0; r < Z::RL; ++r) {
          const int kk = b + r * Z::NSL;          // retained: kk < h or kk >= pz - (nz - h)
          const int k = kk < h ? kk : kk - (PZ - nz);
          if (kk < h || k >= h) {
            double2 o;
            o.x = sz * (double)v[r].x;
            o.y = sz * (double)v[r].y;
            gst(dst, (size_t)k * mx, o);
          }
        }
      }
    }
  }
}

template <int PZ, int L, typename TR, int V = 0>
__global__ void __launch_bounds__((L * Plan<PZ, V>::T))
    __attribute__((amdgpu_waves_per_eu(sizeof(TR) == 4 && PZ >= 384 ? LH_Z_WPE32 : LH_Z_WPE)))
    k_zinv(ZJobs jobs, Dims dm, const cx<TR>* __restrict__ tw) {
  zinv_body<PZ, L, TR, V>(jobs, dm, tw, blockIdx.x, gridDim.x);
}
template <int PZ, int L, typename TR, int V = 0>
__global__ void __launch_bounds__((L * Plan<PZ, V>::T))
    __attribute__((amdgpu_waves_per_eu(sizeof(TR) == 4 && PZ >= 384 ? LH_Z_WPE32 : LH_Z_WPE)))
    k_zfwd(ZJobs jobs, Dims dm, const cx<TR>* __restrict__ tw) {
  zfwd_body<PZ, L, TR, V>(jobs, dm, tw, blockIdx.x, gridDim.x);
}

// Job tables: the launches of many small channels (lengths up to 144) as one launch per
// kernel, length and plan.  Job j covers blocks blk0 .. blk0 + nblk - 1 of the launch; block
// b works on the channel map[b] as block b - blk0 of that channel's own grid.  A single
// small channel launches the same kernels with a table of its own, so that it computes
// exactly what it computes inside a batch.
struct ZJobT {
  ZJobs jobs;
  Dims dm;
  const void* tw;
  int blk0, nblk;
};
template <int PZ, int L, typename TR, int V = 0>
__global__ void __launch_bounds__((L * Plan<PZ, V>::T)) k_zinv_jobs(const ZJobT* __restrict__ jt,
                                                                     const int* __restrict__ map) {
  const ZJobT& J = jt[map[blockIdx.x]];
  const Dims dm = J.dm;
  // the job list is read where it lies: a local copy, indexed at run time, would sit in
  // per-lane scratch memory
  zinv_body<PZ, L, TR, V>(J.jobs, dm, (const cx<TR>*)J.tw, blockIdx.x - J.blk0, J.nblk);
}
template <int PZ, int L, typename TR, int V = 0>
__global__ void __launch_bounds__((L * Plan<PZ, V>::T)) k_zfwd_jobs(const ZJobT* __restrict__ jt,
                                                                     const int* __restrict__ map) {
  const ZJobT& J = jt[map[blockIdx.x]];
  const Dims dm = J.dm;
  // the job list is read where it lies: a local copy, indexed at run time, would sit in
  // per-lane scratch memory
  zfwd_body<PZ, L, TR, V>(J.jobs, dm, (const cx<TR>*)J.tw, blockIdx.x - J.blk0, J.nblk);
}

// ------------------------------------------------------------------------------------
// x pass helpers: rows z0, z0+1 of a Z slot packed into one complex line
//   line = A + iB (A, B Hermitian half spectra) -> inverse -> Re = row z0, Im = row z0+1
// ------------------------------------------------------------------------------------

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
  // non-negative doubles order like their bit patterns; a NaN (exponent all ones, non-zero
  // mantissa) compares above +inf, so a NaN operand always wins the maximum
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// Maximum that keeps a NaN: fmax would drop it and report a finite maximum of a field that
// holds non-finite values.

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
  const IBEntry* ent;           // this launch's entries
  const unsigned char* masks;   // [nmask][pz][px]
  const double* scal;
  double bre;                   // bRK(kRK)/Re
  int paired;                   // 0: tip entries (rhs and Lu_y of one row pair per block)
                                // 1: other entries (two row pairs z0, z0+2 per block)
};

template <int PX, typename TR, int V>
__device__ __forceinline__ void xpass_ib_body(const XIB& a, const Dims& dm, const cx<TR>* __restrict__ tw,
                                              int bx, int by) {
  using C = cx<TR>;
  using P = Plan<PX, V>;
  constexpr int T = P::T;
  constexpr int LS = PX + 1;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const bool paired = a.paired != 0;
  const int z0 = (paired ? 4 : 2) * bx;
  IBEntry en;
  {
    const gptr<const int> e4 = gp(reinterpret_cast<const int*>(a.ent + by));
    en.slot = e4[0];
    en.lslot = e4[1];
    en.mask = e4[2];
  }
  const int mx = dm.mx, pz = dm.pz;
  const size_t slot = (size_t)pz * mx;
  C* z = reinterpret_cast<C*>(a.z);
  // line 0: rhs rows (z0, z0+1); line 1: Lu_y rows (z0, z0+1) of a tip entry, or the rhs
  // rows (z0+2, z0+3)
  const C* A = z + en.slot * slot + (size_t)z0 * mx;
  const C* B = paired ? A + 2 * mx : z + en.lslot * slot + (size_t)z0 * mx;
  const unsigned char* m0 = a.masks + ((size_t)en.mask * pz + z0) * dm.px;
  // Every load of the block is issued before any is used, so the block waits for memory
  // once: the canopy mask bytes of its rows as 32-bit words (px is a multiple of 4), and
  // the half spectra of the two lines, item (line, k) for k < mx.  The host holds
  // px = 3 nx / 2, so mx = px / 3 + 1 and 2 mx items cover both lines.
  constexpr int NT = 2 * T;
  constexpr int WR = PX / 4;                            // mask words per row
  constexpr int MW = (4 * WR + NT - 1) / NT;            // mask words per thread (4 rows)
  constexpr int NI = 2 * (PX / 3 + 1);                  // items of the two lines
  constexpr int MH = (NI + NT - 1) / NT;                // items per thread
  const int nw = (paired ? 4 : 2) * WR;
  const unsigned* mwp = reinterpret_cast<const unsigned*>(m0);
  unsigned mw[MW];
  {
    C va[MH], vb[MH];
#pragma unroll
    for (int q = 0; q < MW; ++q) mw[q] = gp(mwp)[min((int)threadIdx.x + q * NT, nw - 1)];
#pragma unroll
    for (int q = 0; q < MH; ++q) {
      const int it = min((int)threadIdx.x + q * NT, 2 * mx - 1);
      const int ln = it / mx, k = it % mx;
      const C* R = ln ? B : A;
      va[q] = gld(R, k);
      vb[q] = gld(R, mx + k);
    }
    // position k holds a + ib, position px - k its mirror (rft_buff.f packing of two rows)
#pragma unroll
    for (int q = 0; q < MH; ++q) {
      const int it = threadIdx.x + q * NT;
      if (it < 2 * mx) {
        const int ln = it / mx, k = it % mx;
        const C x = va[q], y = vb[q];
        if (k == 0) {
          lds[ln * LS] = mk<C>(x.x, y.x);
        } else {
          lds[ln * LS + k] = mk<C>(x.x - y.y, x.y + y.x);
          lds[ln * LS + PX - k] = mk<C>(x.x + y.y, -x.y + y.x);
        }
      }
    }
    // zero between the halves: positions mx .. px - mx of both lines
    const int nzr = PX - 2 * mx + 1;
    for (int e = threadIdx.x; e < 2 * nzr; e += NT) lds[(e / nzr) * LS + mx + e % nzr] = mk<C>(0, 0);
  }
  const int ln = threadIdx.x / T, t = threadIdx.x % T;
  typename P::template Tw<C> w;
  w.load(t, tw);                 // in flight across the barrier
  __syncthreads();
  P::template run<true>(lds + ln * LS, t, w);
  const double beta = a.bre * gp(a.scal)[S_DT];
  // mask word q covers bytes x0 .. x0 + 3 of block row r; row r is component r % 2 of
  // line r / 2 (paired: rhs = 0 there), or, on a tip plane (rows z0, z0 + 1 of line 0),
  // rhs = -beta Lu_y with Lu_y the same component of line 1.  Each thread writes single
  // components, so no two threads touch the same value.
  TR* ldsr = reinterpret_cast<TR*>(lds);
#pragma unroll
  for (int q = 0; q < MW; ++q) {
    const int wd = threadIdx.x + q * NT;
    if (wd < nw && mw[q] != 0u) {
      const int r = wd / WR, x0 = 4 * (wd % WR);
      const int c = r & 1, l0 = paired ? r >> 1 : 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if ((mw[q] >> (8 * b)) & 0xFFu) {
          const int o = 2 * (l0 * LS + x0 + b) + c;
          ldsr[o] = paired ? (TR)0 : (TR)(-beta * (double)ldsr[o + 2 * LS]);
        }
      }
    }
  }
  __syncthreads();
  P::template run<false>(lds + ln * LS, t, w);
  C* dst = z + en.slot * slot + (size_t)z0 * mx;
  const int nout = paired ? 2 : 1;
  for (int e = threadIdx.x; e < nout * mx; e += blockDim.x) {
    const int l2 = e / mx, k = e % mx;
    C va, vb;
    unpack_rows<PX>(lds + l2 * LS, k, va, vb);
    C* d = dst + (size_t)(2 * l2) * mx;
    gst(d, k, va);
    gst(d, mx + k, vb);
  }
}
template <int PX, typename TR, int V = 0>
__global__ void __launch_bounds__((2 * Plan<PX, V>::T))
    k_xpass_ib(XIB a, Dims dm, const cx<TR>* __restrict__ tw) {
  xpass_ib_body<PX, TR, V>(a, dm, tw, blockIdx.x, blockIdx.y);
}
// the canopy x pass from a job table (a channel's grid: gx row-pair blocks x entries)
struct XibJob {
  XIB a;
  Dims dm;
  const void* tw;
  int blk0, gx;
};
template <int PX, typename TR, int V = 0>
__global__ void __launch_bounds__((2 * Plan<PX, V>::T)) k_xpass_ib_jobs(const XibJob* __restrict__ jt,
                                                                          const int* __restrict__ map) {
  const XibJob& J = jt[map[blockIdx.x]];
  const int lb = blockIdx.x - J.blk0;
  const XIB a = J.a;
  const Dims dm = J.dm;
  xpass_ib_body<PX, TR, V>(a, dm, (const cx<TR>*)J.tw, lb % J.gx, lb / J.gx);
}

__device__ __forceinline__ double2 d2(double a, double b) { double2 r; r.x = a; r.y = b; return r; }
// Complex (double2) buffers through a global pointer (gptr): g2(p)[i] reads and writes
// element i as two doubles (double2 itself cannot be copied through one)
typedef double gd2v __attribute__((ext_vector_type(2)));
struct G2ref {
  __attribute__((address_space(1))) gd2v* q;
  __device__ __forceinline__ operator double2() const {
    const gd2v v = *q;
    return d2(v.x, v.y);
  }
  __device__ __forceinline__ const G2ref& operator=(double2 v) const {
    gd2v w;
    w.x = v.x;
    w.y = v.y;
    *q = w;
    return *this;
  }
};
struct G2 {
  __attribute__((address_space(1))) gd2v* p;
  __device__ __forceinline__ G2ref operator[](size_t i) const { return G2ref{p + i}; }
};
__device__ __forceinline__ G2 g2(const double2* p) {
  return G2{(__attribute__((address_space(1))) gd2v*)const_cast<double2*>(p)};
}
__device__ __forceinline__ double2 dadd(double2 a, double2 b) { return d2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dsub(double2 a, double2 b) { return d2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 drm(double s, double2 a) { return d2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 dim(double k, double2 a) { return d2(-k * a.y, k * a.x); }

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

// LU_build + LU_dec of the two shared y matrices (tridLU_3D_IB.f90:427-493, 601-640), one
// thread per grid.  The recurrence runs in registers (the previous row's scaled upper
// entry is the only carried value) with the operations in the order of LU_dec.
__device__ __forceinline__ void prep_body(const PREP& a, int t) {
  if (t > 1) return;
  const int nj = t == 0 ? a.nyu : a.nyv;
  const gptr<const double> dd = gp(t == 0 ? a.dyu2i : a.dyv2i);
  const gptr<double> ay = gp(t == 0 ? a.ayu : a.ayv);
  const double beta = a.bre * gp(a.scal)[S_DT];
  // row 0: (0, 1, gw(1)) on the u grid, Dirichlet (0, 1, 0) on the v grid
  double up = t == 0 ? a.gw0 : 0.0;
  double di = 1.0 / 1.0;
  ay[0] = 0.0;
  ay[1] = di;
  for (int r = 1; r < nj; ++r) {
    double lo, dg, upr;
    if (r < nj - 1) {
      lo = -beta * dd[3 * r + 0];
      dg = 1.0 - beta * dd[3 * r + 1];
      upr = -beta * dd[3 * r + 2];
    } else {                      // last row: (gw(2), 1, 0) on the u grid, (0, 1, 0) on v
      lo = t == 0 ? a.gw1 : 0.0;
      dg = 1.0;
      upr = 0.0;
    }
    const double a3 = up * di;    // upper of row r-1 scaled by its 1/diag
    ay[3 * (r - 1) + 2] = a3;
    di = 1.0 / (dg - lo * a3);
    ay[3 * r + 0] = lo;
    ay[3 * r + 1] = di;
    up = upr;
  }
  ay[3 * (nj - 1) + 2] = up;
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

// Row r of the Poisson operator D G + k^2 of one column (LU_buildP, tridLU_3D_IB.f90:497-571):
// lower lo, diagonal dg and upper up; p row r <-> plane j = r - d + 1, u/v slot r + 1.  Row 0
// of column 0 is the pin (0, 1/(yu(2)-yu(1))^2, 0).
struct PoisRow {
  double lo, dg, up;
};
// Row r of the Poisson operator from the metrics it touches: hu = dthdyu(r+1), hv0 = dthdyv(r),
// hv1 = dthdyv(r+1); dypin = yu(2+d) - yu(1+d), the row pinned in column 0.
__device__ __forceinline__ PoisRow poisson_row_v(int r, bool pin, double k2x, double k2z, int np, double dv,
                                                 double hu, double hv0, double hv1, double dypin) {
  // D_vec_b(j) = -dv*dthdyu(j+1), G_vec_b(j) = -dv*dthdyv(j), D_vec_t(j) = dv*dthdyu(j),
  // G_vec_t(j) = dv*dthdyv(j-1); in row r: DB(r-1) = -dv hu, GB(r-1) = -dv hv0, DT(r) = dv hu,
  // GT(r) = dv hv0, GB(r) = -dv hv1, GT(r+1) = dv hv1
  PoisRow w;
  if (r == 0) {
    w.lo = 0.0;
    if (pin) {
      w.dg = 1.0 / (dypin * dypin);
      w.up = 0.0;
    } else {
      w.dg = ((dv * hu) * (-dv * hv1)) + k2x + k2z;
      w.up = ((dv * hu) * (dv * hv1));
    }
  } else {
    w.lo = ((-dv * hu) * (-dv * hv0));
    if (r < np - 1) {
      w.dg = ((-dv * hu) * (dv * hv0)) + ((dv * hu) * (-dv * hv1)) + k2x + k2z;
      w.up = ((dv * hu) * (dv * hv1));
    } else {
      w.dg = ((-dv * hu) * (dv * hv0)) + k2x + k2z;
      w.up = 0.0;
    }
  }
  return w;
}
__device__ __forceinline__ PoisRow poisson_row(int r, bool pin, double k2x, double k2z, int np, double dv,
                                               const double* __restrict__ dthdyu,
                                               const double* __restrict__ dthdyv,
                                               const double* __restrict__ yu, int d) {
  return poisson_row_v(r, pin, k2x, k2z, np, dv, dthdyu[r + 1], dthdyv[r], dthdyv[r + 1],
                       yu[2 + d] - yu[1 + d]);
}

// LU_decP once at start (tridLU_3D_IB.f90:643-684): the modified uppers a3 = up / pivot,
// which the back substitution of every solve reads; the lower and the inverse pivots are
// formed again inside the forward sweep (k_pres) with the same operations.
struct BUILDP {
  double* dg3;                  // [nyp][ncol]
  const double *dthdyu, *dthdyv, *yu, *k2x, *k2z;
};

__global__ void __launch_bounds__(256) k_buildP(BUILDP a, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dm.ncol) return;
  const int i = c % dm.mx, k = c / dm.mx;
  const double k2x = a.k2x[i], k2z = a.k2z[k];
  const int np = dm.nyp;
  double a3p = 0.0;
  for (int r = 0; r < np; ++r) {
    const PoisRow w = poisson_row(r, c == 0, k2x, k2z, np, dm.ddthetavi, a.dthdyu, a.dthdyv, a.yu, dm.d);
    const double a2 = 1.0 / (w.dg - w.lo * a3p);
    a3p = w.up * a2;
    a.dg3[(size_t)r * dm.ncol + c] = a3p;
  }
}

// meanflow_ctP (littleharsh.f90:505-521): Qx of the (0,0) column, one block of 256 threads
// (strided partial sums, then a fixed-order tree: the same sum on every run)
// flowrateRe of u1 into scal[S_QX], one block of 256 threads (part: 256 doubles of LDS)
__device__ __forceinline__ void flowrate_body(const double2* u1p, const double* yvp, double* scal, int ny,
                                              int d, int ncol, double* part) {
  const double2* u1 = u1p;
  const gptr<const double> yv = gp(yvp);
  double q = 0.0;
  for (int j = 1 + (int)threadIdx.x; j <= ny - 1; j += 256)
    q += (yv[j + d] - yv[j - 1 + d]) * u1[(size_t)(j + d) * ncol].x;
  part[threadIdx.x] = q;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) gp(scal)[S_QX] = part[0];
}

// error (littleharsh.f90:323-352): max |div| over the p grid
__global__ void __launch_bounds__(256) k_maxdiv(const double2* u1, const double2* u2, const double2* u3,
                                                const double* dthdyu, const double* k1x,
                                                const double* k1z, double* scal, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (c < dm.ncol) e = col_maxdiv(c, u1, u2, u3, dthdyu, k1x, k1z, dm.ddthetavi, dm.ncol, dm.mx, dm.nyp);
  for (int off = 32; off > 0; off >>= 1) e = nanmax(e, __shfl_down(e, off, 64));
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
