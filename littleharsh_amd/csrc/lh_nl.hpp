// lh_nl.hpp -- nonlinear-term kernels of the RK3 substep (gfx950).
//
// The nonlinear term of the reference lives in the absent FOU3D.f90 (call at
// littleharsh.f90:171); it is restated as the 3/2-dealiased divergence form on the
// staggered grid:
//   Nu1(j) = ikx uu(j) + ikz uw(j) + m_u(j) (uv(j) - uv(j-1))      u centres
//   Nu3(j) = ikx uw(j) + ikz ww(j) + m_u(j) (vw(j) - vw(j-1))      u centres
//   Nu2(j) = ikx uv(j) + ikz vw(j) + m_v(j) (vv(j+1) - vv(j))      v faces
// with m = ddthetavi*dthdy (divergence, littleharsh.f90:238-241), v averaged to centres and
// u, w to faces (theta midpoints), products of the zero-padded physical fields.
//
// Because the z transform is linear and ikx is constant along a z line, each Nu is
// F_z(A) + ikz F_z(B) with A, B combined before the z transform:
//   u:  A1 = ikx uu + m_u (uv - uv_prev),   B1 = uw
//   w:  A3 = ikx uw + m_u (vw - vw_prev),   B3 = ww
//   v:  A2 = ikx uv_prev + m_v (vv - vv_prev), B2 = vw_prev   (face j-1 at level j)
// The x pass (k_xraw2) produces the six raw products per level (their x spectra, two rows
// packed per complex transform); the forward z pass (k_zasm) separates the rows, forms A
// and B and transforms them into Nu_new.  RHS0 and the rhs of solveU/solveV are built
// inside the velocity sweeps (k_vel_rhs, lh_col.hpp).
#pragma once
#include <type_traits>
#include "lh_kernels.hpp"

// z assembly block order: 0 = field slowest, 1 = field fastest (the three blocks of one
// column group and level chunk, which share product rows, run on one XCD together)
#ifndef LH_ZASM_ORDER
#define LH_ZASM_ORDER 1
#endif

namespace lh {

// Nonlinear x pass: products uu_c(j), uw_c(j), ww_c(j), vv_c(j), uv_f(j), vw_f(j) of each
// level j, forward-transformed in x.  Rows z0, z0+1 (z0 even) share one complex transform;
// their row of the product buffer holds that packed spectrum P unscaled: P[k] at k and
// P[px-k] at mx + k for k = 0..mx-1 (the z assembly separates the two rows).
struct XRAW {
  const void* zu;
  const void* zw;
  const void* zv;
  void* zp;                  // [6][level][pz/2][2 mx] packed row pairs
  const double* dthdyu;      // per u plane
  double* scal;
  int want_cfl;
};

// Sliding-window x pass: one row pair and a chunk of `lc` consecutive levels per block.
// Each thread owns fixed x positions of every line, so the physical U(j), W(j), V(j-1) of
// those positions stay in registers from one level to the next and a level costs three
// inverse transforms instead of six.  Levels are taken two at a time, so that all six
// line groups of the block transform together: six new lines (U, W at centres j+1, j+2 and
// V at faces j, j+1) are inverse-transformed, then the six products of level j and those
// of level j+1 are forward-transformed in turn.
#ifndef LH_XRAW_PAD
#define LH_XRAW_PAD 5
#endif
#ifndef LH_XRAW_LSP
#define LH_XRAW_LSP 1
#endif
#ifndef LH_XRAW2_WPE
#define LH_XRAW2_WPE 2
#endif
// FP32 lines take half the LDS, so two blocks would fit per CU with three waves per SIMD;
// measured, that is no faster than one block (the chunk model counts resident blocks)
#ifndef LH_XRAW2_WPE32
#define LH_XRAW2_WPE32 2
#endif
template <int PX, typename TR, int V = 0>
__global__ void __launch_bounds__((6 * Plan<PX, V>::T))
    __attribute__((amdgpu_waves_per_eu(sizeof(TR) == 4 ? LH_XRAW2_WPE32 : LH_XRAW2_WPE)))
    k_xraw2(XRAW a, Dims dm, const cx<TR>* __restrict__ tw, int lc) {
  using C = cx<TR>;
  using P = Plan<PX, V>;
  using ZS = ZShape<PX, V>;
  constexpr int T = P::T;
  constexpr int PS = LH_XRAW_PAD;                  // line padding (one slot per 2^PS)
  constexpr int LS = PadLen<PX, PS>::value + LH_XRAW_LSP;
  constexpr int NT = 6 * T;
  constexpr int NX = (PX + NT - 1) / NT;      // x positions per thread
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int z0 = 2 * blockIdx.x;               // row pair
  const int nlev = dm.nyu - 2;
  const int l0 = blockIdx.y * lc;               // level chunk
  const int l1 = min(l0 + lc, nlev);
  const int mx = dm.mx;
  const size_t slot = (size_t)dm.pz * mx;
  const C* zu = reinterpret_cast<const C*>(a.zu);
  const C* zw = reinterpret_cast<const C*>(a.zw);
  const C* zv = reinterpret_cast<const C*>(a.zv);
  C* zp = reinterpret_cast<C*>(a.zp);
  const int ln = tid / T, t = tid % T;
  // two sets of six lines (out-of-place transforms), three lines that keep the upper level
  // of a level pair until its turn, then the twiddle table
  C* const b0 = lds;
  C* const b1 = lds + 6 * LS;
  C* const keep = lds + 12 * LS;
  // An inverse transform of line f starts in in_line(f) and ends in out_line(f): lines 3..5
  // (the upper level of a pair) end in `keep` whatever the parity of the stage count.
  constexpr bool ODD = P::NSTAGE % 2 == 1;
  auto in_line = [&](int f) { return ODD || f < 3 ? b0 + f * LS : keep + (f - 3) * LS; };
  auto other_line = [&](int f) { return !ODD || f < 3 ? b1 + f * LS : keep + (f - 3) * LS; };
  C* const bres = ODD ? b1 : b0;               // where the transforms of lines 0..2 end
  C* const mine_a = in_line(ln);
  C* const mine_b = other_line(ln);
  typename P::template TwL<C> w;
  w.init(lds + 15 * LS, t, tw, ln == 0);

  // packed rows (z0, z0+1) of n Z slots -> line slots 0..n-1
  auto load_lines = [&](const C* const* src, int n) {
    constexpr int M = (6 * (PX / 3 + 1) + NT - 1) / NT;
    C va[M], vb[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int e = tid + q * NT;
      if (e < n * mx) {
        const C* A = src[e / mx] + (size_t)z0 * mx;
        va[q] = A[e % mx];
        vb[q] = A[mx + e % mx];
      }
    }
#pragma unroll
    for (int q = 0; q < M; ++q) {
      const int e = tid + q * NT;
      if (e < n * mx) {
        const int f = e / mx, k = e % mx;
        C* line = in_line(f);
        if (k == 0) {
          line[0] = mk<C>(va[q].x, vb[q].x);
        } else {
          line[lpad<PS>(k)] = mk<C>(va[q].x - vb[q].y, va[q].y + vb[q].x);
          line[lpad<PS>(PX - k)] = mk<C>(va[q].x + vb[q].y, -va[q].y + vb[q].x);
        }
      }
    }
    const int nzr = PX - 2 * mx + 1;
    for (int e = tid; e < n * nzr; e += NT) in_line(e / nzr)[lpad<PS>(mx + e % nzr)] = mk<C>(0, 0);
    lds_bar();
  };
  auto read_lines = [&](const C* src, C (*dst)[NX]) {
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int m = 0; m < NX; ++m) {
        const int x = tid + m * NT;
        if (x < PX) dst[f][m] = src[f * LS + lpad<PS>(x)];
      }
  };
  double cmax = 0.0;
  const TR half = (TR)0.5;
  // products of level l from old (U(j), W(j), V(j-1)) and new (U(j+1), W(j+1), V(j)) values,
  // forward transforms, and the six product rows out
  auto level_out = [&](int l, const C (*o)[NX], const C (*nw)[NX], C* pa, C* pb) {
    const double dyi = a.dthdyu[l + 1];
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      const int x = tid + m * NT;
      if (x < PX) {
        const C U = o[0][m], W = o[1][m], Vb = o[2][m];
        const C Ua = nw[0][m], Wa = nw[1][m], Vt = nw[2][m];
        const TR vc0 = half * (Vb.x + Vt.x), vc1 = half * (Vb.y + Vt.y);
        const TR uf0 = half * (U.x + Ua.x), uf1 = half * (U.y + Ua.y);
        const TR wf0 = half * (W.x + Wa.x), wf1 = half * (W.y + Wa.y);
        pa[0 * LS + lpad<PS>(x)] = mk<C>(U.x * U.x, U.y * U.y);
        pa[1 * LS + lpad<PS>(x)] = mk<C>(U.x * W.x, U.y * W.y);
        pa[2 * LS + lpad<PS>(x)] = mk<C>(W.x * W.x, W.y * W.y);
        pa[3 * LS + lpad<PS>(x)] = mk<C>(vc0 * vc0, vc1 * vc1);
        pa[4 * LS + lpad<PS>(x)] = mk<C>(uf0 * Vt.x, uf1 * Vt.y);
        pa[5 * LS + lpad<PS>(x)] = mk<C>(Vt.x * wf0, Vt.y * wf1);
        if (a.want_cfl) {
          const double s0 = fabs((double)U.x) * dm.dxi + fabs((double)vc0) * dyi + fabs((double)W.x) * dm.dzi;
          const double s1 = fabs((double)U.y) * dm.dxi + fabs((double)vc1) * dyi + fabs((double)W.y) * dm.dzi;
          cmax = nanmax(cmax, nanmax(s0, s1));
        }
      }
    }
    lds_bar();
    // all stages but the last ping-pong between pa and pb; the last one leaves P in registers
    C* const ha = pa + ln * LS;
    C* const hb = pb + ln * LS;
    stage_pp<PX, T, ZS::R1, 1, false, PS>(ha, hb, t, NoTw{});
    if constexpr (ZS::THREE) stage_pp<PX, T, ZS::R2, ZS::R1, false, PS>(hb, ha, t, w.s2);
    const C* const hl = ZS::THREE ? ha : hb;
    C* dst = zp + ((size_t)ln * nlev + l) * slot + (size_t)z0 * mx;
#pragma unroll
    for (int q = 0; q < ZS::PERL; ++q) {
      const int b = t + q * T;
      if (ZS::PERL * T == ZS::NSL || b < ZS::NSL) {
        C v[ZS::RL];
        z_last<PX, V, false, PS>(hl, b, q, w, v);
#pragma unroll
        for (int r = 0; r < ZS::RL; ++r) {
          const int k = b + r * ZS::NSL;
          if (k < mx) dst[k] = v[r];
          else if (k > PX - mx) dst[mx + (PX - k)] = v[r];
        }
      }
    }
  };

  // c0: physical U(j), W(j), V(j-1) of the current level; n1: those of the next one
  C c0[3][NX], n1[3][NX];
  {
    const int su = l0 + 1;
    const C* s[3] = {zu + su * slot, zw + su * slot, zv + (su - 1) * slot};
    load_lines(s, 3);
    P::template run_pp<true, PS>(mine_a, mine_b, t, w);
    read_lines(bres, c0);
    if constexpr (!ODD) lds_bar();   // odd: read b1 while the next writes go to b0
  }
  int l = l0;
  for (; l + 1 < l1; l += 2) {
    const int su = l + 1;
    const C* s[6] = {zu + (su + 1) * slot, zw + (su + 1) * slot, zv + su * slot,
                     zu + (su + 2) * slot, zw + (su + 2) * slot, zv + (su + 1) * slot};
    load_lines(s, 6);
    P::template run_pp<true, PS>(mine_a, mine_b, t, w);
    read_lines(bres, n1);
    if constexpr (!ODD) lds_bar();
    level_out(l, c0, n1, b0, b1);
    // the upper level waited in `keep`; it becomes the current level of the next pair
    read_lines(keep, c0);
    // with an odd stage count the last stage of level l reads b0: level l+1 starts in b1
    level_out(l + 1, n1, c0, ODD ? b1 : b0, ODD ? b0 : b1);
  }
  if (l < l1) {
    const int su = l + 1;
    const C* s[3] = {zu + (su + 1) * slot, zw + (su + 1) * slot, zv + su * slot};
    load_lines(s, 3);
    P::template run_pp<true, PS>(mine_a, mine_b, t, w);
    read_lines(bres, n1);
    if constexpr (!ODD) lds_bar();
    level_out(l, c0, n1, b0, b1);
  }
  if (a.want_cfl) {
    const int wv = tid / 64, lane = tid % 64;
    const int act = min(64, (int)blockDim.x - wv * 64);
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(cmax, off, 64);
      if (lane + off < act) cmax = nanmax(cmax, o);
    }
    if (lane == 0) red[wv] = cmax;
    __syncthreads();
    if (tid == 0) {
      double m = 0.0;
      for (int q = 0; q < (int)((blockDim.x + 63) / 64); ++q) m = nanmax(m, red[q]);
      atomic_max_pos(a.scal + S_CFLMAX, m);
    }
  }
}

// Forward z transform of the combined rows -> Nu_new: one block per (L consecutive kx
// columns, field, chunk of lc levels).  Walking up the levels, the upper product row of the
// y difference of level l is the lower one of level l+1 and stays in registers, so a level
// reads three product rows instead of four.  Block ids are remapped so that consecutive
// column groups run on one XCD.
struct ZASM {
  const void* zo;
  double2 *nu1, *nu2, *nu3;                  // Nu_new of this substep
  const double *dthdyu, *dthdyv, *k1x, *k1z;
  int nib;                                   // column blocks
  int lc;                                    // levels per block
};

// twiddle table of the z assembly's transforms
template <int PZ, typename C, int V = 0>
using ZasmTw = TwRowsPlan<PZ, Plan<PZ, V>::T, Plan<PZ, V>::RA, Plan<PZ, V>::RB, Plan<PZ, V>::RC, C>;

// items per load batch of the z assembly (more loads in flight against more registers)
#ifndef LH_ZASM_ZB
#define LH_ZASM_ZB 1
#endif
// FP32 lines of the long transforms (pz >= 384) at 6 waves per SIMD (80 VGPRs, a few
// spilled): two 768-thread blocks then share a CU (77 KB of LDS each), so one block's
// transforms overlap the other's loads (mixed precision at 576: 3.51 -> 3.27 ms per step)
#ifndef LH_ZASM_WPE32
#define LH_ZASM_WPE32 6
#endif
template <int PZ, int L, typename TR, int V = 0>
__global__ void __launch_bounds__((2 * L * Plan<PZ, V>::T))
    __attribute__((amdgpu_waves_per_eu(sizeof(TR) == 4 && PZ >= 384 ? LH_ZASM_WPE32 : 3)))
    k_zasm(ZASM a, Dims dm, const cx<TR>* __restrict__ tw) {
  using C = cx<TR>;
  using P = Plan<PZ, V>;
  constexpr int T = P::T;
  constexpr int NT = 2 * L * T;
  constexpr int LS = PZ + 1;
  constexpr int NPR = PZ / 2;                          // row pairs
  constexpr int ITZ = (L * NPR + NT - 1) / NT;         // (column, row pair) items per thread
  constexpr int NZ = 2 * PZ / 3;
  constexpr int ZB = LH_ZASM_ZB < ITZ ? LH_ZASM_ZB : ITZ;   // items per load batch
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x;
  const int nlev = dm.nyu - 2;
  // XCD-aware remap of the linear block id (8 XCDs, round-robin dispatch)
  const int nblk = gridDim.x;
  int bid = blockIdx.x;
  const int per = nblk / 8;
  if (bid < per * 8) bid = (bid % 8) * per + bid / 8;
  const int nch = (nlev + a.lc - 1) / a.lc;
  int ib, ch, f;                              // f: 0 u1, 1 u3, 2 v
  if (LH_ZASM_ORDER) {
    f = bid % 3;
    ib = (bid / 3) % a.nib;
    ch = bid / (3 * a.nib);
  } else {
    ib = bid % a.nib;
    ch = (bid / a.nib) % nch;
    f = bid / (a.nib * nch);
  }
  const int l0 = ch * a.lc;
  // face ny+d-1 is the wall: no Nu2 there
  const int l1 = min(l0 + a.lc, f == 2 ? nlev - 1 : nlev);
  if (l0 >= l1) return;
  const int i0 = ib * L;
  const int mx = dm.mx, nz = dm.nz, h = nz / 2, ncol = dm.ncol;
  const size_t slot = (size_t)PZ * mx;
  // raw products [6][nlev][pz][mx]: 0 uu, 1 uw, 2 ww, 3 vv, 4 uv, 5 vw (centre/face of level)
  // u centre j (level l):  A = ikx uu(l) + m_u (uv(l) - uv(l-1)),   B = uw(l)
  // w centre j (level l):  A = ikx uw(l) + m_u (vw(l) - vw(l-1)),   B = ww(l)
  // v face j   (level l):  A = ikx uv(l) + m_v (vv(l+1) - vv(l)),   B = vw(l)
  const C* zo = reinterpret_cast<const C*>(a.zo);
  const size_t ps = (size_t)nlev * slot;
  const int q0 = f == 0 ? 0 : (f == 1 ? 1 : 4);       // ikx term
  const int qd = f == 0 ? 4 : (f == 1 ? 5 : 3);       // y-difference term
  const int qb = f == 0 ? 1 : (f == 1 ? 2 : 5);       // B
  double2* __restrict__ nu = f == 0 ? a.nu1 : (f == 1 ? a.nu3 : a.nu2);
  const int ln = tid / T, t = tid % T;
  // twiddles: one LDS row per distinct twiddle set behind the 2L lines, shared by every
  // line (registers hold the carried rows); read after the barrier that precedes the
  // transform
  using TWP = ZasmTw<PZ, C, V>;
  TWP w;
  C* const twt = lds + 2 * L * LS;
  TWP::fill(twt, tw, tid, NT);
  w.init(twt, t);
  // The product rows hold the packed x spectrum P of each row pair (z, z+1): rows z and z+1
  // of column k are (P[k] + conj P[px-k]) / 2 and (P[k] - conj P[px-k]) / 2i, times 1/px.
  // (1/px)(x/2) = (1/(2px)) x exactly: halving is exact
  const TR sxh = (TR)0.5 * ((TR)1 / (TR)dm.px);
  // Item q of this thread: column i0 + l2 and row pair pr0 + q PST (NT is a multiple of L,
  // so every item of a thread has the same column).  Row pairs past the plane read the last
  // one (clamped) and are masked afterwards, so that every load of a level is issued
  // unconditionally and the level waits for memory once.  Column 0 reads its packed value
  // twice (slot mx + 0 is never written); the other columns read P[k] and P[px-k].
  static_assert(NT % L == 0, "z assembly: the block is a whole number of column groups");
  constexpr int PST = NT / L;
  const int l2 = tid % L, pr0 = tid / L;
  const bool col_ok = i0 + l2 < mx;
  const int col = min(i0 + l2, mx - 1);
  const bool k0 = col == 0;
  const int dn = k0 ? 0 : mx;
  auto pr_of = [&](int q) { return ITZ * PST == NPR ? pr0 + q * PST : min(pr0 + q * PST, NPR - 1); };
  // element e of a row at base b (uniform): a 32-bit byte offset from a scalar base, so that
  // the three rows of a level share each lane's offsets
  using GB = const char __attribute__((address_space(1)))*;
  typedef TR V2 __attribute__((ext_vector_type(2)));
  auto ld = [](GB b, unsigned e) {
    const V2 v = *reinterpret_cast<const V2 __attribute__((address_space(1)))*>(b + e * (unsigned)sizeof(C));
    return mk<C>(v.x, v.y);
  };
  // a row base held in scalar registers (readfirstlane), which also keeps the compiler from
  // folding each lane's offset into a 64-bit pointer of its own
  auto sbase = [](const void* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<GB>(((unsigned long long)hi << 32) | lo);
  };
  auto pr_ok = [&](int q) { return ITZ * PST == NPR || pr0 + q * PST < NPR; };
  const TR kx = (TR)a.k1x[col];
  auto split = [&](C zk, C zn, C& ra, C& rb) {
    ra = mk<C>(sxh * (zk.x + zn.x), sxh * (zk.y - zn.y));
    rb = mk<C>(sxh * (zk.y + zn.y), -(sxh * (zk.x - zn.x)));
    if (k0) { ra.y = 0; rb.y = 0; }
  };
  // lower operand of the y difference at the first level (rows 2p, 2p+1 of item q)
  C rl[ITZ][2];
  {
    const int l_lo = f < 2 ? l0 - 1 : l0;
    const GB row = sbase(zo + qd * ps + (size_t)max(l_lo, 0) * slot);
    C zk[ITZ], zn[ITZ];
#pragma unroll
    for (int q = 0; q < ITZ; ++q) {
      const unsigned o = pr_of(q) * 2 * mx + col;
      zk[q] = ld(row, o);
      zn[q] = ld(row, o + dn);
    }
#pragma unroll
    for (int q = 0; q < ITZ; ++q) {
      split(zk[q], zn[q], rl[q][0], rl[q][1]);
      if (!col_ok || l_lo < 0) rl[q][0] = rl[q][1] = mk<C>(0, 0);
    }
  }
  // Nu stores of this thread: mode k = pr0 + n PST of column i0 + l2
  constexpr int NW = (L * NZ + NT - 1) / NT;
  const double sz = 1.0 / (double)PZ;
  for (int l = l0; l < l1; ++l) {
    const int l_hi = f < 2 ? l : l + 1;
    const TR mlev = (TR)(dm.ddthetavi * (f < 2 ? a.dthdyu[l + 1] : a.dthdyv[l + 1]));
    const GB row0 = sbase(zo + q0 * ps + (size_t)l * slot);
    const GB rowb = sbase(zo + qb * ps + (size_t)l * slot);
    const GB rowh = sbase(zo + qd * ps + (size_t)l_hi * slot);
    // items in batches of ZB: the loads of a batch are issued together, then combined
#pragma unroll
    for (int q1 = 0; q1 < ITZ; q1 += ZB) {
      C z0[ZB][2], zb[ZB][2], zh[ZB][2];
#pragma unroll
      for (int u = 0; u < ZB; ++u) {
        const unsigned o = pr_of(min(q1 + u, ITZ - 1)) * 2 * mx + col, on = o + dn;
        z0[u][0] = ld(row0, o);
        z0[u][1] = ld(row0, on);
        zb[u][0] = ld(rowb, o);
        zb[u][1] = ld(rowb, on);
        zh[u][0] = ld(rowh, o);
        zh[u][1] = ld(rowh, on);
      }
#pragma unroll
      for (int u = 0; u < ZB; ++u) {
        const int q = q1 + u;
        const int pr = pr0 + q * PST;
        if (q < ITZ && pr_ok(q)) {
          C r0[2], rb[2], rh[2];
          split(z0[u][0], z0[u][1], r0[0], r0[1]);
          split(zb[u][0], zb[u][1], rb[0], rb[1]);
          split(zh[u][0], zh[u][1], rh[0], rh[1]);
#pragma unroll
          for (int h2 = 0; h2 < 2; ++h2) {
            C A = mk<C>(0, 0), B = mk<C>(0, 0);
            if (col_ok) {
              const C x0 = r0[h2], xh = rh[h2], xl = rl[q][h2];
              A = mk<C>(-kx * x0.y + mlev * (xh.x - xl.x), kx * x0.x + mlev * (xh.y - xl.y));
              B = rb[h2];
              rl[q][h2] = xh;
            }
            // A of column l2 in line l2, B in line L + l2: the eight columns of a write then
            // start LS * 4 dwords = 4 banks apart, eight distinct bank groups
            lds[l2 * LS + 2 * pr + h2] = A;
            lds[(L + l2) * LS + 2 * pr + h2] = B;
          }
        }
      }
    }
    __syncthreads();
    P::template run<false>(lds + ln * LS, t, w);
    // Nu = (F(A) + ikz F(B)) / pz on the retained modes
    const int s = l + 1;                      // u slot of centre j / v slot of face j
    double kzn[NW];
#pragma unroll
    for (int n = 0; n < NW; ++n) kzn[n] = a.k1z[min(pr0 + n * PST, NZ - 1)];
#pragma unroll
    for (int n = 0; n < NW; ++n) {
      const int k = pr0 + n * PST;
      if ((NW * PST == NZ || k < NZ) && col_ok) {
        const int kk = k < h ? k : k + (PZ - nz);
        const C A = lds[l2 * LS + kk], B = lds[(L + l2) * LS + kk];
        const double kz = kzn[n];
        typedef double D2 __attribute__((ext_vector_type(2)));
        const GB nus = sbase(nu + (size_t)s * ncol);
        D2 r;
        r.x = sz * ((double)A.x - kz * (double)B.y);
        r.y = sz * ((double)A.y + kz * (double)B.x);
        *const_cast<D2 __attribute__((address_space(1)))*>(
            reinterpret_cast<const D2 __attribute__((address_space(1)))*>(nus + (unsigned)(k * mx + i0 + l2) * 16u)) = r;
      }
    }
    __syncthreads();
  }
}

}  // namespace lh
