// lh_xw.hpp -- the nonlinear x pass with one wavefront per 576-point line (gfx950; FP64, and
// FP32 for the transforms of the mixed-precision mode).
//
// Each wave owns one row pair (z0, z0+1) of the Z slots and walks up a chunk of levels, as
// the block-level pass k_xraw2 does, but every 576-point transform is done by that wave
// alone: 9 points per lane, no workgroup barrier.  576 = 9 x 64:
//
//   inverse  x[n1 + 9 n2] = sum_t w64^{n2 t} ( w576^{n1 t} sum_r X[t + 64 r] w9^{n1 r} )
//
//   1. lane t loads X[t + 64 r], r = 0..8, straight from the Z rows (natural order) and
//      does the radix-9 DFT over r in registers, then the twiddle w576^{n1 t};
//   2. the 64-point DFTs over t of slots n1 = 0..7 run as 8 x 8 (t = t1 + 8 t2,
//      n2 = m2 + 8 m1) through one wave-private LDS line: a transpose, radix-8 over t2, the
//      twiddle w64^{m2 t1}, a second transpose, radix-8 over t1.  Lane L = (n1, m2) =
//      m2 + 8 n1 then holds x[n1 + 9 (m2 + 8 m1)] in slot m1;
//   3. slot n1 = 8 (the 72nd group of eight that the 64 lanes cannot share out) takes its
//      64-point DFT across the lanes: six radix-2 decimation-in-frequency stages with lane
//      exchanges, leaving x[8 + 9 bitrev6(t)] in lane t.
//
// The forward transform applies the transposed stages in reverse order (the DFT matrix is
// symmetric), so it consumes exactly the lane layout the inverse leaves and ends with
// X[t + 64 r] in lane t, slot r, ready for coalesced stores.  The products of the physical
// fields are pointwise and never leave the registers.
//
// A wave's LDS operations complete in issue order, so the transposes need no barrier; the
// empty asm statements only keep the compiler from moving memory accesses across them.
#pragma once
#include "lh_kernels.hpp"

namespace lh {
namespace xw {

constexpr int N = 576;
constexpr int MX = 193;                  // nx/2 + 1 for nx = 384
constexpr int TW576 = 0;                 // table: w576^{+n1 t} at n1*64 + t (n1 = 0..8)
constexpr int TW64 = 9 * 64;             // table: w64^{+j} at TW64 + j (j = 0..63)
constexpr int TWN = 9 * 64 + 64;         // table entries
constexpr int LINE = 8 * 72;             // wave-private LDS line (the larger transpose)

// C: double2 (FP64) or float2 (the FP32 transforms of the mixed-precision mode)
template <typename C> __device__ __forceinline__ C cm(C a, C b) { return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C> __device__ __forceinline__ C cmc(C a, C b) { return mk<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__device__ __forceinline__ void mem_fence_compiler() { asm volatile("" ::: "memory"); }
template <typename C> __device__ __forceinline__ C cadd2(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub2(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
// multiply by +i (INV) or -i
template <bool INV, typename C> __device__ __forceinline__ C ri(C a) { return INV ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x); }

// radix-3 in place on (a, b, c): a <- a+b+c, b <- a + w b + w^2 c, c <- a + w^2 b + w c,
// w = e^{+2 pi i/3} (INV) or e^{-2 pi i/3}
template <bool INV, typename C> __device__ __forceinline__ void r3(C& a, C& b, C& c) {
  using R = decltype(C::x);
  const R s3 = (R)0.86602540378443864676;
  const C t1 = cadd2(b, c);
  const C t2 = mk<C>(a.x - (R)0.5 * t1.x, a.y - (R)0.5 * t1.y);
  const C d = csub2(b, c);
  const C r = ri<INV>(mk<C>(s3 * d.x, s3 * d.y));
  a = cadd2(a, t1);
  b = cadd2(t2, r);
  c = csub2(t2, r);
}

// 9-point DFT in place, natural order in; X[k1 + 3 k2] ends in v[3 k1 + k2]
template <bool INV, typename C> __device__ __forceinline__ void dft9_t(C (&v)[9]) {
  r3<INV>(v[0], v[3], v[6]);
  r3<INV>(v[1], v[4], v[7]);
  r3<INV>(v[2], v[5], v[8]);
  // a[n2][k1] sits in v[n2 + 3 k1]; twiddle W9^{n2 k1}
  const C w1 = mk<C>(0.76604444311897803520, INV ? 0.64278760968653932632 : -0.64278760968653932632);
  const C w2 = mk<C>(0.17364817766693034885, INV ? 0.98480775301220805936 : -0.98480775301220805936);
  const C w4 = mk<C>(-0.93969262078590838405, INV ? 0.34202014332566873304 : -0.34202014332566873304);
  v[4] = cm(v[4], w1);
  v[7] = cm(v[7], w2);
  v[5] = cm(v[5], w2);
  v[8] = cm(v[8], w4);
  r3<INV>(v[0], v[1], v[2]);
  r3<INV>(v[3], v[4], v[5]);
  r3<INV>(v[6], v[7], v[8]);
}
// natural order of dft9_t's output: X[k] = v[perm9(k)], perm9(k1 + 3 k2) = 3 k1 + k2
__device__ __host__ constexpr int perm9(int k) { return 3 * (k % 3) + k / 3; }

// 8-point DFT in place (radix-2 decimation in frequency), natural order in; X[k] ends in
// v[bitrev3(k)]
template <bool INV, typename C> __device__ __forceinline__ void dft8_t(C (&v)[8]) {
  using R = decltype(C::x);
  const R h = (R)0.70710678118654752440;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const C a = cadd2(v[j], v[j + 4]);
    C b = csub2(v[j], v[j + 4]);
    if (j == 1) b = INV ? mk<C>(h * (b.x - b.y), h * (b.x + b.y)) : mk<C>(h * (b.x + b.y), h * (b.y - b.x));
    if (j == 2) b = ri<INV>(b);
    if (j == 3) b = INV ? mk<C>(-h * (b.x + b.y), h * (b.x - b.y)) : mk<C>(h * (b.y - b.x), -h * (b.x + b.y));
    v[j] = a;
    v[j + 4] = b;
  }
#pragma unroll
  for (int g = 0; g < 8; g += 4) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const C a = cadd2(v[g + j], v[g + j + 2]);
      C b = csub2(v[g + j], v[g + j + 2]);
      if (j == 1) b = ri<INV>(b);
      v[g + j] = a;
      v[g + j + 2] = b;
    }
  }
#pragma unroll
  for (int g = 0; g < 8; g += 2) {
    const C a = cadd2(v[g], v[g + 1]);
    const C b = csub2(v[g], v[g + 1]);
    v[g] = a;
    v[g + 1] = b;
  }
}
__device__ __host__ constexpr int brev3(int k) { return ((k & 1) << 2) | (k & 2) | ((k >> 2) & 1); }

// Value of lane t ^ H (t: this lane), for H = 1, 2, 8 by DPP (quad_perm [1,0,3,2], [2,3,0,1]; row_ror:8),
// 4 and 16 by ds_swizzle in bit mode (xor within 32 lanes, no memory access) and 32 by
// v_permlane32_swap, which exchanges the two halves of the wave.
template <int H> __device__ __forceinline__ double lane_xor(double v, int t) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  if constexpr (H == 1 || H == 2 || H == 8) {
    constexpr int ctl = H == 1 ? 0xB1 : (H == 2 ? 0x4E : 0x128);
    lo = __builtin_amdgcn_mov_dpp(lo, ctl, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(hi, ctl, 0xF, 0xF, false);
  } else if constexpr (H == 4 || H == 16) {
    constexpr int off = 0x1F | (H << 10);
    lo = __builtin_amdgcn_ds_swizzle(lo, off);
    hi = __builtin_amdgcn_ds_swizzle(hi, off);
  } else {
    static_assert(H == 32, "lane_xor: H in 1, 2, 4, 8, 16, 32");
    const bool upper = (t & 32) != 0;
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    lo = upper ? a[0] : a[1];
    hi = upper ? b[0] : b[1];
  }
  return __hiloint2double(hi, lo);
}
template <int H> __device__ __forceinline__ float lane_xor(float v, int t) {
  int x = __float_as_int(v);
  if constexpr (H == 1 || H == 2 || H == 8) {
    x = __builtin_amdgcn_mov_dpp(x, H == 1 ? 0xB1 : (H == 2 ? 0x4E : 0x128), 0xF, 0xF, false);
  } else if constexpr (H == 4 || H == 16) {
    x = __builtin_amdgcn_ds_swizzle(x, 0x1F | (H << 10));
  } else {
    static_assert(H == 32, "lane_xor: H in 1, 2, 4, 8, 16, 32");
    const auto a = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    x = (t & 32) != 0 ? a[0] : a[1];
  }
  return __int_as_float(x);
}
template <int H, typename C> __device__ __forceinline__ C lane_xor(C v, int t) {
  return mk<C>(lane_xor<H>(v.x, t), lane_xor<H>(v.y, t));
}

// slot 8: the 64-point DFT across the lanes, decimation in frequency (inverse sign);
// lane t ends with output bitrev6(t).  Stage of span H: the lower lane of a pair keeps
// v + p, the upper one (p - v) w with w = e^{+2 pi i (t mod H) / 2H}; written without
// branches (the lower lane multiplies by table entry 0, which is 1).
template <int H, typename C> __device__ __forceinline__ C dif_stage(C v, int t, const C* tw) {
  using R = decltype(C::x);
  const C p = lane_xor<H>(v, t);
  const bool up = (t & H) != 0;
  const R sg = up ? (R)-1 : (R)1;
  const C d = mk<C>(fma(sg, v.x, p.x), fma(sg, v.y, p.y));
  if constexpr (H > 1) {
    const int e = up ? (t % H) * (32 / H) : 0;
    return cm(d, tw[TW64 + e]);
  } else {
    return d;
  }
}
template <typename C> __device__ __forceinline__ C dif64_inv(C v, int t, const C* tw) {
  v = dif_stage<32>(v, t, tw);
  v = dif_stage<16>(v, t, tw);
  v = dif_stage<8>(v, t, tw);
  v = dif_stage<4>(v, t, tw);
  v = dif_stage<2>(v, t, tw);
  return dif_stage<1>(v, t, tw);
}
// its transpose with the forward sign: input in bit-reversed lane order, output natural.
// Lower lane: v + conj(w) p; upper lane: p - conj(w) v.
template <int H, typename C> __device__ __forceinline__ C dit_stage(C v, int t, const C* tw) {
  using R = decltype(C::x);
  const C p = lane_xor<H>(v, t);
  const bool up = (t & H) != 0;
  const C a = up ? p : v;
  C b = up ? v : p;
  if constexpr (H > 1) b = cmc(b, tw[TW64 + (t % H) * (32 / H)]);
  const R sg = up ? (R)-1 : (R)1;
  return mk<C>(fma(sg, b.x, a.x), fma(sg, b.y, a.y));
}
template <typename C> __device__ __forceinline__ C dit64_fwd(C v, int t, const C* tw) {
  v = dit_stage<1>(v, t, tw);
  v = dit_stage<2>(v, t, tw);
  v = dit_stage<4>(v, t, tw);
  v = dit_stage<8>(v, t, tw);
  v = dit_stage<16>(v, t, tw);
  return dit_stage<32>(v, t, tw);
}

// inverse 576-point transform of one wave: v[r] = X[t + 64 r] in; v[s] = x at the lane
// layout described above out.  buf: this wave's LDS line.
template <typename C> __device__ __forceinline__ void inv576(C (&v)[9], int t, C* buf, const C* tw) {
  dft9_t<true>(v);                                   // Y[n1] in v[perm9(n1)]
#pragma unroll
  for (int n1 = 1; n1 < 9; ++n1) v[perm9(n1)] = cm(v[perm9(n1)], tw[TW576 + n1 * 64 + t]);
  const int t1 = t & 7, g = t >> 3;
  mem_fence_compiler();
#pragma unroll
  for (int n1 = 0; n1 < 8; ++n1) buf[n1 * 64 + t] = v[perm9(n1)];
  mem_fence_compiler();
  C b[8];
#pragma unroll
  for (int t2 = 0; t2 < 8; ++t2) b[t2] = buf[g * 64 + t1 + 8 * t2];
  mem_fence_compiler();
  dft8_t<true>(b);                                   // V[m2] in b[brev3(m2)]
#pragma unroll
  for (int m2 = 1; m2 < 8; ++m2) b[brev3(m2)] = cm(b[brev3(m2)], tw[TW64 + m2 * t1]);
#pragma unroll
  for (int m2 = 0; m2 < 8; ++m2) buf[g * 72 + m2 * 9 + t1] = b[brev3(m2)];
  mem_fence_compiler();
  C c8[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) c8[q] = buf[g * 72 + t1 * 9 + q];
  mem_fence_compiler();
  v[0] = dif64_inv(v[perm9(8)], t, tw);              // slot 8 (v[8] = Y[8])
  dft8_t<true>(c8);
  // physical x[n1 + 9 (m2 + 8 m1)] in slot m1 (+1: slot 0 holds the slot-8 value)
  const C s8 = v[0];
#pragma unroll
  for (int m1 = 0; m1 < 8; ++m1) v[m1] = c8[brev3(m1)];
  v[8] = s8;
}

// forward 576-point transform (unscaled): v in the inverse's output layout; out
// v[r] = X[t + 64 r]
template <typename C> __device__ __forceinline__ void fwd576(C (&v)[9], int t, C* buf, const C* tw) {
  const int t1 = t & 7, g = t >> 3;
  v[8] = dit64_fwd(v[8], t, tw);                     // Z[8] of lane t
  C c8[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) c8[q] = v[q];
  dft8_t<false>(c8);                                 // V'[t1'] in c8[brev3(t1')]
#pragma unroll
  for (int q = 1; q < 8; ++q) c8[brev3(q)] = cmc(c8[brev3(q)], tw[TW64 + t1 * q]);
  mem_fence_compiler();
#pragma unroll
  for (int q = 0; q < 8; ++q) buf[g * 72 + t1 * 9 + q] = c8[brev3(q)];
  mem_fence_compiler();
  C b[8];
#pragma unroll
  for (int m2 = 0; m2 < 8; ++m2) b[m2] = buf[g * 72 + m2 * 9 + t1];
  mem_fence_compiler();
  dft8_t<false>(b);                                  // Z[t1 + 8 t2] in b[brev3(t2)]
#pragma unroll
  for (int t2 = 0; t2 < 8; ++t2) buf[g * 64 + t1 + 8 * t2] = b[brev3(t2)];
  mem_fence_compiler();
  const C z8 = v[8];
  C y[9];
#pragma unroll
  for (int n1 = 0; n1 < 8; ++n1) y[n1] = buf[n1 * 64 + t];
  mem_fence_compiler();
  y[8] = z8;
#pragma unroll
  for (int n1 = 1; n1 < 9; ++n1) y[n1] = cmc(y[n1], tw[TW576 + n1 * 64 + t]);
  dft9_t<false>(y);                                  // X[t + 64 r] in y[perm9(r)]
#pragma unroll
  for (int r = 0; r < 9; ++r) v[r] = y[perm9(r)];
}

// packed rows (z0, z0+1) of one Z slot -> the natural-order line of lane t:
// line[n] = A[n] + i B[n] for n < MX, A[576-n] conj... (rows' Hermitian halves), 0 between
// Every lane issues every load (no load is left to one lane), so that the line waits for
// memory once: lane 0's mode 0 drops the imaginary parts, and mode 192 (lane 0 of slot 3;
// the other lanes of that slot are zero padding), by factors 0 and 1.
template <typename C> __device__ __forceinline__ void load_line(C (&v)[9], const C* __restrict__ A, int t) {
  using R = decltype(C::x);
  const C* B = A + MX;
  const R lane0 = t == 0 ? (R)1 : (R)0, other = (R)1 - lane0;
  C a[3], b[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    a[r] = A[t + 64 * r];
    b[r] = B[t + 64 * r];
  }
  const C a3 = A[192], b3 = B[192];
  v[0] = mk<C>(fma(-other, b[0].y, a[0].x), fma(other, a[0].y, b[0].x));
#pragma unroll
  for (int r = 1; r < 3; ++r) v[r] = mk<C>(a[r].x - b[r].y, a[r].y + b[r].x);
  v[3] = mk<C>(lane0 * (a3.x - b3.y), lane0 * (a3.y + b3.x));
  v[4] = mk<C>(0, 0);
  v[5] = mk<C>(0, 0);
#pragma unroll
  for (int r = 6; r < 9; ++r) {
    const int m = N - (t + 64 * r);           // 1..192
    const C a = A[m], b = B[m];
    v[r] = mk<C>(a.x + b.y, -a.y + b.x);
  }
}

// forward spectrum P of a packed product line -> its row of the product buffer: P[k] at k
// (k < MX) and P[576 - k] at MX + k (k = 1..MX-1)
template <typename C> __device__ __forceinline__ void store_line(const C (&v)[9], C* __restrict__ dst, int t) {
#pragma unroll
  for (int r = 0; r < 3; ++r) dst[t + 64 * r] = v[r];
  if (t == 0) dst[192] = v[3];
#pragma unroll
  for (int r = 6; r < 9; ++r) dst[MX + N - (t + 64 * r)] = v[r];
}

}  // namespace xw

struct XW {
  const void *zu, *zw, *zv;      // Z slots [slot][576][193] (cx<TR>)
  void* zp;                      // [6][level][288][2 * 193] packed row pairs (cx<TR>)
  const void* tw;                // xw::TWN twiddles (cx<TR>)
  const double* dthdyu;
  double* scal;
  int want_cfl;
  int lc;                        // levels per wave
};

// LDS of one block of WPB waves: the twiddle table, then per wave a transpose line and a
// park line (one field of 9 values per lane, lane-major), then each lane's running CFL
// maximum (a double per lane, kept out of the registers)
template <int WPB, typename TR> constexpr size_t xw_lds_bytes() {
  return sizeof(cx<TR>) * ((size_t)xw::TWN + (size_t)WPB * (xw::LINE + 9 * 64)) + sizeof(double) * WPB * 64;
}
// waves per SIMD: FP64 two (the registers' limit); FP32 follows the block size, with one
// block of up to 16 waves per CU
template <typename TR, int WPB> constexpr int xw_wpe() { return sizeof(TR) == 8 ? 2 : (WPB + 3) / 4; }

// One wave per (row pair, chunk of lc levels); WPB waves per block.  Products of level l
// (u centre j = l+1): uu, uw, ww, vv at the centre, uv, vw at the face above, from U, W at
// centres j, j+1 and V at faces j-1, j.  U(j), W(j) and V(j-1) are carried from one level to
// the next, so a level costs three inverse and six forward transforms.  At most three fields
// stay in registers: the one the next products do not need waits in the wave's park line
// (V(j-1) between levels, W(j) while the centre products are formed) and is read back slot
// by slot as an operand.
// CFL: whether the launch forms the CFL maximum (a template flag, so that the slot loops
// carry no branch on it)
template <int WPB, bool CFL, typename TR>
__global__ void __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(xw_wpe<TR, WPB>())))
    k_xw576(XW a, Dims dm) {
  using C = cx<TR>;
  const TR half = (TR)0.5;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  C* tw = lds;                                   // twiddle table, shared by the block
  for (int e = threadIdx.x; e < xw::TWN; e += blockDim.x) tw[e] = static_cast<const C*>(a.tw)[e];
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = threadIdx.x & 63;
  C* buf = lds + xw::TWN + wv * (xw::LINE + 9 * 64);
  C* pk = buf + xw::LINE + t;                    // park line: slot s at pk[64 s]
  const int nlev = dm.nyu - 2;
  const int npair = xw::N / 2;
  const int item = blockIdx.x * WPB + wv;
  const int pair = item % npair, chunk = item / npair;
  const int l0 = chunk * a.lc;
  const int l1 = min(l0 + a.lc, nlev);
  if (l0 >= l1) return;
  const size_t slot = (size_t)xw::N * xw::MX;
  const size_t row = (size_t)pair * 2 * xw::MX;
  const C* zu = static_cast<const C*>(a.zu);
  const C* zw = static_cast<const C*>(a.zw);
  const C* zv = static_cast<const C*>(a.zv);
  C* zp = static_cast<C*>(a.zp);
  double* const cm = reinterpret_cast<double*>(lds + xw::TWN + WPB * (xw::LINE + 9 * 64)) + wv * 64 + t;
  if constexpr (CFL) *cm = 0.0;
  C cu[9], cw[9], p[9];
  auto park = [&](const C (&x)[9]) {
    xw::mem_fence_compiler();
#pragma unroll
    for (int s2 = 0; s2 < 9; ++s2) pk[64 * s2] = x[s2];
    xw::mem_fence_compiler();
  };
  auto out = [&](int q, int l, C (&x)[9]) {
    xw::fwd576(x, t, buf, tw);
    xw::store_line(x, zp + ((size_t)q * nlev + l) * slot + row, t);
  };
  xw::load_line(cu, zu + (size_t)(l0 + 1) * slot + row, t);
  xw::inv576(cu, t, buf, tw);
  xw::load_line(cw, zw + (size_t)(l0 + 1) * slot + row, t);
  xw::inv576(cw, t, buf, tw);
  xw::load_line(p, zv + (size_t)l0 * slot + row, t);
  xw::inv576(p, t, buf, tw);
  park(p);                                        // V(j-1)
  for (int l = l0; l < l1; ++l) {
    // V(j), the face above: vv and the CFL maximum at the centre, from V(j-1) parked
    C vj[9];
    xw::load_line(vj, zv + (size_t)(l + 1) * slot + row, t);
    xw::inv576(vj, t, buf, tw);
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const C vb = pk[64 * s];
      p[s] = mk<C>(half * (vb.x + vj[s].x), half * (vb.y + vj[s].y));   // v at the centre
    }
    xw::mem_fence_compiler();
    park(cw);                                     // W(j) waits while V(j) stays resident
    if constexpr (CFL) {
      // |u|/dx + |v|/dy + |w|/dz at the centre, W(j) read back from the park line (its
      // registers are free from here to the end of the level)
      const double dyi = a.dthdyu[l + 1];
      double cmax = *cm;
#pragma unroll
      for (int s = 0; s < 9; ++s) {
        const C w = pk[64 * s];
        const double s0 = fabs((double)cu[s].x) * dm.dxi + fabs((double)p[s].x) * dyi + fabs((double)w.x) * dm.dzi;
        const double s1 = fabs((double)cu[s].y) * dm.dxi + fabs((double)p[s].y) * dyi + fabs((double)w.y) * dm.dzi;
        cmax = nanmax(cmax, nanmax(s0, s1));
      }
      *cm = cmax;
    }
#pragma unroll
    for (int s = 0; s < 9; ++s) p[s] = mk<C>(p[s].x * p[s].x, p[s].y * p[s].y);
    out(3, l, p);
#pragma unroll
    for (int s = 0; s < 9; ++s) p[s] = mk<C>(cu[s].x * cu[s].x, cu[s].y * cu[s].y);
    out(0, l, p);
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const C w = pk[64 * s];
      p[s] = mk<C>(cu[s].x * w.x, cu[s].y * w.y);
    }
    out(1, l, p);
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const C w = pk[64 * s];
      p[s] = mk<C>(w.x * w.x, w.y * w.y);
    }
    out(2, l, p);
    // U(j+1): uv at the face, then carried
    xw::load_line(p, zu + (size_t)(l + 2) * slot + row, t);
    xw::inv576(p, t, buf, tw);
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const C uf = mk<C>(half * (cu[s].x + p[s].x), half * (cu[s].y + p[s].y));
      cu[s] = mk<C>(uf.x * vj[s].x, uf.y * vj[s].y);
    }
    out(4, l, cu);
#pragma unroll
    for (int s = 0; s < 9; ++s) cu[s] = p[s];
    // W(j+1): vw at the face, then carried; W(j) comes back from the park line, which then
    // takes V(j) for the next level
    xw::load_line(p, zw + (size_t)(l + 2) * slot + row, t);
    xw::inv576(p, t, buf, tw);
    xw::mem_fence_compiler();
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const C w = pk[64 * s];
      pk[64 * s] = vj[s];                         // V(j) parked for the next level
      const C wf = mk<C>(half * (w.x + p[s].x), half * (w.y + p[s].y));
      vj[s] = mk<C>(vj[s].x * wf.x, vj[s].y * wf.y);
    }
    xw::mem_fence_compiler();
    out(5, l, vj);
#pragma unroll
    for (int s = 0; s < 9; ++s) cw[s] = p[s];
  }
  if constexpr (CFL) {
    double cmax = *cm;
    for (int off = 32; off > 0; off >>= 1) cmax = nanmax(cmax, __shfl_down(cmax, off, 64));
    if (t == 0) atomic_max_pos(a.scal + S_CFLMAX, cmax);
  }
}

// Canopy x pass at 576 points, one wavefront per (entry, row pair): the rhs rows (z0, z0+1)
// of a canopy plane to physical space, rhs = 0 on the canopy points of a plane below the
// tips (XIB::paired = 1) or -beta Lu_y there on a tip plane (paired = 0, the Lu_y rows of
// the entry's lslot transformed alongside), then back to Fourier and the two rows separated
// (tridLU_3D_IB.f90:224-245, 306-325; the operations of k_xpass_ib on the transforms of
// k_xw576).  After inv576 lane t holds x = (t>>3) + 9 ((t&7) + 8 s) in slot s < 8 and
// x = 8 + 9 bitrev6(t) in slot 8, which is where its mask bytes are read.
template <int WPB>
__global__ void __launch_bounds__(64 * WPB) k_xpass_ib576(XIB a, Dims dm, int nent,
                                                          const double2* __restrict__ twg) {
  using C = double2;
  extern __shared__ char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  C* tw = lds;                                   // twiddle table, shared by the block
  for (int e = threadIdx.x; e < xw::TWN; e += blockDim.x) tw[e] = twg[e];
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = threadIdx.x & 63;
  C* buf = lds + xw::TWN + wv * xw::LINE;
  const int npair = dm.pz / 2;
  const int item = blockIdx.x * WPB + wv;
  if (item >= nent * npair) return;              // whole waves (no barrier follows)
  const int e = item / npair, z0 = 2 * (item % npair);
  const IBEntry en = a.ent[e];
  const size_t slot = (size_t)dm.pz * xw::MX;
  C* z = reinterpret_cast<C*>(a.z);
  C* const row = z + en.slot * slot + (size_t)z0 * xw::MX;
  const bool tip = a.paired == 0;
  C v[9], w[9];
  xw::load_line(v, row, t);
  xw::inv576(v, t, buf, tw);
  if (tip) {
    xw::load_line(w, z + en.lslot * slot + (size_t)z0 * xw::MX, t);
    xw::inv576(w, t, buf, tw);
  }
  // canopy points: mask bytes of rows z0 (real parts) and z0 + 1 (imaginary parts)
  const unsigned char* m0 = a.masks + ((size_t)en.mask * dm.pz + z0) * xw::N;
  const double beta = a.bre * a.scal[S_DT];
  const int n1 = t >> 3, m2 = t & 7;
  const int x8 = 8 + 9 * (int)(__brev((unsigned)t) >> 26);
#pragma unroll
  for (int s = 0; s < 9; ++s) {
    const int x = s < 8 ? n1 + 9 * (m2 + 8 * s) : x8;
    const bool c0 = m0[x] != 0, c1 = m0[xw::N + x] != 0;
    if (c0) v[s].x = tip ? -beta * w[s].x : 0.0;
    if (c1) v[s].y = tip ? -beta * w[s].y : 0.0;
  }
  xw::fwd576(v, t, buf, tw);                     // v[r] = P[t + 64 r]
  // rows z0, z0+1 from P[k] and P[576 - k], x 1/576, k = t + 64 r < 193: P[576 - k] is
  // slot 8 - r of lane 64 - t, or, in lane 0, slot 9 - r (P[0] itself for k = 0)
  const double sx = 1.0 / (double)xw::N;
  auto put = [&](int k, C zk, C zn) {
    C ra = mk<C>(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
    C rb = mk<C>(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x));
    ra = mk<C>(sx * ra.x, sx * ra.y);
    rb = mk<C>(sx * rb.x, sx * rb.y);
    if (k == 0) { ra.y = 0; rb.y = 0; }
    row[k] = ra;
    row[xw::MX + k] = rb;
  };
  const int src = (64 - t) & 63;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    C zn = mk<C>(__shfl(v[8 - r].x, src, 64), __shfl(v[8 - r].y, src, 64));
    if (t == 0) zn = r == 0 ? v[0] : v[9 - r];
    put(t + 64 * r, v[r], zn);
  }
  if (t == 0) put(192, v[3], v[6]);
}


}  // namespace lh
