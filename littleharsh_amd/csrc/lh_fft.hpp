// lh_fft.hpp -- batched LDS-staged Stockham FFT for gfx950, FP64 or FP32.
//
// A "line" of N complex values lives in LDS; T threads of the workgroup work on one line.
// Each stage of radix R reads the line with stride N/R (conflict-free, consecutive
// butterflies on consecutive threads), twiddles, does an in-register DFT_R and writes the
// autosorted (Stockham) positions back to the same LDS line.  Every stage brackets its
// read and write phases with __syncthreads(), so a workgroup of L lines x T threads
// transforms L lines at once.  Radices are composed at compile time from 2, 3 and 4.
//
// Conventions: forward = sum x_n e^{-2 pi i nk/N} (no scaling), inverse = e^{+...}; the
// 1/N of the reference's forward rft/cft (rft_buff.f:63-73, cft_buff.f:44-45) is applied
// by the callers.
#pragma once
#include <hip/hip_runtime.h>

namespace lh {

template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename T> using cx = typename cvec<T>::type;
template <typename C> struct cscalar;
template <> struct cscalar<double2> { using type = double; };
template <> struct cscalar<float2> { using type = float; };

template <typename C> __device__ __forceinline__ C mk(decltype(C::x) a, decltype(C::x) b) {
  C r; r.x = a; r.y = b; return r;
}

// A device buffer held as a global-address-space pointer.  A pointer that a kernel loads
// from memory (a job table) is generic to the compiler, which then uses flat loads and
// stores; through a gptr every access is a global one.
template <typename T> using gptr = __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> gp(T* p) { return (gptr<T>)p; }
// element i of a complex buffer in global memory, read or written as a two-element vector
// (double2 and float2 themselves cannot be copied through a gptr)
template <typename C> __device__ __forceinline__ C gld(const C* p, size_t i) {
  typedef typename cscalar<C>::type S __attribute__((ext_vector_type(2)));
  const S v = ((const __attribute__((address_space(1))) S*)p)[i];
  return mk<C>(v.x, v.y);
}
template <typename C> __device__ __forceinline__ void gst(C* p, size_t i, C a) {
  typedef typename cscalar<C>::type S __attribute__((ext_vector_type(2)));
  S v;
  v.x = a.x;
  v.y = a.y;
  ((__attribute__((address_space(1))) S*)p)[i] = v;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  return mk<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C rot(C a) {
  return INV ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x);
}

// Workgroup barrier for LDS hand-offs only: waits for this wave's LDS operations, not for
// its outstanding global loads, so prefetched loads stay in flight across FFT stages.
__device__ __forceinline__ void lds_bar() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// W_R^m = exp(-2 pi i m / R) for the composite radices, cos/sin(2 pi m/R) to double precision
__device__ constexpr double kW6[6][2] = {{1.0, -0.0}, {5.00000000000000111022e-01, -8.66025403784438596588e-01}, {-4.99999999999999777955e-01, -8.66025403784438707611e-01}, {-1.0, -1.22464679914735320717e-16}, {-5.00000000000000444089e-01, 8.66025403784438374544e-01}, {5.00000000000000111022e-01, 8.66025403784438596588e-01}};
__device__ constexpr double kW8[8][2] = {{1.0, -0.0}, {7.07106781186547572737e-01, -7.07106781186547461715e-01}, {6.12323399573676603587e-17, -1.0}, {-7.07106781186547461715e-01, -7.07106781186547572737e-01}, {-1.0, -1.22464679914735320717e-16}, {-7.07106781186547683760e-01, 7.07106781186547461715e-01}, {-1.83697019872102968750e-16, 1.0}, {7.07106781186547350693e-01, 7.07106781186547683760e-01}};
__device__ constexpr double kW9[9][2] = {{1.0, -0.0}, {7.66044443118978013452e-01, -6.42787609686539251896e-01}, {1.73648177666930414453e-01, -9.84807753012208020316e-01}, {-4.99999999999999777955e-01, -8.66025403784438707611e-01}, {-9.39692620785908316883e-01, -3.42020143325668879442e-01}, {-9.39692620785908427905e-01, 3.42020143325668657397e-01}, {-5.00000000000000444089e-01, 8.66025403784438374544e-01}, {1.73648177666929970364e-01, 9.84807753012208131338e-01}, {7.66044443118977791407e-01, 6.42787609686539584963e-01}};
__device__ constexpr double kW12[12][2] = {{1.0, -0.0}, {8.66025403784438707611e-01, -4.99999999999999944489e-01}, {5.00000000000000111022e-01, -8.66025403784438596588e-01}, {6.12323399573676603587e-17, -1.0}, {-4.99999999999999777955e-01, -8.66025403784438707611e-01}, {-8.66025403784438707611e-01, -4.99999999999999944489e-01}, {-1.0, -1.22464679914735320717e-16}, {-8.66025403784438818633e-01, 4.99999999999999722444e-01}, {-5.00000000000000444089e-01, 8.66025403784438374544e-01}, {-1.83697019872102968750e-16, 1.0}, {5.00000000000000111022e-01, 8.66025403784438596588e-01}, {8.66025403784438374544e-01, 5.00000000000000444089e-01}};
__device__ constexpr double kW16[16][2] = {{1.0, -0.0}, {9.23879532511286738483e-01, -3.82683432365089781779e-01}, {7.07106781186547572737e-01, -7.07106781186547461715e-01}, {3.82683432365089837290e-01, -9.23879532511286738483e-01}, {6.12323399573676603587e-17, -1.0}, {-3.82683432365089726268e-01, -9.23879532511286738483e-01}, {-7.07106781186547461715e-01, -7.07106781186547572737e-01}, {-9.23879532511286738483e-01, -3.82683432365089892802e-01}, {-1.0, -1.22464679914735320717e-16}, {-9.23879532511286849505e-01, 3.82683432365089670757e-01}, {-7.07106781186547683760e-01, 7.07106781186547461715e-01}, {-3.82683432365090336891e-01, 9.23879532511286516439e-01}, {-1.83697019872102968750e-16, 1.0}, {3.82683432365090003824e-01, 9.23879532511286627461e-01}, {7.07106781186547350693e-01, 7.07106781186547683760e-01}, {9.23879532511286516439e-01, 3.82683432365090392402e-01}};

template <int R> __device__ __forceinline__ double wre(int m);
template <int R> __device__ __forceinline__ double wim(int m);
#define LH_WTAB(R)                                                                   \
  template <> __device__ __forceinline__ double wre<R>(int m) { return kW##R[m][0]; } \
  template <> __device__ __forceinline__ double wim<R>(int m) { return kW##R[m][1]; }
LH_WTAB(6)
LH_WTAB(8)
LH_WTAB(9)
LH_WTAB(12)
LH_WTAB(16)
#undef LH_WTAB

template <int R> struct split;
template <> struct split<6> { static constexpr int A = 2, B = 3; };
template <> struct split<8> { static constexpr int A = 4, B = 2; };
template <> struct split<9> { static constexpr int A = 3, B = 3; };
template <> struct split<12> { static constexpr int A = 4, B = 3; };
template <> struct split<16> { static constexpr int A = 4, B = 4; };

// In-register DFT of size R
template <int R, bool INV, typename C> struct DFT {
  // composite: x[n2 + B*n1] -> X[k1 + A*k2]
  static __device__ __forceinline__ void run(C* v) {
    constexpr int A = split<R>::A, B = split<R>::B;
    using T = decltype(C::x);
    C y[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      C t[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) t[n1] = v[n2 + B * n1];
      DFT<A, INV, C>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) {
        const int m = (n2 * k1) % R;
        if (m == 0) {
          y[n2 * A + k1] = t[k1];
        } else if (4 * m == R) {            // W = -i (forward) / +i (inverse)
          y[n2 * A + k1] = rot<INV>(t[k1]);
        } else if (2 * m == R) {            // W = -1
          y[n2 * A + k1] = mk<C>(-t[k1].x, -t[k1].y);
        } else if (4 * m == 3 * R) {        // W = +i (forward) / -i (inverse)
          y[n2 * A + k1] = rot<!INV>(t[k1]);
        } else {
          const C w = mk<C>((T)wre<R>(m), INV ? -(T)wim<R>(m) : (T)wim<R>(m));
          y[n2 * A + k1] = cmul(t[k1], w);
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      C t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = y[n2 * A + k1];
      DFT<B, INV, C>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
    }
  }
};
template <bool INV, typename C> struct DFT<1, INV, C> {
  static __device__ __forceinline__ void run(C*) {}
};
template <bool INV, typename C> struct DFT<2, INV, C> {
  static __device__ __forceinline__ void run(C* v) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <bool INV, typename C> struct DFT<3, INV, C> {
  static __device__ __forceinline__ void run(C* v) {
    using T = decltype(C::x);
    const T s = (T)0.86602540378443864676;
    const C t1 = cadd(v[1], v[2]);
    const C t2 = mk<C>(v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y);
    const C dd = csub(v[1], v[2]);
    const C r = rot<INV>(mk<C>(s * dd.x, s * dd.y));  // -i s (v1-v2) forward
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, r);
    v[2] = csub(t2, r);
  }
};
template <bool INV, typename C> struct DFT<4, INV, C> {
  static __device__ __forceinline__ void run(C* v) {
    const C s0 = cadd(v[0], v[2]), s1 = csub(v[0], v[2]);
    const C s2 = cadd(v[1], v[3]), s3 = rot<INV>(csub(v[1], v[3]));
    v[0] = cadd(s0, s2);
    v[2] = csub(s0, s2);
    v[1] = cadd(s1, s3);
    v[3] = csub(s1, s3);
  }
};

// Per-thread twiddles of one stage: the thread's butterflies b = t + q*T are fixed for a
// given line position t, so their twiddles W_N^{(b%NS)*r*N/(NS*R)} are loaded once into
// registers and reused for every line the thread transforms.
template <int N, int T, int R, int NS, typename C>
struct StageTw {
  static constexpr int NB = N / R;
  static constexpr int PER = (NB + T - 1) / T;
  C w[PER][R > 1 ? R - 1 : 1];
  __device__ __forceinline__ void load(int t, const C* __restrict__ tw) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int b = t + q * T;
      const int j = (b < NB ? b : 0) % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) w[q][r - 1] = gld(tw, j * r * (N / (NS * R)));
    }
  }
  __device__ __forceinline__ C get(int q, int r) const { return w[q][r - 1]; }
  template <bool INV, int RR> __device__ __forceinline__ void apply(int q, C (&v)[RR]) const {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = INV ? cmulc(v[r], w[q][r - 1]) : cmul(v[r], w[q][r - 1]);
  }
};

// The same twiddles kept in an LDS table (one row of PER*(R-1) values per line position t)
// when registers are the occupancy limit.
template <int N, int T, int R, int NS, typename C>
struct StageTwL {
  static constexpr int NB = N / R;
  static constexpr int PER = (NB + T - 1) / T;
  static constexpr int CNT = PER * (R - 1);
  static constexpr int SIZE = T * CNT;
  const C* p;
  __device__ __forceinline__ void init(C* tab, int t, const C* __restrict__ tw, bool fill) {
    p = tab + t * CNT;
    if (fill) {
      for (int q = 0; q < PER; ++q) {
        const int b = t + q * T;
        const int j = (b < NB ? b : 0) % NS;
        for (int r = 1; r < R; ++r) tab[t * CNT + q * (R - 1) + r - 1] = tw[j * r * (N / (NS * R))];
      }
    }
  }
  __device__ __forceinline__ C get(int q, int r) const { return p[q * (R - 1) + r - 1]; }
  template <bool INV, int RR> __device__ __forceinline__ void apply(int q, C (&v)[RR]) const {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = INV ? cmulc(v[r], get(q, r)) : cmul(v[r], get(q, r));
  }
};

// Stage twiddles from an LDS table with one row per distinct j = b mod NS, shared by the
// block: row j holds w^(j r N/(NS R)) for r = 1..M; a higher power r > M (at most 2M) is the
// product of the row's entries r - M and M (one rounding more).  Rows of butterflies b < NS
// need no modulo, so their reads are a per-thread base plus compile-time offsets.
template <int N, int T, int R, int NS, int M, typename C>
struct StageTwD {
  static_assert(M >= 1 && 2 * M >= R - 1, "StageTwD: every power r < R must be r <= M or a product of two stored ones");
  static constexpr int NB = N / R;
  static constexpr int PER = (NB + T - 1) / T;
  static constexpr int SIZE = NS * M;
  const C* tab;
  int t;
  __device__ __forceinline__ static void fill(C* dst, const C* __restrict__ tw, int tid, int nthreads) {
    for (int e = tid; e < SIZE; e += nthreads) {
      const int j = e / M, r = e % M + 1;
      dst[e] = tw[j * r * (N / (NS * R))];
    }
  }
  __device__ __forceinline__ void init(const C* tb, int tt) {
    tab = tb;
    t = tt;
  }
  __device__ __forceinline__ C get(int q, int r) const {
    const int b = t + q * T;
    const C* row = tab + ((b < NB ? b : 0) % NS) * M;
    if (r <= M) return row[r - 1];
    return cmul(row[r - M - 1], row[M - 1]);
  }
  template <bool INV, int RR> __device__ __forceinline__ void apply(int q, C (&v)[RR]) const {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = INV ? cmulc(v[r], get(q, r)) : cmul(v[r], get(q, r));
  }
};
// StageTwD for both later stages of Plan<N>: the middle stage with every power in its rows,
// the last stage of a three-stage plan with the first R/2 of them
template <int N, int T, int RA, int RB, int RC, typename C>
struct TwRowsPlan {
  using S2 = StageTwD<N, T, RB, RA, RB - 1, C>;
  using S3 = StageTwD<N, T, (RC > 1 ? RC : 2), RA * RB, (RC > 2 ? RC / 2 : 1), C>;
  static constexpr int SIZE = S2::SIZE + (RC > 1 ? S3::SIZE : 0);
  S2 s2;
  S3 s3;
  __device__ __forceinline__ static void fill(C* dst, const C* __restrict__ tw, int tid, int nthreads) {
    S2::fill(dst, tw, tid, nthreads);
    if constexpr (RC > 1) S3::fill(dst + S2::SIZE, tw, tid, nthreads);
  }
  __device__ __forceinline__ void init(const C* tb, int t) {
    s2.init(tb, t);
    if constexpr (RC > 1) s3.init(tb + S2::SIZE, t);
  }
};

// Position of element e of a line stored with one pad slot after every 2^S elements
// (S = 0: no padding); spreads the strided stage writes over the LDS banks.
template <int S> __device__ __forceinline__ constexpr int lpad(int e) { return S ? e + (e >> S) : e; }
template <int N, int S> struct PadLen { static constexpr int value = S ? N + (N >> S) : N; };

// One Stockham stage of radix R on a line of N values; NS = product of earlier radices.
// w: the thread's twiddles for this stage (unused when NS == 1).
template <int N, int T, int R, int NS, bool INV, int S, typename C, typename W>
__device__ __forceinline__ void stage(C* line, int t, const W& w) {
  constexpr int NB = N / R;
  constexpr int PER = (NB + T - 1) / T;
  C v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = t + q * T;
    if (PER * T == NB || b < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = line[lpad<S>(b + r * NB)];
    }
  }
  lds_bar();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = t + q * T;
    if (PER * T == NB || b < NB) {
      const int j = b % NS;
      if constexpr (NS > 1) w.template apply<INV>(q, v[q]);
      DFT<R, INV, C>::run(v[q]);
      const int idx = (b / NS) * NS * R + j;
#pragma unroll
      for (int r = 0; r < R; ++r) line[lpad<S>(idx + r * NS)] = v[q][r];
    }
  }
  lds_bar();
}

// The same stage out of place: read the line from src, write the autosorted result to dst.
// Only the barrier after the writes is needed (the next stage reads dst and overwrites src,
// whose reads all precede that barrier).
template <int N, int T, int R, int NS, bool INV, int S, typename C, typename W>
__device__ __forceinline__ void stage_pp(const C* src, C* dst, int t, const W& w) {
  constexpr int NB = N / R;
  constexpr int PER = (NB + T - 1) / T;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = t + q * T;
    if (PER * T == NB || b < NB) {
      C v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = src[lpad<S>(b + r * NB)];
      const int j = b % NS;
      if constexpr (NS > 1) w.template apply<INV>(q, v);
      DFT<R, INV, C>::run(v);
      const int idx = (b / NS) * NS * R + j;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[lpad<S>(idx + r * NS)] = v[r];
    }
  }
  lds_bar();
}

struct NoTw {};

// Stage pieces for transforms whose first stage takes its inputs from registers (loaded
// from global memory by the caller) or whose last stage leaves its outputs in registers
// (stored by the caller).  Butterfly b of a stage of radix R reads positions b + r*N/R and
// writes the autosorted positions stage_pos(b, r).
template <int N, int R, int NS> __device__ __forceinline__ int stage_pos(int b, int r) {
  return (b / NS) * NS * R + b % NS + r * NS;
}
// twiddle (butterfly q of this thread), DFT and store to dst
template <int N, int R, int NS, bool INV, int S, typename C, typename W>
__device__ __forceinline__ void stage_regs_store(C (&v)[R], int b, int q, const W& w, C* dst) {
  if constexpr (NS > 1) w.template apply<INV>(q, v);
  DFT<R, INV, C>::run(v);
#pragma unroll
  for (int r = 0; r < R; ++r) dst[lpad<S>(stage_pos<N, R, NS>(b, r))] = v[r];
}
// load from src, twiddle and DFT; the outputs stay in v (positions stage_pos(b, r))
template <int N, int R, int NS, bool INV, int S, typename C, typename W>
__device__ __forceinline__ void stage_load_regs(const C* src, int b, int q, const W& w, C (&v)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src[lpad<S>(b + r * (N / R))];
  if constexpr (NS > 1) w.template apply<INV>(q, v);
  DFT<R, INV, C>::run(v);
}

// Plans: threads per line T, the radix sequence, and the per-thread twiddle set Tw<C>.
// run() works in place; run_pp() ping-pongs between two lines and leaves the result in
// the second one when NSTAGE is odd, in the first one when it is even.  Plan<N, 0> is the
// default factorisation of length N; Plan<N, 1>, where a length has one, is an
// alternative (other radices and threads per line) that the autotuner may choose.
#define LH_PLAN2(NN, VV, TT, R1, R2)                                                    \
  template <> struct Plan<NN, VV> {                                                     \
    static constexpr int T = TT;                                                        \
    static constexpr int RA = R1, RB = R2, RC = 1;                                      \
    template <typename C> struct Tw {                                                   \
      StageTw<NN, TT, R2, R1, C> s2;                                                    \
      __device__ __forceinline__ void load(int t, const C* tw) { s2.load(t, tw); }      \
    };                                                                                  \
    template <typename C> struct TwL {                                                  \
      StageTwL<NN, TT, R2, R1, C> s2;                                                   \
      static constexpr int SIZE = StageTwL<NN, TT, R2, R1, C>::SIZE;                    \
      __device__ __forceinline__ void init(C* tab, int t, const C* tw, bool fill) {     \
        s2.init(tab, t, tw, fill);                                                      \
      }                                                                                 \
    };                                                                                  \
    template <bool INV, int S = 0, typename C, typename W>                              \
    static __device__ void run(C* l, int t, const W& w) {                               \
      stage<NN, TT, R1, 1, INV, S>(l, t, NoTw{});                                       \
      stage<NN, TT, R2, R1, INV, S>(l, t, w.s2);                                        \
    }                                                                                   \
    static constexpr int NSTAGE = 2;                                                    \
    template <bool INV, int S = 0, typename C, typename W>                              \
    static __device__ void run_pp(C* a, C* b, int t, const W& w) {                      \
      stage_pp<NN, TT, R1, 1, INV, S>(a, b, t, NoTw{});                                 \
      stage_pp<NN, TT, R2, R1, INV, S>(b, a, t, w.s2);                                  \
    }                                                                                   \
  };
#define LH_PLAN3(NN, VV, TT, R1, R2, R3)                                                \
  template <> struct Plan<NN, VV> {                                                     \
    static constexpr int T = TT;                                                        \
    static constexpr int RA = R1, RB = R2, RC = R3;                                     \
    template <typename C> struct Tw {                                                   \
      StageTw<NN, TT, R2, R1, C> s2;                                                    \
      StageTw<NN, TT, R3, R1 * R2, C> s3;                                               \
      __device__ __forceinline__ void load(int t, const C* tw) { s2.load(t, tw); s3.load(t, tw); } \
    };                                                                                  \
    template <typename C> struct TwL {                                                  \
      StageTwL<NN, TT, R2, R1, C> s2;                                                   \
      StageTwL<NN, TT, R3, R1 * R2, C> s3;                                              \
      static constexpr int SIZE = StageTwL<NN, TT, R2, R1, C>::SIZE +                   \
                                  StageTwL<NN, TT, R3, R1 * R2, C>::SIZE;               \
      __device__ __forceinline__ void init(C* tab, int t, const C* tw, bool fill) {     \
        s2.init(tab, t, tw, fill);                                                      \
        s3.init(tab + StageTwL<NN, TT, R2, R1, C>::SIZE, t, tw, fill);                  \
      }                                                                                 \
    };                                                                                  \
    template <bool INV, int S = 0, typename C, typename W>                              \
    static __device__ void run(C* l, int t, const W& w) {                               \
      stage<NN, TT, R1, 1, INV, S>(l, t, NoTw{});                                       \
      stage<NN, TT, R2, R1, INV, S>(l, t, w.s2);                                        \
      stage<NN, TT, R3, R1 * R2, INV, S>(l, t, w.s3);                                   \
    }                                                                                   \
    static constexpr int NSTAGE = 3;                                                    \
    template <bool INV, int S = 0, typename C, typename W>                              \
    static __device__ void run_pp(C* a, C* b, int t, const W& w) {                      \
      stage_pp<NN, TT, R1, 1, INV, S>(a, b, t, NoTw{});                                 \
      stage_pp<NN, TT, R2, R1, INV, S>(b, a, t, w.s2);                                  \
      stage_pp<NN, TT, R3, R1 * R2, INV, S>(a, b, t, w.s3);                             \
    }                                                                                   \
  };

template <int N, int V = 0> struct Plan;

LH_PLAN3(576, 0, 48, 12, 12, 4)
LH_PLAN3(384, 0, 48, 8, 12, 4)
LH_PLAN2(144, 0, 12, 12, 12)
LH_PLAN3(96, 0, 8, 4, 4, 6)
LH_PLAN2(72, 0, 6, 12, 6)
LH_PLAN2(64, 0, 4, 16, 4)
LH_PLAN2(48, 0, 4, 12, 4)
LH_PLAN2(32, 0, 4, 8, 4)
// alternatives for the lengths of the small channels (nx, nz of 32..96 give 48..144)
LH_PLAN2(144, 1, 16, 9, 16)
LH_PLAN2(96, 1, 12, 8, 12)
LH_PLAN2(72, 1, 9, 8, 9)
LH_PLAN2(48, 1, 8, 6, 8)
#undef LH_PLAN2
#undef LH_PLAN3

// number of plans of length N (1 or 2)
template <int N> constexpr int plan_count() { return (N == 48 || N == 72 || N == 96 || N == 144) ? 2 : 1; }

}  // namespace lh
