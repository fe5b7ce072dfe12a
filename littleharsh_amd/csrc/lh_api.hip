// lh_api.hip -- C-ABI of the MI355X solver (declarations in include/littleharsh_amd.h).
//
// The handle owns one HIP stream, the device-resident state in [plane][column] layout and
// all work buffers.  lh_substep issues the kernels of one RK3 substep:
//   NL stage   : k_zinv (u,w,v) -> k_xw576 at 576 points in x, k_xraw2 at the other sizes
//                (x transforms, six products, CFL max)
//                -> k_zasm (A/B recombination, forward z transforms -> Nu_new)
//   IB stage   : k_rhs_ib (rhs and Lu_y on canopy planes) -> k_zinv -> k_xpass_ib -> k_zfwd
//   column     : k_prep (LU of the shared velocity matrices) -> k_vel_rhs (RHS0 + rhs of
//                solveU/V on the fly, Thomas u, w, v)
//                -> k_pres (constant-flow-rate correction, solveP, v_corr); these read the
//                channel's parameters from a job table, so that a batch launches them once
//                for all its members
// Statistics and spectra (lh_stats.hpp) run on request, outside the substep.
// lh_step replays a captured HIP graph of three substeps (+ meanflow_ctP).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/littleharsh_amd.h"
#include "lh_kernels.hpp"
#include "lh_nl.hpp"
#include "lh_xw.hpp"
#include "lh_col.hpp"
#include "lh_stats.hpp"
#include "lh_internal.hpp"

using namespace lh;

struct ColJob;
struct IbSets;

// Default launch parameters (per handle; lh_autotune may change them):
//   LH_XRAW_LC  levels per block of the sliding x pass (0: chosen per grid, see chunk_levels)
//   LH_ZASM_LC  levels per block of the z assembly (0: chosen per grid)
//   LH_COL_BS   threads per block of the canopy rhs kernel (one column per thread; the column
//               stage itself runs 64 columns per block, which spread the 74k columns of the
//               production grid evenly over the 256 CUs)
#ifndef LH_XRAW_LC
#define LH_XRAW_LC 0
#endif
#ifndef LH_ZASM_LC
#define LH_ZASM_LC 0
#endif

// Levels per block for a kernel whose blocks each fill a CU (LDS-bound, one resident block):
// it runs in ceil(blocks / CUs) rounds of (levels + start-up) level times; the start-up of a
// level chunk costs about `start` levels.  Returns the level count in [lo, hi] of least cost.
static int chunk_levels(int nlev, long colblocks, int ncu, int lo, int hi, double start) {
  int best = lo;
  double bc = 1e300;
  for (int lc = lo; lc <= hi; ++lc) {
    const long nblk = colblocks * ((nlev + lc - 1) / lc);
    const double rounds = (double)((nblk + ncu - 1) / ncu);
    const double cost = rounds * (lc + start);
    if (cost < bc - 1e-9) {
      bc = cost;
      best = lc;
    }
  }
  return best;
}
#ifndef LH_COL_BS
#define LH_COL_BS 64
#endif

struct lh_handle {
  std::string err;
  int dev = 0;
  hipStream_t st = nullptr;
  // side stream for the LU of the velocity matrices (k_prep), which needs only dt: it runs
  // beside the nonlinear stage, forked and joined by the two events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  Dims dm{};
  int prec = LH_PREC_FP64;
  int flag_ctpress = 1;
  double rk[15];
  double gw[2];
  int tip_u[2], tip_v[2];
  int ntilex = 1, ntilez = 1, band_v[2] = {0, 0};
  // metrics (device, slot indexed)
  double *yu = nullptr, *yv = nullptr, *dthdyu = nullptr, *dthdyv = nullptr, *dyu2i = nullptr,
         *dyv2i = nullptr;
  double *k1x = nullptr, *k2x = nullptr, *k1z = nullptr, *k2z = nullptr;
  double* scal = nullptr;
  // state
  double2 *u1 = nullptr, *u2 = nullptr, *u3 = nullptr, *p = nullptr;
  // nonlinear terms, two buffers: substeps 1 and 3 write A and read B, substep 2 the reverse
  double2 *nu1 = nullptr, *nu2 = nullptr, *nu3 = nullptr;     // A
  double2 *nb1 = nullptr, *nb2 = nullptr, *nb3 = nullptr;     // B
  int nu_last = 0;                                             // buffer written last (0: A)
  // optional forcing added to Nu_new in every substep (lh_set_forcing; allocated on first use)
  double2 *fo1 = nullptr, *fo2 = nullptr, *fo3 = nullptr;
  bool forcing = false;
  double2 *r1 = nullptr, *r2 = nullptr, *r3 = nullptr, *psi = nullptr, *tipl = nullptr;
  void *zu = nullptr, *zw = nullptr, *zv = nullptr, *zp = nullptr;
  double* dg3 = nullptr;
  double *ayu = nullptr, *ayv = nullptr, *ctwork = nullptr;
  ColJob* coljobs = nullptr;    // the column-stage jobs of the three substeps
  int* colmap = nullptr;        // block -> channel (all 0)
  IbSets* ib = nullptr;         // canopy-stage job tables (small channels with a canopy)
  void *twx = nullptr, *twz = nullptr;         // twiddles in the NL precision
  void *twx64 = nullptr, *twz64 = nullptr;     // FP64 twiddles (IB, statistics)
  // the one-wave-per-line x pass (576 collocation points in x): its twiddle table, and the
  // same rounded to FP32 in the mixed-precision mode
  double2* twxw = nullptr;
  float2* twxw32 = nullptr;
  // canopy
  unsigned char* masks = nullptr;
  int nmask = 0;
  int *ibu_planes = nullptr, *ibv_planes = nullptr;
  unsigned char *ibflag_u = nullptr, *ibflag_v = nullptr;   // per slot: canopy plane
  int nibu = 0, nibv = 0;
  IBEntry* ibent = nullptr;
  int nibent = 0;
  int nibtip = 0;                 // entries with a Lu_y slot (tip planes), stored first
  // statistics and spectra (allocated on first use)
  double *st_psum = nullptr, *st_pfold = nullptr, *st_acc = nullptr, *cst_acc = nullptr;
  double* sp_acc = nullptr;
  int nstat = 0;
  // host copies
  std::vector<int> h_ibu, h_ibv;
  std::vector<int> mask_of_u, mask_of_v;
  // graphs
  hipGraphExec_t gexec[2] = {nullptr, nullptr};   // [0] fixed dt, [1] device CFL dt
  double graph_dt = -1.0;
  // launch parameters (lh_autotune picks them per instance)
  int xraw_lc = LH_XRAW_LC;
  int zasm_lc = LH_ZASM_LC;
  int col_bs = LH_COL_BS;
  int xplan = 0, zplan = 0;     // FFT plan of the x and z lengths (Plan<n, v>)
  // rows per step of the velocity sweep: 1 on large grids (fewer registers, more resident
  // waves), kVelUSmall when few columns leave the sweeps latency-bound (set at creation)
  int vel_u = 1;
  // per-kernel timing (direct issue, an event pair per launch) or, with timing_steps, one
  // event pair around each replay of the captured step graph
  bool timing = false;
  bool timing_steps = false;
  // batches the handle belongs to: while any, the handle's own graphs keep k_prep on its
  // stream (a batch runs the column stage of all its members itself; a side branch per
  // substep and instance, 48 x 3 cross-stream joins per step, serialised the instances)
  int in_batch = 0;
  // bumped whenever the captured graphs are dropped (launch parameters, forcing switched on
  // or off, batch membership): a batch re-captures when a member's generation changed
  unsigned gen = 0;
  std::vector<hipEvent_t> evpool;
  size_t evnext = 0;
  struct Pend { int k; hipEvent_t a, b; };
  std::vector<Pend> pending;
  double tms[LH_K_COUNT] = {0};
  long long tcnt[LH_K_COUNT] = {0};
  double tbytes[LH_K_COUNT] = {0};
};

// event pair around one launch (timing mode only; never inside graph capture)
static int tbeg(lh_handle* h, int k, double bytes, bool steps = false) {
  if (!(steps ? h->timing_steps : h->timing)) return -1;
  if (h->evnext + 2 > h->evpool.size()) {
    for (int q = 0; q < 64; ++q) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      h->evpool.push_back(e);
    }
  }
  hipEvent_t a = h->evpool[h->evnext++], b = h->evpool[h->evnext++];
  (void)hipEventRecord(a, h->st);
  h->pending.push_back({k, a, b});
  h->tbytes[k] += bytes;
  return (int)h->pending.size() - 1;
}
static void tend(lh_handle* h, int idx) {
  if (idx < 0) return;
  (void)hipEventRecord(h->pending[idx].b, h->st);
}
static void tflush(lh_handle* h) {
  for (auto& q : h->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) {
      h->tms[q.k] += ms;
      h->tcnt[q.k] += 1;
    }
  }
  h->pending.clear();
  h->evnext = 0;
}
#define TIMED(h, K, BYTES, ...)        \
  do {                                 \
    int ti_ = tbeg(h, K, BYTES);       \
    __VA_ARGS__;                       \
    tend(h, ti_);                      \
  } while (0)

static int fail(lh_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, x)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      return fail(h, LH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));            \
  } while (0)

// Non-finite detection after a step (the reference's only monitor is `error`,
// littleharsh.f90:323-352).  The first x pass of every step takes the CFL maximum of the
// state at the start of the step with a reduction that keeps NaN, so a non-finite value
// anywhere in u, v or w shows there.  With need_dt the device CFL time step must also be
// finite and positive.  Reads back the device scalars into s.
static int check_scalars(lh_handle* h, bool need_dt, const double* s) {
  char msg[160];
  if (!std::isfinite(s[S_CFLMAX])) {
    snprintf(msg, sizeof msg, "non-finite velocity at the start of the step (CFL maximum %g)", s[S_CFLMAX]);
    return fail(h, LH_ERR_NUMERIC, msg);
  }
  if (need_dt && !(std::isfinite(s[S_DT]) && s[S_DT] > 0.0)) {
    snprintf(msg, sizeof msg, "CFL time step %g is not finite and positive (CFL maximum %g)", s[S_DT],
             s[S_CFLMAX]);
    return fail(h, LH_ERR_NUMERIC, msg);
  }
  return LH_OK;
}
static int numeric_check(lh_handle* h, bool need_dt, double* s) {
  HIPCHK(h, hipMemcpy(s, h->scal, S_NSCAL * sizeof(double), hipMemcpyDeviceToHost));
  return check_scalars(h, need_dt, s);
}

template <typename T>
static hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, n * sizeof(T) + 64);
}

static bool supported_len(int n) {
  switch (n) {
    case 32: case 48: case 64: case 72: case 96: case 144: case 384: case 576: return true;
    default: return false;
  }
}

// ---------------------------------------------------------------------------------------
// templated launchers, dispatched on the transform lengths
// ---------------------------------------------------------------------------------------
#ifndef LH_ZLINES_LONG
#define LH_ZLINES_LONG 16
#endif
#ifndef LH_ZASM_COLS
#define LH_ZASM_COLS 8
#endif
#ifndef LH_COL_U
#define LH_COL_U 8
#endif
#ifndef LH_VEL_U
#define LH_VEL_U 1
#endif
// Grids of at most kVelSmallCols spectral columns run the velocity sweep kVelUSmall rows per
// step: with one thread per column too few waves are resident to hide each row's loads
// (96x145x96: 1.37 -> 1.24 ms per step; 64x97x64: 0.89 -> 0.82 ms; the pressure sweep's
// unroll and the block size change nothing there).
constexpr int kVelSmallCols = 16384;
constexpr int kVelUSmall = 4;
template <int N, int V> constexpr int zlines() {
  return N >= 384 ? LH_ZLINES_LONG : (Plan<N, V>::T >= 8 ? 16 : 32);
}

template <int PZ, typename TR, int V>
static void launch_zinv_t(lh_handle* h, const ZJobs& jobs, int njobs, int maxn, const void* tw) {
  constexpr int L = zlines<PZ, V>();
  ZJobs jb = jobs;
  jb.maxn = maxn;
  dim3 grid(((h->dm.mx + L - 1) / L) * maxn * njobs);
  size_t sh = (size_t)L * zls<PZ>() * sizeof(cx<TR>);
  hipLaunchKernelGGL((k_zinv<PZ, L, TR, V>), grid, dim3(L * Plan<PZ, V>::T), sh, h->st, jb, h->dm,
                     (const cx<TR>*)tw);
}
template <int PZ, typename TR, int V>
static void launch_zfwd_t(lh_handle* h, const ZJobs& jobs, int njobs, int maxn, const void* tw) {
  constexpr int L = zlines<PZ, V>();
  ZJobs jb = jobs;
  jb.maxn = maxn;
  dim3 grid(((h->dm.mx + L - 1) / L) * maxn * njobs);
  size_t sh = (size_t)L * zls<PZ>() * sizeof(cx<TR>);
  hipLaunchKernelGGL((k_zfwd<PZ, L, TR, V>), grid, dim3(L * Plan<PZ, V>::T), sh, h->st, jb, h->dm,
                     (const cx<TR>*)tw);
}
template <int PZ, typename TR, int V>
static void attr_z_t(int* dummy) {
  constexpr int L = zlines<PZ, V>();
  (void)dummy;
  const int sh = (int)(L * zls<PZ>() * sizeof(cx<TR>));
  (void)hipFuncSetAttribute((const void*)k_zinv<PZ, L, TR, V>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, sh);
  (void)hipFuncSetAttribute((const void*)k_zfwd<PZ, L, TR, V>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, sh);
}
template <int PX, typename TR, int V>
static size_t xraw_lds() {
  return ((size_t)15 * (PadLen<PX, LH_XRAW_PAD>::value + LH_XRAW_LSP) +
          Plan<PX, V>::template TwL<cx<TR>>::SIZE) * sizeof(cx<TR>);
}
template <int PX, typename TR, int V>
static void attr_xraw_t(int* dummy) {
  (void)dummy;
  (void)hipFuncSetAttribute((const void*)k_xraw2<PX, TR, V>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)xraw_lds<PX, TR, V>());
}
// resident blocks per CU of the x pass (LDS and registers)
template <int PX, typename TR, int V>
static void xraw_occ_t(int* nb) {
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, k_xraw2<PX, TR, V>, 6 * Plan<PX, V>::T,
                                                   xraw_lds<PX, TR, V>()) != hipSuccess)
    *nb = 1;
}
template <int PX, typename TR, int V>
static void launch_xraw_t(lh_handle* h, const XRAW& a, const void* tw) {
  const size_t sh = xraw_lds<PX, TR, V>();
  const int nlev = h->dm.nyu - 2, lc = h->xraw_lc;
  dim3 grid(h->dm.pz / 2, (nlev + lc - 1) / lc);
  hipLaunchKernelGGL((k_xraw2<PX, TR, V>), grid, dim3(6 * Plan<PX, V>::T), sh, h->st, a, h->dm,
                     (const cx<TR>*)tw, lc);
}
// the one-wave-per-line x pass: one block per CU (LDS), LH_XW_WAVES32 waves per block in FP32
// (8: two per SIMD; 12: three, with a few spills), 8 in FP64 (two per SIMD, the registers'
// limit); each wave takes one row pair and xraw_lc levels
#ifndef LH_XW_WAVES32
#define LH_XW_WAVES32 8
#endif
template <typename TR> struct XwWaves { static constexpr int value = sizeof(TR) == 8 ? 8 : LH_XW_WAVES32; };
static bool use_xw(const lh_handle* h) { return h->dm.px == 576; }
template <typename TR>
static void launch_xw_t(lh_handle* h, int want_cfl) {
  XW a{h->zu, h->zw, h->zv, h->zp, sizeof(TR) == 8 ? (const void*)h->twxw : (const void*)h->twxw32,
       h->dthdyu, h->scal, want_cfl, h->xraw_lc};
  const int nlev = h->dm.nyu - 2;
  const long waves = (long)(h->dm.px / 2) * ((nlev + h->xraw_lc - 1) / h->xraw_lc);
  constexpr int W = XwWaves<TR>::value;
  const dim3 grid((unsigned)((waves + W - 1) / W)), block(64 * W);
  const size_t sh = xw_lds_bytes<W, TR>();
  if (want_cfl)
    hipLaunchKernelGGL((k_xw576<W, true, TR>), grid, block, sh, h->st, a, h->dm);
  else
    hipLaunchKernelGGL((k_xw576<W, false, TR>), grid, block, sh, h->st, a, h->dm);
}
static void launch_xw(lh_handle* h, int want_cfl) {
  if (h->prec == LH_PREC_FP64) launch_xw_t<double>(h, want_cfl);
  else launch_xw_t<float>(h, want_cfl);
}
template <typename TR>
static hipError_t attr_xw_t() {
  constexpr int W = XwWaves<TR>::value;
  hipError_t e = hipFuncSetAttribute((const void*)k_xw576<W, true, TR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)xw_lds_bytes<W, TR>());
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_xw576<W, false, TR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)xw_lds_bytes<W, TR>());
}
// the canopy x pass at 576 points: one wave per (entry, row pair), LH_XIBW_WAVES per block
// (12: three per SIMD at 132 VGPRs, 121 KB of LDS; 16 would spill)
#ifndef LH_XIBW_WAVES
#define LH_XIBW_WAVES 12
#endif
static size_t xibw_lds_bytes() { return sizeof(double2) * ((size_t)xw::TWN + (size_t)LH_XIBW_WAVES * xw::LINE); }
static void launch_xib576(lh_handle* h, const XIB& a, int nent) {
  constexpr int W = LH_XIBW_WAVES;
  const long waves = (long)nent * (h->dm.pz / 2);
  hipLaunchKernelGGL((k_xpass_ib576<W>), dim3((unsigned)((waves + W - 1) / W)), dim3(64 * W), xibw_lds_bytes(),
                     h->st, a, h->dm, nent, (const double2*)h->twxw);
}
constexpr int kZasmCols = LH_ZASM_COLS;
template <int PZ, typename TR, int V>
static size_t zasm_lds() {
  return ((size_t)2 * kZasmCols * (PZ + 1) + ZasmTw<PZ, cx<TR>, V>::SIZE) * sizeof(cx<TR>);
}
template <int PZ, typename TR, int V>
static void launch_zasm_t(lh_handle* h, const ZASM& a, const void* tw) {
  constexpr int L = kZasmCols;
  const int nlev = h->dm.nyu - 2;
  dim3 grid(a.nib * ((nlev + a.lc - 1) / a.lc) * 3);
  const size_t sh = zasm_lds<PZ, TR, V>();
  hipLaunchKernelGGL((k_zasm<PZ, L, TR, V>), grid, dim3(2 * L * Plan<PZ, V>::T), sh, h->st, a, h->dm,
                     (const cx<TR>*)tw);
}
template <int PZ, typename TR, int V>
static void zasm_occ_t(int* nb) {
  constexpr int L = kZasmCols;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, k_zasm<PZ, L, TR, V>, 2 * L * Plan<PZ, V>::T,
                                                   zasm_lds<PZ, TR, V>()) != hipSuccess)
    *nb = 1;
}
template <int PZ, typename TR, int V>
static void attr_zasm_t(int* dummy) {
  constexpr int L = kZasmCols;
  (void)dummy;
  (void)hipFuncSetAttribute((const void*)k_zasm<PZ, L, TR, V>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)zasm_lds<PZ, TR, V>());
}
template <int PX, typename TR, int V>
static void launch_xib_t(lh_handle* h, const XIB& a, int nent, const void* tw) {
  dim3 grid(h->dm.pz / (a.paired ? 4 : 2), nent);
  size_t sh = (size_t)2 * (PX + 1) * sizeof(cx<TR>);
  hipLaunchKernelGGL((k_xpass_ib<PX, TR, V>), grid, dim3(2 * Plan<PX, V>::T), sh, h->st, a, h->dm,
                     (const cx<TR>*)tw);
}

template <int PX, typename TR>
static void launch_stats_t(lh_handle* h, const STATS& a) {
  dim3 grid(a.dnz, a.n);
  // lines, twiddle rows, the fold [6][dnx] and the fold parts [parts][6][dnx]
  const size_t sh = stats_lds_bytes<PX>(a.dnx, a.ntx);
  // more than 64 KB when a tile is wide (dnx = px without a canopy lattice)
  (void)hipFuncSetAttribute((const void*)k_stats<PX, TR>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
  hipLaunchKernelGGL((k_stats<PX, TR>), grid, dim3(3 * Plan<PX>::T), sh, h->st, a, h->dm,
                     (const cx<double>*)h->twx64);
}

#define LH_DISPATCH_LEN(n, F, TR, ...)          \
  switch (n) {                                   \
    case 32: F<32, TR>(__VA_ARGS__); break;      \
    case 48: F<48, TR>(__VA_ARGS__); break;      \
    case 64: F<64, TR>(__VA_ARGS__); break;      \
    case 72: F<72, TR>(__VA_ARGS__); break;      \
    case 96: F<96, TR>(__VA_ARGS__); break;      \
    case 144: F<144, TR>(__VA_ARGS__); break;    \
    case 384: F<384, TR>(__VA_ARGS__); break;    \
    case 576: F<576, TR>(__VA_ARGS__); break;    \
    default: break;                              \
  }
// the same for the launchers templated on a plan: v selects Plan<n, v> where n has two
#define LH_PLAN_CASE2(N, v, F, TR, ...)                          \
  case N:                                                         \
    if (v) F<N, TR, 1>(__VA_ARGS__); else F<N, TR, 0>(__VA_ARGS__); \
    break;
#define LH_DISPATCH_PLAN(n, v, F, TR, ...)         \
  switch (n) {                                      \
    case 32: F<32, TR, 0>(__VA_ARGS__); break;      \
    LH_PLAN_CASE2(48, v, F, TR, __VA_ARGS__)        \
    case 64: F<64, TR, 0>(__VA_ARGS__); break;      \
    LH_PLAN_CASE2(72, v, F, TR, __VA_ARGS__)        \
    LH_PLAN_CASE2(96, v, F, TR, __VA_ARGS__)        \
    LH_PLAN_CASE2(144, v, F, TR, __VA_ARGS__)       \
    case 384: F<384, TR, 0>(__VA_ARGS__); break;    \
    case 576: F<576, TR, 0>(__VA_ARGS__); break;    \
    default: break;                                 \
  }
// the same over the lengths that run from job tables (up to 144)
#define LH_DISPATCH_PLAN_SMALL(n, v, F, TR, ...)    \
  switch (n) {                                      \
    case 32: F<32, TR, 0>(__VA_ARGS__); break;      \
    LH_PLAN_CASE2(48, v, F, TR, __VA_ARGS__)        \
    case 64: F<64, TR, 0>(__VA_ARGS__); break;      \
    LH_PLAN_CASE2(72, v, F, TR, __VA_ARGS__)        \
    LH_PLAN_CASE2(96, v, F, TR, __VA_ARGS__)        \
    LH_PLAN_CASE2(144, v, F, TR, __VA_ARGS__)       \
    default: break;                                 \
  }
// number of FFT plans compiled for length n
static int plans_of_len(int n) {
  switch (n) {
    case 48: return plan_count<48>();
    case 72: return plan_count<72>();
    case 96: return plan_count<96>();
    case 144: return plan_count<144>();
    default: return 1;
  }
}

static void launch_zinv(lh_handle* h, const ZJobs& jobs, int njobs, int maxn, bool fp64) {
  if (fp64 || h->prec == LH_PREC_FP64) {
    LH_DISPATCH_PLAN(h->dm.pz, h->zplan, launch_zinv_t, double, h, jobs, njobs, maxn, fp64 ? h->twz64 : h->twz)
  } else {
    LH_DISPATCH_PLAN(h->dm.pz, h->zplan, launch_zinv_t, float, h, jobs, njobs, maxn, h->twz)
  }
}
static void launch_zfwd(lh_handle* h, const ZJobs& jobs, int njobs, int maxn, bool fp64) {
  if (fp64 || h->prec == LH_PREC_FP64) {
    LH_DISPATCH_PLAN(h->dm.pz, h->zplan, launch_zfwd_t, double, h, jobs, njobs, maxn, fp64 ? h->twz64 : h->twz)
  } else {
    LH_DISPATCH_PLAN(h->dm.pz, h->zplan, launch_zfwd_t, float, h, jobs, njobs, maxn, h->twz)
  }
}

// ---------------------------------------------------------------------------------------
// substep
// ---------------------------------------------------------------------------------------
// Nu_new += forcing, field blockIdx.y (0: u, 1: w, 2: v), whole plane ranges
__global__ void k_add_forcing(double2* n1, double2* n2, double2* n3, const double2* f1,
                              const double2* f2, const double2* f3, size_t nu, size_t nv) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  double2* n = f == 0 ? n1 : (f == 1 ? n3 : n2);
  const double2* g = f == 0 ? f1 : (f == 1 ? f3 : f2);
  if (i < (f == 2 ? nv : nu)) {
    const double2 a = n[i], b = g[i];
    n[i] = make_double2(a.x + b.x, a.y + b.y);
  }
}

// stages of a substep: the nonlinear term, the velocity update (the canopy forcing, then
// RHS0 and solveU/V fused in the sweep), the pressure step (ctU, solveP, v_corr, fused).
// The nonlinear and the canopy stage are split into their launches (z, x, assembly; rhs,
// z, x, z) so that a batch can issue each launch for all its members in turn.
constexpr int kStageNLz = 1, kStageIBr = 2, kStagePres = 4, kStageVelCol = 8, kStageNLx = 16,
              kStageNLa = 32, kStageIBz = 64, kStageIBx = 128, kStageIBf = 256;
// with a stage: the caller clears the CFL maximum and forms the time step of a CFL step
// (a batch does both for all its members in one launch each)
constexpr int kCallerCfl = 512;
constexpr int kStageNL = kStageNLz | kStageNLx | kStageNLa;
constexpr int kStageIB = kStageIBr | kStageIBz | kStageIBx | kStageIBf;
constexpr int kStageVel = kStageIB | kStageVelCol, kStageAll = kStageNL | kStageVel | kStagePres;

// parameters of the column stage of substep krk (0-based); the batch builds them for every
// member and launches the column kernels of all members at once
static void nu_buffers(const lh_handle* h, int krk, double2* nw[3], double2* ol[3]) {
  // Nu_new into the buffer of this substep; RHS0 + rhs per column reads both
  const bool wb = (krk % 2) == 1;
  nw[0] = wb ? h->nb1 : h->nu1; nw[1] = wb ? h->nb2 : h->nu2; nw[2] = wb ? h->nb3 : h->nu3;
  ol[0] = wb ? h->nu1 : h->nb1; ol[1] = wb ? h->nu2 : h->nb2; ol[2] = wb ? h->nu3 : h->nb3;
}
static PREP make_prep(const lh_handle* h, int krk) {
  const Dims& dm = h->dm;
  const double bre = h->rk[3 + krk] / dm.re;
  return PREP{h->dyu2i, h->dyv2i, h->scal, h->ayu, h->ayv, h->gw[0], h->gw[1], bre, dm.nyu, dm.nyv};
}
static RHS make_rhs(const lh_handle* h, int krk) {
  const Dims& dm = h->dm;
  const double aR = h->rk[krk], bR = h->rk[3 + krk], gR = h->rk[6 + krk], cR = h->rk[9 + krk],
               dR = h->rk[12 + krk];
  double2 *nw[3], *ol[3];
  nu_buffers(h, krk, nw, ol);
  RHS ra{};
  ra.u1 = h->u1; ra.u2 = h->u2; ra.u3 = h->u3; ra.p = h->p;
  ra.no1 = ol[0]; ra.no2 = ol[1]; ra.no3 = ol[2];
  ra.nn1 = nw[0]; ra.nn2 = nw[1]; ra.nn3 = nw[2];
  ra.r1 = h->r1; ra.r2 = h->r2; ra.r3 = h->r3; ra.tipl = h->tipl;
  ra.dyu2i = h->dyu2i; ra.dyv2i = h->dyv2i; ra.dthdyv = h->dthdyv;
  ra.k1x = h->k1x; ra.k2x = h->k2x; ra.k1z = h->k1z; ra.k2z = h->k2z; ra.scal = h->scal;
  ra.planes_u = h->ibu_planes; ra.planes_v = h->ibv_planes;
  ra.ib_u = h->ibflag_u; ra.ib_v = h->ibflag_v;
  ra.np_u = h->nibu; ra.np_v = h->nibv;
  ra.tip_u0 = h->tip_u[0] + dm.d; ra.tip_u1 = h->tip_u[1] + dm.d;
  ra.tip_v0 = h->tip_v[0] + dm.d; ra.tip_v1 = h->tip_v[1] + dm.d;
  ra.C1 = aR / dm.re; ra.C2 = -cR; ra.dR = dR; ra.g = gR; ra.bre = bR / dm.re;
  return ra;
}
static COL make_col(const lh_handle* h, int krk) {
  const Dims& dm = h->dm;
  const double aR = h->rk[krk], bR = h->rk[3 + krk], dR = h->rk[12 + krk];
  COL a{};
  a.r1 = h->r1; a.r2 = h->r2; a.r3 = h->r3;
  a.u1 = h->u1; a.u2 = h->u2; a.u3 = h->u3; a.p = h->p; a.psi = h->psi;
  a.ayu = h->ayu; a.ayv = h->ayv; a.dg3 = h->dg3;
  a.k2x = h->k2x; a.k2z = h->k2z; a.yu = h->yu;
  a.dthdyu = h->dthdyu; a.dthdyv = h->dthdyv; a.k1x = h->k1x; a.k1z = h->k1z;
  a.scal = h->scal; a.ab = aR + bR; a.gw0 = h->gw[0]; a.gw1 = h->gw[1];
  a.ctu = h->flag_ctpress == 0 ? 1 : 0;
  a.ct = CTU{h->u1, h->scal, h->ctwork, h->dyu2i, h->yv, h->gw[0], h->gw[1], bR / dm.re, dR, dm.ny,
             dm.d, dm.ncol};
  return a;
}

// The column stage (k_prep, the velocity sweeps, the pressure step, the flow rate) runs from
// job tables: a job holds one channel's parameters for one substep, and block b of a launch
// works on channel colmap[b], columns (b - blk0) * 64 + lane.  A single channel launches
// them with its own table (one job per substep); a batch launches them once for all its
// members.  Small channels have a few thousand spectral columns, so one channel's column
// kernels fill a few percent of the device for the time of their sequential sweep over the
// planes; launched together, the members' sweeps run side by side.  Both launch the same
// kernels, so a batch member computes exactly what the channel computes alone.
struct ColJob {
  RHS ra;
  COL ca;
  PREP prep;
  Dims dm;
  const double* yv;
  int vel_u;        // rows per step of the member's velocity sweep
  int blk0;         // first block of the member in the launches of its sweep kernels
  int flowrate;     // flag_ctpress == 1: flowrateRe after the step
};
constexpr int kMultiCols = 64;
__global__ void __launch_bounds__(64) k_prep(const ColJob* __restrict__ jobs) {
  prep_body(jobs[blockIdx.x].prep, threadIdx.x);
}
// one kernel per row unroll of the velocity sweep (their register needs differ: 72 VGPRs at
// one row per step, 150 at four); a batch orders its members' blocks by unroll and launches
// each variant over its range, blocks boff, boff + 1, ...
template <int U>
__global__ void __launch_bounds__(kMultiCols) k_vel_rhs(const ColJob* __restrict__ jobs,
                                                          const int* __restrict__ colmap, int boff) {
  const int b = blockIdx.x + boff;
  const ColJob& J = jobs[colmap[b]];
  const Dims dm = J.dm;                // loaded once, ahead of the component branches
  const int c = (b - J.blk0) * kMultiCols + threadIdx.x;
  if (c >= dm.ncol) return;
  const RHS ra = J.ra;
  vel_column<U>(ra, dm, c, blockIdx.y, J.ca.ayu, J.ca.ayv);
}
__global__ void __launch_bounds__(kMultiCols) k_pres(const ColJob* __restrict__ jobs,
                                                           const int* __restrict__ colmap) {
  const ColJob& J = jobs[colmap[blockIdx.x]];
  const Dims dm = J.dm;
  const int c = (blockIdx.x - J.blk0) * kMultiCols + threadIdx.x;
  if (c >= dm.ncol) return;
  const COL ca = J.ca;
  pres_column<LH_COL_U>(ca, dm, c);
}
__global__ void __launch_bounds__(256) k_flowrate(const ColJob* __restrict__ jobs) {
  __shared__ double part[256];
  const ColJob& J = jobs[blockIdx.x];
  if (!J.flowrate) return;
  flowrate_body(J.ca.u1, J.yv, J.ca.scal, J.dm.ny, J.dm.d, J.dm.ncol, part);
}
// The scalars of the n channels of a job table, one thread per channel: the set-up of a CFL
// step (the CFL number, and the maximum cleared for the x pass to accumulate), its time
// step (the division k_cfl_dt makes), and a copy of every scalar into out[n][S_NSCAL].
__global__ void k_jobs_cfl_begin(const ColJob* __restrict__ jobs, int n, double cfl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const gptr<double> s = gp(jobs[i].ca.scal);
  s[S_CFL] = cfl;
  s[S_CFLMAX] = 0.0;
}
__global__ void k_jobs_cfl_dt(const ColJob* __restrict__ jobs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const gptr<double> s = gp(jobs[i].ca.scal);
  s[S_DT] = s[S_CFL] / s[S_CFLMAX];
}
__global__ void k_jobs_scalars(const ColJob* __restrict__ jobs, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const gptr<const double> s = gp(jobs[i].ca.scal);
  for (int q = 0; q < S_NSCAL; ++q) out[(size_t)i * S_NSCAL + q] = s[q];
}

static ColJob make_job(const lh_handle* h, int krk, int blk0) {
  ColJob J;
  J.ra = make_rhs(h, krk);
  J.ca = make_col(h, krk);
  J.prep = make_prep(h, krk);
  J.dm = h->dm;
  J.yv = h->yv;
  J.vel_u = h->vel_u;
  J.blk0 = blk0;
  J.flowrate = h->flag_ctpress == 1 ? 1 : 0;
  return J;
}
static int col_blocks(const lh_handle* h) { return (h->dm.ncol + kMultiCols - 1) / kMultiCols; }

// ---- the canopy stage from job tables (channels whose transform lengths are all <= 144)
// A channel of that size runs its canopy stage from tables, alone or in a batch; within a
// batch the channels that share the kernel, the length and the plan form one launch.
static bool small_channel(const lh_handle* h) { return h->dm.px <= 144 && h->dm.pz <= 144; }
static bool ib_jobs(const lh_handle* h) { return small_channel(h) && h->nibent > 0; }
template <int N, typename TR, int V> static void zlines_of_t(int* l) { *l = zlines<N, V>(); }
static int zlines_of(int n, int v) {
  int l = 16;
  LH_DISPATCH_PLAN_SMALL(n, v, zlines_of_t, double, &l)
  return l;
}

// Groups of one kernel: job units that share (length, plan), a unit being one job of a
// channel (a channel has nu of them, u = 0 .. nu-1); per group a job table per substep
// (nk = 3) or one for all substeps (nk = 1), and one block map.
template <typename J>
struct GroupSet {
  struct Unit { lh_handle* h; int u; };
  struct G { int len, plan; std::vector<Unit> mem; size_t job0[3]; size_t map0; int nblk; };
  std::vector<G> gs;
  J* djobs = nullptr;
  int* dmap = nullptr;
  // key(h, u, &len, &plan) -> whether unit u of the channel takes part; blocks(h, u);
  // make(h, u, k, blk0)
  template <class Key, class Blocks, class Make>
  hipError_t build(const std::vector<lh_handle*>& hs, int nu, int nk, Key key, Blocks blocks, Make make) {
    release();
    for (int i = 0; i < (int)hs.size(); ++i)
      for (int u = 0; u < nu; ++u) {
        int len = 0, plan = 0;
        if (!key(hs[i], u, &len, &plan)) continue;
        size_t g = 0;
        while (g < gs.size() && !(gs[g].len == len && gs[g].plan == plan)) ++g;
        if (g == gs.size()) gs.push_back(G{len, plan, {}, {0, 0, 0}, 0, 0});
        gs[g].mem.push_back(Unit{hs[i], u});
      }
    std::vector<J> jobs;
    std::vector<int> map;
    for (auto& G : gs) {
      for (int k = 0; k < nk; ++k) {
        G.job0[k] = jobs.size();
        int blk = 0;
        for (const Unit& m : G.mem) {
          jobs.push_back(make(m.h, m.u, k, blk));
          blk += blocks(m.h, m.u);
        }
        G.nblk = blk;
      }
      for (int k = nk; k < 3; ++k) G.job0[k] = G.job0[0];
      G.map0 = map.size();
      for (size_t j = 0; j < G.mem.size(); ++j) map.insert(map.end(), blocks(G.mem[j].h, G.mem[j].u), (int)j);
    }
    if (jobs.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void**)&djobs, jobs.size() * sizeof(J));
    if (e == hipSuccess) e = hipMalloc((void**)&dmap, map.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(djobs, jobs.data(), jobs.size() * sizeof(J), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dmap, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice);
    return e;
  }
  void release() {
    if (djobs) (void)hipFree(djobs);
    if (dmap) (void)hipFree(dmap);
    djobs = nullptr;
    dmap = nullptr;
    gs.clear();
  }
};

template <int PZ, typename TR, int V>
static void launch_zinv_jobs_t(hipStream_t st, const ZJobT* jt, const int* map, int nblk) {
  constexpr int L = zlines<PZ, V>();
  const size_t sh = (size_t)L * zls<PZ>() * sizeof(cx<TR>);
  hipLaunchKernelGGL((k_zinv_jobs<PZ, L, TR, V>), dim3(nblk), dim3(L * Plan<PZ, V>::T), sh, st, jt, map);
}
template <int PZ, typename TR, int V>
static void launch_zfwd_jobs_t(hipStream_t st, const ZJobT* jt, const int* map, int nblk) {
  constexpr int L = zlines<PZ, V>();
  const size_t sh = (size_t)L * zls<PZ>() * sizeof(cx<TR>);
  hipLaunchKernelGGL((k_zfwd_jobs<PZ, L, TR, V>), dim3(nblk), dim3(L * Plan<PZ, V>::T), sh, st, jt, map);
}
template <int PX, typename TR, int V>
static void launch_xib_jobs_t(hipStream_t st, const XibJob* jt, const int* map, int nblk) {
  const size_t sh = (size_t)2 * (PX + 1) * sizeof(cx<TR>);
  hipLaunchKernelGGL((k_xpass_ib_jobs<PX, TR, V>), dim3(nblk), dim3(2 * Plan<PX, V>::T), sh, st, jt, map);
}

// the canopy-stage jobs of one channel (the launches of the by-value path, same arguments)
static int ib_npl(const lh_handle* h) { return h->nibu > h->nibv ? h->nibu : h->nibv; }
static int ib_maxn(const lh_handle* h) { return ib_npl(h) > 6 ? ib_npl(h) : 6; }
static ZJobs ib_zinv_jobs(const lh_handle* h) {
  ZJobs jobs{};
  int slot = 0;
  jobs.j[0] = ZJob{h->r1, h->zp, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
  jobs.j[1] = ZJob{h->r3, h->zp, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
  jobs.j[2] = ZJob{h->r2, h->zp, h->ibv_planes, 0, h->nibv, slot}; slot += h->nibv;
  jobs.j[3] = ZJob{h->tipl, h->zp, nullptr, 0, 6, slot};
  jobs.maxn = ib_maxn(h);
  return jobs;
}
static ZJobs ib_zfwd_jobs(const lh_handle* h) {
  ZJobs fj{};
  int slot = 0;
  fj.j[0] = ZJob{(const double2*)h->zp, h->r1, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
  fj.j[1] = ZJob{(const double2*)h->zp, h->r3, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
  fj.j[2] = ZJob{(const double2*)h->zp, h->r2, h->ibv_planes, 0, h->nibv, slot};
  fj.maxn = ib_npl(h);
  return fj;
}
static XIB ib_xib(const lh_handle* h, int krk, int paired) {
  const double bre = h->rk[3 + krk] / h->dm.re;
  return paired ? XIB{h->zp, h->ibent + h->nibtip, h->masks, h->scal, bre, 1}
                : XIB{h->zp, h->ibent, h->masks, h->scal, bre, 0};
}
static int zgrid_cols(const lh_handle* h) { return (h->dm.mx + zlines_of(h->dm.pz, h->zplan) - 1) / zlines_of(h->dm.pz, h->zplan); }

struct IbSets {
  GroupSet<RhsIbJob> ri;
  GroupSet<ZJobT> zi, zf;
  GroupSet<XibJob> xb;
  hipError_t build(const std::vector<lh_handle*>& hs) {
    hipError_t e = ri.build(
        hs, 1, 3, [](const lh_handle* h, int, int* l, int* v) { *l = 0; *v = 0; return ib_jobs(h); },
        [](const lh_handle* h, int) { return ((h->dm.ncol + 63) / 64) * ib_npl(h) * 3; },
        [](const lh_handle* h, int, int k, int blk0) {
          return RhsIbJob{make_rhs(h, k), h->dm, blk0, (h->dm.ncol + 63) / 64, ib_npl(h)};
        });
    if (e == hipSuccess)
      e = zi.build(
          hs, 1, 1, [](const lh_handle* h, int, int* l, int* v) { *l = h->dm.pz; *v = h->zplan; return ib_jobs(h); },
          [](const lh_handle* h, int) { return zgrid_cols(h) * ib_maxn(h) * 4; },
          [](const lh_handle* h, int, int, int blk0) {
            return ZJobT{ib_zinv_jobs(h), h->dm, h->twz64, blk0, zgrid_cols(h) * ib_maxn(h) * 4};
          });
    if (e == hipSuccess)
      e = zf.build(
          hs, 1, 1, [](const lh_handle* h, int, int* l, int* v) { *l = h->dm.pz; *v = h->zplan; return ib_jobs(h); },
          [](const lh_handle* h, int) { return zgrid_cols(h) * ib_npl(h) * 3; },
          [](const lh_handle* h, int, int, int blk0) {
            return ZJobT{ib_zfwd_jobs(h), h->dm, h->twz64, blk0, zgrid_cols(h) * ib_npl(h) * 3};
          });
    // x pass: unit 0 the tip entries (one row pair per block), unit 1 the others (two), in one
    // launch per group
    if (e == hipSuccess)
      e = xb.build(
          hs, 2, 3,
          [](const lh_handle* h, int u, int* l, int* v) {
            *l = h->dm.px; *v = h->xplan;
            return ib_jobs(h) && (u == 0 ? h->nibtip > 0 : h->nibent > h->nibtip);
          },
          [](const lh_handle* h, int u) {
            return u == 0 ? (h->dm.pz / 2) * h->nibtip : (h->dm.pz / 4) * (h->nibent - h->nibtip);
          },
          [](const lh_handle* h, int u, int k, int blk0) {
            return XibJob{ib_xib(h, k, u), h->dm, h->twx64, blk0, h->dm.pz / (u == 0 ? 2 : 4)};
          });
    return e;
  }
  void release() { ri.release(); zi.release(); zf.release(); xb.release(); }
  // the launches of the canopy stage of substep k of every channel in the sets, in stage
  // order: rhs, inverse z, x (tip and other entries together), forward z
  void launch_rhs(int k, hipStream_t st) const {
    for (const auto& G : ri.gs)
      hipLaunchKernelGGL(k_rhs_ib_jobs, dim3(G.nblk), dim3(64), 0, st, ri.djobs + G.job0[k], ri.dmap + G.map0);
  }
  void launch_zinv(int k, hipStream_t st) const {
    for (const auto& G : zi.gs)
      LH_DISPATCH_PLAN_SMALL(G.len, G.plan, launch_zinv_jobs_t, double, st, zi.djobs + G.job0[k], zi.dmap + G.map0, G.nblk)
  }
  void launch_x(int k, hipStream_t st) const {
    for (const auto& G : xb.gs)
      LH_DISPATCH_PLAN_SMALL(G.len, G.plan, launch_xib_jobs_t, double, st, xb.djobs + G.job0[k], xb.dmap + G.map0, G.nblk)
  }
  void launch_zfwd(int k, hipStream_t st) const {
    for (const auto& G : zf.gs)
      LH_DISPATCH_PLAN_SMALL(G.len, G.plan, launch_zfwd_jobs_t, double, st, zf.djobs + G.job0[k], zf.dmap + G.map0, G.nblk)
  }
  void launch(int k, hipStream_t st) const {
    launch_rhs(k, st);
    launch_zinv(k, st);
    launch_x(k, st);
    launch_zfwd(k, st);
  }
};
// (re)build a small channel's own canopy tables: at creation and whenever its plans change
static hipError_t refresh_ib(lh_handle* h) {
  if (!ib_jobs(h)) return hipSuccess;
  if (!h->ib) h->ib = new IbSets();
  return h->ib->build(std::vector<lh_handle*>{h});
}

static void issue_substep(lh_handle* h, int krk /*0-based*/, bool cfl_mode, int stages = kStageAll) {
  const Dims& dm = h->dm;
  const double bre = h->rk[3 + krk] / dm.re;
  const int nlev = dm.nyu - 2;
  const bool f64 = h->prec == LH_PREC_FP64;
  const double zel = f64 ? 16.0 : 8.0;                       // bytes per Z element
  const double zsl = zel * dm.pz * dm.mx;                    // bytes per Z slot
  const double spl = 16.0 * dm.ncol;                         // bytes per spectral plane
  const double nuvw = 2.0 * dm.nyu + dm.nyv;                 // u, w, v planes
  // k_prep needs only dt: in graph and direct issue it runs on the side stream beside the
  // nonlinear stage (timing mode and batch members keep it in line); at kRK = 1 of a CFL
  // step dt comes from the first x pass, so it forks after k_cfl_dt
  const bool side = !h->timing && h->in_batch == 0 && stages == kStageAll;
  auto fork_prep = [&]() {
    (void)hipEventRecord(h->ev_fork, h->st);
    (void)hipStreamWaitEvent(h->side, h->ev_fork, 0);
    hipLaunchKernelGGL(k_prep, dim3(1), dim3(64), 0, h->side, (const ColJob*)h->coljobs + krk);
    (void)hipEventRecord(h->ev_join, h->side);
  };
  const bool late = krk == 0 && cfl_mode;
  if (side && !late) fork_prep();
  double2 *nw[3], *ol[3];
  nu_buffers(h, krk, nw, ol);
  const bool wb = (krk % 2) == 1;
  if (stages & kStageNL) {
  // ---- NL stage: inverse z transforms of u, w (all u planes) and v (all v planes)
  if (stages & kStageNLz) {
    ZJobs jobs{};
    jobs.j[0] = ZJob{h->u1, h->zu, nullptr, 0, dm.nyu, 0};
    jobs.j[1] = ZJob{h->u3, h->zw, nullptr, 0, dm.nyu, 0};
    jobs.j[2] = ZJob{h->u2, h->zv, nullptr, 0, dm.nyv, 0};
    TIMED(h, LH_K_ZINV, nuvw * (spl + zsl), launch_zinv(h, jobs, 3, dm.nyu, false));
  }
  if (stages & kStageNLx) {
  if (krk == 0 && !(stages & kCallerCfl))
    TIMED(h, LH_K_MISC, 0.0,
          hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_CFLMAX, 0.0));
  {
    XRAW a{h->zu, h->zw, h->zv, h->zp, h->dthdyu, h->scal, krk == 0 ? 1 : 0};
    TIMED(h, LH_K_XNL, nuvw * zsl + 6.0 * nlev * zsl, {
      if (use_xw(h)) {
        launch_xw(h, krk == 0 ? 1 : 0);
      } else if (f64) {
        LH_DISPATCH_PLAN(dm.px, h->xplan, launch_xraw_t, double, h, a, h->twx)
      } else {
        LH_DISPATCH_PLAN(dm.px, h->xplan, launch_xraw_t, float, h, a, h->twx)
      }
    });
  }
  if (krk == 0 && cfl_mode && !(stages & kCallerCfl))
    TIMED(h, LH_K_MISC, 0.0, hipLaunchKernelGGL(k_cfl_dt, dim3(1), dim3(1), 0, h->st, h->scal));
  if (side && late) fork_prep();
  }
  if (stages & kStageNLa) {
  h->nu_last = wb ? 1 : 0;
  {
    ZASM a{};
    a.zo = h->zp;
    a.nu1 = nw[0]; a.nu2 = nw[1]; a.nu3 = nw[2];
    a.dthdyu = h->dthdyu; a.dthdyv = h->dthdyv; a.k1x = h->k1x; a.k1z = h->k1z;
    a.nib = (dm.mx + kZasmCols - 1) / kZasmCols;
    a.lc = h->zasm_lc;
    // in: six product Z rows per level; out: Nu_new
    const double by = 6.0 * nlev * zsl + spl * (2.0 * nlev + (nlev - 1));
    TIMED(h, LH_K_ASM, by, {
      if (f64) {
        LH_DISPATCH_PLAN(dm.pz, h->zplan, launch_zasm_t, double, h, a, h->twz)
      } else {
        LH_DISPATCH_PLAN(dm.pz, h->zplan, launch_zasm_t, float, h, a, h->twz)
      }
    });
  }
  if (h->forcing) {
    const size_t nu = (size_t)dm.nyu * dm.ncol, nv = (size_t)dm.nyv * dm.ncol;
    TIMED(h, LH_K_MISC, 3.0 * nuvw * spl,
          hipLaunchKernelGGL(k_add_forcing, dim3((unsigned)((nu + 255) / 256), 3), dim3(256), 0,
                             h->st, nw[0], nw[1], nw[2], (const double2*)h->fo1,
                             (const double2*)h->fo2, (const double2*)h->fo3, nu, nv));
  }
  }
  }   // NL stage
  const RHS ra = make_rhs(h, krk);
  // ---- IB stage (FP64 transforms): canopy planes of rhs u, w, v and Lu_y on tip planes
  if ((stages & kStageIB) && ib_jobs(h) && h->ib) {
    const double z64 = 16.0 * dm.pz * dm.mx;
    const double nib = 2.0 * h->nibu + h->nibv;
    if (stages & kStageIBr) TIMED(h, LH_K_RHS, spl * 7.0 * nib, h->ib->launch_rhs(krk, h->st));
    if (stages & kStageIBz) TIMED(h, LH_K_IBZ, (nib + 6.0) * (spl + z64), h->ib->launch_zinv(krk, h->st));
    if (stages & kStageIBx) TIMED(h, LH_K_XIB, (2.0 * nib + 6.0) * z64, h->ib->launch_x(krk, h->st));
    if (stages & kStageIBf) TIMED(h, LH_K_IBZ, nib * (spl + z64), h->ib->launch_zfwd(krk, h->st));
  } else if ((stages & kStageIB) && h->nibent > 0) {
    const int npl = h->nibu > h->nibv ? h->nibu : h->nibv;
    // in: u (3 rows), p, Nu_old, Nu_new per canopy plane; out: rhs (+ Lu_y of the tips)
    if (stages & kStageIBr)
    TIMED(h, LH_K_RHS, spl * 7.0 * (2.0 * h->nibu + h->nibv),
          hipLaunchKernelGGL(k_rhs_ib, dim3((dm.ncol + h->col_bs - 1) / h->col_bs, npl, 3),
                             dim3(h->col_bs), 0, h->st, ra, dm));
    const double z64 = 16.0 * dm.pz * dm.mx;
    const double nib = 2.0 * h->nibu + h->nibv;
    ZJobs jobs{};
    int slot = 0;
    jobs.j[0] = ZJob{h->r1, h->zp, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
    jobs.j[1] = ZJob{h->r3, h->zp, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
    jobs.j[2] = ZJob{h->r2, h->zp, h->ibv_planes, 0, h->nibv, slot}; slot += h->nibv;
    jobs.j[3] = ZJob{h->tipl, h->zp, nullptr, 0, 6, slot};
    int maxn = h->nibu > h->nibv ? h->nibu : h->nibv;
    if (maxn < 6) maxn = 6;
    if (stages & kStageIBz)
    TIMED(h, LH_K_IBZ, (nib + 6.0) * (spl + z64), launch_zinv(h, jobs, 4, maxn, true));
    // entries are sorted tip first: the tips take one row pair per block, the rest two
    XIB a{h->zp, h->ibent, h->masks, h->scal, bre, 0};
    XIB b{h->zp, h->ibent + h->nibtip, h->masks, h->scal, bre, 1};
    if (stages & kStageIBx)
    TIMED(h, LH_K_XIB, (2.0 * nib + 6.0) * z64, {
      if (use_xw(h)) {
        if (h->nibtip > 0) launch_xib576(h, a, h->nibtip);
        if (h->nibent > h->nibtip) launch_xib576(h, b, h->nibent - h->nibtip);
      } else {
        if (h->nibtip > 0) LH_DISPATCH_PLAN(dm.px, h->xplan, launch_xib_t, double, h, a, h->nibtip, h->twx64)
        if (h->nibent > h->nibtip)
          LH_DISPATCH_PLAN(dm.px, h->xplan, launch_xib_t, double, h, b, h->nibent - h->nibtip, h->twx64)
      }
    });
    ZJobs fj{};
    slot = 0;
    fj.j[0] = ZJob{(const double2*)h->zp, h->r1, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
    fj.j[1] = ZJob{(const double2*)h->zp, h->r3, h->ibu_planes, 0, h->nibu, slot}; slot += h->nibu;
    fj.j[2] = ZJob{(const double2*)h->zp, h->r2, h->ibv_planes, 0, h->nibv, slot};
    if (stages & kStageIBf)
    TIMED(h, LH_K_IBZ, nib * (spl + z64),
          launch_zfwd(h, fj, 3, h->nibu > h->nibv ? h->nibu : h->nibv, true));
  }
  // ---- column stage
  if (side)
    (void)hipStreamWaitEvent(h->st, h->ev_join, 0);
  else if (stages & kStageVelCol)
    TIMED(h, LH_K_MISC, 0.0,
          hipLaunchKernelGGL(k_prep, dim3(1), dim3(64), 0, h->st, (const ColJob*)h->coljobs + krk));
  {
    const ColJob* jk = h->coljobs + krk;
    const int nb = col_blocks(h);
    // in: rhs u, w, v, p, LU factors (8 B each); out: u, w, v, p
    const double by = spl * (2.0 * nuvw + 2.0 * dm.nyp) + 3.0 * 8.0 * dm.ncol * (double)dm.nyp;
    // in: u, p, Nu_old, Nu_new (rhs on canopy planes); eliminated rows out and back; out: u
    if (stages & kStageVelCol)
    TIMED(h, LH_K_VEL, spl * (7.0 * nuvw), {
      if (h->vel_u == kVelUSmall)
        hipLaunchKernelGGL(k_vel_rhs<kVelUSmall>, dim3(nb, 3), dim3(kMultiCols), 0, h->st, jk,
                           (const int*)h->colmap, 0);
      else
        hipLaunchKernelGGL(k_vel_rhs<LH_VEL_U>, dim3(nb, 3), dim3(kMultiCols), 0, h->st, jk,
                           (const int*)h->colmap, 0);
    });
    if (stages & kStagePres)
    TIMED(h, LH_K_PRES, by - 2.0 * nuvw * spl + 2.0 * nuvw * spl,
          hipLaunchKernelGGL(k_pres, dim3(nb), dim3(kMultiCols), 0, h->st, jk,
                             (const int*)h->colmap));
  }
}

static void issue_step(lh_handle* h, bool cfl_mode) {
  for (int k = 0; k < 3; ++k) issue_substep(h, k, cfl_mode);
  if (h->flag_ctpress == 1)
    TIMED(h, LH_K_MISC, 0.0,
          hipLaunchKernelGGL(k_flowrate, dim3(1), dim3(256), 0, h->st, (const ColJob*)h->coljobs));
}

// ---------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------
extern "C" {

const char* lh_last_error(const lh_handle* h) { return h ? h->err.c_str() : "null handle"; }

int lh_create(const lh_config* c, lh_handle** out) {
  if (!c || !out) return LH_ERR_ARG;
  *out = nullptr;
  lh_handle* h = new lh_handle();
  auto bad = [&](const std::string& m) { h->err = m; *out = h; return LH_ERR_ARG; };
  if (c->nx <= 0 || c->nz <= 0 || c->nx % 2 || c->nz % 2) return bad("nx, nz must be positive and even");
  if (c->px != 3 * c->nx / 2 || c->pz != 3 * c->nz / 2 || c->nx % 4 || c->nz % 4)
    return bad("px = 3*nx/2 and pz = 3*nz/2 are required (nx, nz multiples of 4)");
  if (!supported_len(c->px) || !supported_len(c->pz)) return bad("unsupported collocation size");
  if (c->dsty < 1 || c->ny < 4) return bad("ny >= 4 and dsty >= 1 required");
  if (c->precision != LH_PREC_FP64 && c->precision != LH_PREC_MIXED) return bad("bad precision");
  if (!c->yu || !c->yv || !c->dthdyu || !c->dthdyv || !c->dyu2i || !c->dyv2i || !c->gridweighting ||
      !c->rk || !c->k1x || !c->k2x || !c->k1z || !c->k2z)
    return bad("null set-up array");
  if ((c->nlist_u > 0 && !c->list_u) || (c->nlist_v > 0 && !c->list_v)) return bad("null list");
  Dims& dm = h->dm;
  dm.nx = c->nx; dm.nz = c->nz; dm.px = c->px; dm.pz = c->pz;
  dm.mx = c->nx / 2 + 1; dm.ncol = dm.mx * c->nz;
  h->vel_u = dm.ncol <= kVelSmallCols ? kVelUSmall : LH_VEL_U;
  dm.ny = c->ny; dm.d = c->dsty;
  dm.nyu = c->ny + 2 * c->dsty + 1; dm.nyv = c->ny + 2 * c->dsty; dm.nyp = c->ny + 2 * c->dsty - 1;
  dm.dxi = c->px * c->alp / (2.0 * M_PI);
  dm.dzi = c->pz * c->bet / (2.0 * M_PI);
  dm.re = c->re; dm.ddthetavi = c->ddthetavi;
  h->prec = c->precision;
  h->flag_ctpress = c->flag_ctpress;
  memcpy(h->rk, c->rk, sizeof(h->rk));
  h->gw[0] = c->gridweighting[0]; h->gw[1] = c->gridweighting[1];
  h->tip_u[0] = c->tip_u[0]; h->tip_u[1] = c->tip_u[1];
  h->tip_v[0] = c->tip_v[0]; h->tip_v[1] = c->tip_v[1];
  h->ntilex = c->ntilex > 0 ? c->ntilex : 1; h->ntilez = c->ntilez > 0 ? c->ntilez : 1;
  h->band_v[0] = c->band_v[0]; h->band_v[1] = c->band_v[1];
  for (int t = 0; t < 2; ++t) {
    if (c->tip_u[t] < -dm.d + 1 || c->tip_u[t] > dm.ny + dm.d - 1) return bad("tip_u out of range");
    if (c->tip_v[t] < -dm.d + 1 || c->tip_v[t] > dm.ny + dm.d - 2) return bad("tip_v out of range");
  }
  *out = h;
  HIPCHK(h, hipGetDevice(&h->dev));
  HIPCHK(h, hipStreamCreate(&h->st));
  HIPCHK(h, hipStreamCreate(&h->side));
  HIPCHK(h, hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
  HIPCHK(h, hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
  {
    int dummy = 0;
    for (int v = 0; v < plans_of_len(c->pz); ++v) {
      LH_DISPATCH_PLAN(c->pz, v, attr_zasm_t, double, &dummy)
      LH_DISPATCH_PLAN(c->pz, v, attr_zasm_t, float, &dummy)
      LH_DISPATCH_PLAN(c->pz, v, attr_z_t, double, &dummy)
      LH_DISPATCH_PLAN(c->pz, v, attr_z_t, float, &dummy)
    }
    for (int v = 0; v < plans_of_len(c->px); ++v) {
      LH_DISPATCH_PLAN(c->px, v, attr_xraw_t, double, &dummy)
      LH_DISPATCH_PLAN(c->px, v, attr_xraw_t, float, &dummy)
    }
  }
  {
    int ncu = 0;
    HIPCHK(h, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->dev));
    if (ncu < 1) ncu = 1;
    // slots = CUs x resident blocks per CU of the kernel at its precision
    int bx = 1, bz = 1;
    if (h->prec == LH_PREC_MIXED) {
      LH_DISPATCH_PLAN(dm.px, h->xplan, xraw_occ_t, float, &bx)
      LH_DISPATCH_PLAN(dm.pz, h->zplan, zasm_occ_t, float, &bz)
    } else {
      LH_DISPATCH_PLAN(dm.px, h->xplan, xraw_occ_t, double, &bx)
      LH_DISPATCH_PLAN(dm.pz, h->zplan, zasm_occ_t, double, &bz)
    }
    if (bx < 1) bx = 1;
    if (bz < 1) bz = 1;
    if (use_xw(h)) {
      // one block of kXwWaves waves per CU; a chunk's start-up is three inverse transforms
      // against nine transforms per level
      const bool f64 = h->prec == LH_PREC_FP64;
      HIPCHK(h, f64 ? attr_xw_t<double>() : attr_xw_t<float>());
      HIPCHK(h, hipFuncSetAttribute((const void*)k_xpass_ib576<LH_XIBW_WAVES>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)xibw_lds_bytes()));
      const int wpb = f64 ? XwWaves<double>::value : XwWaves<float>::value;
      if (h->xraw_lc <= 0)
        h->xraw_lc = chunk_levels(dm.nyu - 2, (long)dm.px / 2, ncu * wpb, 4, 96, 1.0 / 3.0);
    }
    if (h->xraw_lc <= 0) h->xraw_lc = chunk_levels(dm.nyu - 2, dm.pz / 2, ncu * bx, 4, 64, 1.0);
    if (h->zasm_lc <= 0)
      h->zasm_lc = chunk_levels(dm.nyu - 2, 3L * ((dm.mx + LH_ZASM_COLS - 1) / LH_ZASM_COLS), ncu * bz, 2, 12, 0.25);
  }
  const size_t nu = (size_t)dm.nyu * dm.ncol, nv = (size_t)dm.nyv * dm.ncol,
               np = (size_t)dm.nyp * dm.ncol;
  const int nlev = dm.nyu - 2;
  HIPCHK(h, dalloc(&h->yu, dm.nyu)); HIPCHK(h, dalloc(&h->dthdyu, dm.nyu));
  HIPCHK(h, dalloc(&h->yv, dm.nyv)); HIPCHK(h, dalloc(&h->dthdyv, dm.nyv));
  HIPCHK(h, dalloc(&h->dyu2i, 3 * dm.nyu)); HIPCHK(h, dalloc(&h->dyv2i, 3 * dm.nyv));
  HIPCHK(h, dalloc(&h->k1x, dm.mx)); HIPCHK(h, dalloc(&h->k2x, dm.mx));
  HIPCHK(h, dalloc(&h->k1z, dm.nz)); HIPCHK(h, dalloc(&h->k2z, dm.nz));
  HIPCHK(h, dalloc(&h->scal, S_NSCAL));
  HIPCHK(h, hipMemcpy(h->yu, c->yu, dm.nyu * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->dthdyu, c->dthdyu, dm.nyu * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->yv, c->yv, dm.nyv * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->dthdyv, c->dthdyv, dm.nyv * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->dyu2i, c->dyu2i, 3 * dm.nyu * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->dyv2i, c->dyv2i, 3 * dm.nyv * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->k1x, c->k1x, dm.mx * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->k2x, c->k2x, dm.mx * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->k1z, c->k1z, dm.nz * 8, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->k2z, c->k2z, dm.nz * 8, hipMemcpyHostToDevice));
  {
    double s[S_NSCAL] = {0};
    s[S_MPGX] = c->mpgx;
    s[S_QXT] = c->qxt;
    HIPCHK(h, hipMemcpy(h->scal, s, sizeof(s), hipMemcpyHostToDevice));
  }
  // state and work
  HIPCHK(h, dalloc(&h->u1, nu)); HIPCHK(h, dalloc(&h->u3, nu)); HIPCHK(h, dalloc(&h->u2, nv));
  HIPCHK(h, dalloc(&h->p, np)); HIPCHK(h, dalloc(&h->psi, np));
  HIPCHK(h, dalloc(&h->nu1, nu)); HIPCHK(h, dalloc(&h->nu3, nu)); HIPCHK(h, dalloc(&h->nu2, nv));
  HIPCHK(h, dalloc(&h->nb1, nu)); HIPCHK(h, dalloc(&h->nb3, nu)); HIPCHK(h, dalloc(&h->nb2, nv));
  HIPCHK(h, dalloc(&h->r1, nu)); HIPCHK(h, dalloc(&h->r3, nu)); HIPCHK(h, dalloc(&h->r2, nv));
  HIPCHK(h, dalloc(&h->tipl, 6 * (size_t)dm.ncol));
  HIPCHK(h, dalloc(&h->dg3, np));
  HIPCHK(h, dalloc(&h->ayu, 3 * dm.nyu)); HIPCHK(h, dalloc(&h->ayv, 3 * dm.nyv));
  HIPCHK(h, dalloc(&h->ctwork, 4 * (dm.ny + 2)));
  const size_t zel = h->prec == LH_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  const size_t zslot = (size_t)dm.pz * dm.mx;
  HIPCHK(h, hipMalloc(&h->zu, zel * zslot * dm.nyu));
  HIPCHK(h, hipMalloc(&h->zw, zel * zslot * dm.nyu));
  HIPCHK(h, hipMalloc(&h->zv, zel * zslot * dm.nyv));
  // zp also serves as the FP64 scratch of the IB stage
  {
    size_t nzp = zel * zslot * 6 * (size_t)nlev;
    size_t nib = sizeof(double2) * zslot * (3 * (size_t)dm.nyu + 6);
    HIPCHK(h, hipMalloc(&h->zp, nzp > nib ? nzp : nib));
  }
  HIPCHK(h, hipMemset(h->u1, 0, nu * 16)); HIPCHK(h, hipMemset(h->u3, 0, nu * 16));
  HIPCHK(h, hipMemset(h->u2, 0, nv * 16)); HIPCHK(h, hipMemset(h->p, 0, np * 16));
  HIPCHK(h, hipMemset(h->nu1, 0, nu * 16)); HIPCHK(h, hipMemset(h->nu3, 0, nu * 16));
  HIPCHK(h, hipMemset(h->nu2, 0, nv * 16));
  HIPCHK(h, hipMemset(h->nb1, 0, nu * 16)); HIPCHK(h, hipMemset(h->nb3, 0, nu * 16));
  HIPCHK(h, hipMemset(h->nb2, 0, nv * 16));
  // twiddles exp(-2 pi i m/N)
  auto mk_tw = [&](int n, void** d64, void** dnl) -> int {
    std::vector<double2> t(n);
    std::vector<float2> tf(n);
    for (int m = 0; m < n; ++m) {
      const double a = 2.0 * M_PI * (double)m / (double)n;
      t[m].x = cos(a); t[m].y = -sin(a);
      tf[m].x = (float)t[m].x; tf[m].y = (float)t[m].y;
    }
    if (hipMalloc(d64, n * sizeof(double2)) != hipSuccess) return 1;
    if (hipMemcpy(*d64, t.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) return 1;
    if (h->prec == LH_PREC_FP64) {
      *dnl = *d64;
    } else {
      if (hipMalloc(dnl, n * sizeof(float2)) != hipSuccess) return 1;
      if (hipMemcpy(*dnl, tf.data(), n * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) return 1;
    }
    return 0;
  };
  if (mk_tw(dm.px, &h->twx64, &h->twx) || mk_tw(dm.pz, &h->twz64, &h->twz))
    return fail(h, LH_ERR_HIP, "twiddle upload failed");
  if (use_xw(h)) {
    // w576^{+n1 t} (n1 = 0..8, t = 0..63) and w64^{+j} (j = 0..63)
    std::vector<double2> t(xw::TWN);
    for (int n1 = 0; n1 < 9; ++n1)
      for (int l = 0; l < 64; ++l) {
        const double a = 2.0 * M_PI * (double)(n1 * l) / 576.0;
        t[xw::TW576 + n1 * 64 + l] = make_double2(cos(a), sin(a));
      }
    for (int j = 0; j < 64; ++j) {
      const double a = 2.0 * M_PI * (double)j / 64.0;
      t[xw::TW64 + j] = make_double2(cos(a), sin(a));
    }
    HIPCHK(h, dalloc(&h->twxw, t.size()));
    HIPCHK(h, hipMemcpy(h->twxw, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
    if (h->prec != LH_PREC_FP64) {
      // the same table rounded to FP32 for the mixed-precision transforms
      std::vector<float2> t32(t.size());
      for (size_t e = 0; e < t.size(); ++e) t32[e] = make_float2((float)t[e].x, (float)t[e].y);
      HIPCHK(h, dalloc(&h->twxw32, t32.size()));
      HIPCHK(h, hipMemcpy(h->twxw32, t32.data(), t32.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
  }
  // canopy masks: one per distinct footprint; planes -> mask index
  {
    const size_t pp = (size_t)dm.px * dm.pz;
    std::vector<std::vector<unsigned char>> pu(dm.nyu), pv(dm.nyv);
    for (int n = 0; n < c->nlist_u; ++n) {
      const int i = c->list_u[3 * n], k = c->list_u[3 * n + 1], j = c->list_u[3 * n + 2];
      if (i < 1 || i > dm.px || k < 1 || k > dm.pz || j < -dm.d + 1 || j > dm.ny + dm.d - 1)
        return fail(h, LH_ERR_ARG, "list_u point out of range");
      auto& m = pu[j + dm.d];
      if (m.empty()) m.assign(pp, 0);
      m[(size_t)(k - 1) * dm.px + (i - 1)] = 1;
    }
    for (int n = 0; n < c->nlist_v; ++n) {
      const int i = c->list_v[3 * n], k = c->list_v[3 * n + 1], j = c->list_v[3 * n + 2];
      if (i < 1 || i > dm.px || k < 1 || k > dm.pz || j < -dm.d + 1 || j > dm.ny + dm.d - 2)
        return fail(h, LH_ERR_ARG, "list_v point out of range");
      auto& m = pv[j + dm.d];
      if (m.empty()) m.assign(pp, 0);
      m[(size_t)(k - 1) * dm.px + (i - 1)] = 1;
    }
    std::vector<std::vector<unsigned char>> uniq;
    auto idx_of = [&](const std::vector<unsigned char>& m) {
      for (size_t q = 0; q < uniq.size(); ++q)
        if (uniq[q] == m) return (int)q;
      uniq.push_back(m);
      return (int)uniq.size() - 1;
    };
    h->mask_of_u.assign(dm.nyu, -1);
    h->mask_of_v.assign(dm.nyv, -1);
    for (int s = 0; s < dm.nyu; ++s)
      if (!pu[s].empty()) { h->mask_of_u[s] = idx_of(pu[s]); h->h_ibu.push_back(s); }
    for (int s = 0; s < dm.nyv; ++s)
      if (!pv[s].empty()) { h->mask_of_v[s] = idx_of(pv[s]); h->h_ibv.push_back(s); }
    h->nmask = (int)uniq.size();
    h->nibu = (int)h->h_ibu.size();
    h->nibv = (int)h->h_ibv.size();
    std::vector<IBEntry> ent;
    int slot = 0;
    const int su0 = 0, sw0 = h->nibu, sv0 = 2 * h->nibu, st0 = 2 * h->nibu + h->nibv;
    auto tip_slot = [&](int grid, int s) -> int {   // Lu_y slot (tipl rows 0..5) or -1
      if (grid == 0) { if (s == h->tip_u[0] + dm.d) return 0; if (s == h->tip_u[1] + dm.d) return 1; }
      if (grid == 2) { if (s == h->tip_u[0] + dm.d) return 2; if (s == h->tip_u[1] + dm.d) return 3; }
      if (grid == 1) { if (s == h->tip_v[0] + dm.d) return 4; if (s == h->tip_v[1] + dm.d) return 5; }
      return -1;
    };
    for (int q = 0; q < h->nibu; ++q) {
      const int s = h->h_ibu[q];
      int ts = tip_slot(0, s);
      ent.push_back(IBEntry{su0 + q, ts >= 0 ? st0 + ts : -1, h->mask_of_u[s], 0});
    }
    for (int q = 0; q < h->nibu; ++q) {
      const int s = h->h_ibu[q];
      int ts = tip_slot(2, s);
      ent.push_back(IBEntry{sw0 + q, ts >= 0 ? st0 + ts : -1, h->mask_of_u[s], 0});
    }
    for (int q = 0; q < h->nibv; ++q) {
      const int s = h->h_ibv[q];
      int ts = tip_slot(1, s);
      ent.push_back(IBEntry{sv0 + q, ts >= 0 ? st0 + ts : -1, h->mask_of_v[s], 0});
    }
    (void)slot;
    std::stable_partition(ent.begin(), ent.end(), [](const IBEntry& x) { return x.lslot >= 0; });
    h->nibent = (int)ent.size();
    h->nibtip = 0;
    for (const IBEntry& x : ent) h->nibtip += x.lslot >= 0 ? 1 : 0;
    if (h->nmask > 0) {
      HIPCHK(h, hipMalloc((void**)&h->masks, pp * h->nmask));
      for (int q = 0; q < h->nmask; ++q)
        HIPCHK(h, hipMemcpy(h->masks + pp * q, uniq[q].data(), pp, hipMemcpyHostToDevice));
    }
    if (h->nibent > 0) {
      HIPCHK(h, hipMalloc((void**)&h->ibent, sizeof(IBEntry) * h->nibent));
      HIPCHK(h, hipMemcpy(h->ibent, ent.data(), sizeof(IBEntry) * h->nibent, hipMemcpyHostToDevice));
    }
    if (h->nibu) {
      HIPCHK(h, hipMalloc((void**)&h->ibu_planes, sizeof(int) * h->nibu));
      HIPCHK(h, hipMemcpy(h->ibu_planes, h->h_ibu.data(), sizeof(int) * h->nibu, hipMemcpyHostToDevice));
    }
    if (h->nibv) {
      HIPCHK(h, hipMalloc((void**)&h->ibv_planes, sizeof(int) * h->nibv));
      HIPCHK(h, hipMemcpy(h->ibv_planes, h->h_ibv.data(), sizeof(int) * h->nibv, hipMemcpyHostToDevice));
    }
    std::vector<unsigned char> fu(dm.nyu, 0), fv(dm.nyv, 0);
    if (h->nibent > 0) {
      for (int s : h->h_ibu) fu[s] = 1;
      for (int s : h->h_ibv) fv[s] = 1;
    }
    HIPCHK(h, hipMalloc((void**)&h->ibflag_u, dm.nyu));
    HIPCHK(h, hipMalloc((void**)&h->ibflag_v, dm.nyv));
    HIPCHK(h, hipMemcpy(h->ibflag_u, fu.data(), dm.nyu, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->ibflag_v, fv.data(), dm.nyv, hipMemcpyHostToDevice));
  }
  // the column-stage jobs of the three substeps (every buffer they name exists now)
  {
    std::vector<ColJob> jobs;
    for (int k = 0; k < 3; ++k) jobs.push_back(make_job(h, k, 0));
    const std::vector<int> map(col_blocks(h), 0);
    HIPCHK(h, hipMalloc((void**)&h->coljobs, jobs.size() * sizeof(ColJob)));
    HIPCHK(h, hipMalloc((void**)&h->colmap, map.size() * sizeof(int)));
    HIPCHK(h, hipMemcpy(h->coljobs, jobs.data(), jobs.size() * sizeof(ColJob), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->colmap, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  HIPCHK(h, refresh_ib(h));
  // pressure LU factors, once (start.f90:628-630)
  {
    BUILDP a{h->dg3, h->dthdyu, h->dthdyv, h->yu, h->k2x, h->k2z};
    hipLaunchKernelGGL(k_buildP, dim3((dm.ncol + 255) / 256), dim3(256), 0, h->st, a, dm);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipStreamSynchronize(h->st));
  return LH_OK;
}

int lh_destroy(lh_handle* h) {
  if (!h) return LH_OK;
  // a batch keeps pointers to its members: it must be destroyed first
  if (h->in_batch > 0) return fail(h, LH_ERR_ARG, "lh_destroy: the handle belongs to a batch; destroy the batch first");
  if (h->st) hipStreamSynchronize(h->st);
  if (h->side) hipStreamSynchronize(h->side);
  for (auto& g : h->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto e : h->evpool) (void)hipEventDestroy(e);
  if (h->ib) {
    h->ib->release();
    delete h->ib;
    h->ib = nullptr;
  }
  void* ptrs[] = {h->yu, h->yv, h->dthdyu, h->dthdyv, h->dyu2i, h->dyv2i, h->k1x, h->k2x, h->k1z,
                  h->k2z, h->scal, h->u1, h->u2, h->u3, h->p, h->nu1, h->nu2, h->nu3, h->nb1, h->nb2, h->nb3, h->r1,
                  h->r2, h->r3, h->psi, h->tipl, h->zu, h->zw, h->zv, h->zp, h->dg3, h->ayu, h->ayv, h->ctwork, h->masks, h->ibu_planes,
                  h->ibv_planes, h->ibflag_u, h->ibflag_v, h->ibent, h->st_psum, h->st_pfold, h->st_acc,
                  h->cst_acc, h->sp_acc, h->fo1, h->fo2, h->fo3, h->twxw, h->twxw32, h->coljobs, h->colmap};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (h->twx && h->twx != h->twx64) hipFree(h->twx);
  if (h->twz && h->twz != h->twz64) hipFree(h->twz);
  if (h->twx64) hipFree(h->twx64);
  if (h->twz64) hipFree(h->twz64);
  if (h->st) hipStreamDestroy(h->st);
  if (h->side) hipStreamDestroy(h->side);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  delete h;
  return LH_OK;
}

// reference layout f(j, column) <-> device [plane][column]
static int upload_field(lh_handle* h, double2* dst, const double* src, int nj) {
  const int ncol = h->dm.ncol;
  std::vector<double2> t((size_t)nj * ncol);
  const double2* s = reinterpret_cast<const double2*>(src);
  for (int c = 0; c < ncol; ++c)
    for (int j = 0; j < nj; ++j) t[(size_t)j * ncol + c] = s[(size_t)c * nj + j];
  HIPCHK(h, hipMemcpy(dst, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
  return LH_OK;
}
static int download_field(lh_handle* h, double* dst, const double2* src, int nj) {
  const int ncol = h->dm.ncol;
  std::vector<double2> t((size_t)nj * ncol);
  HIPCHK(h, hipMemcpy(t.data(), src, t.size() * sizeof(double2), hipMemcpyDeviceToHost));
  double2* d = reinterpret_cast<double2*>(dst);
  for (int c = 0; c < ncol; ++c)
    for (int j = 0; j < nj; ++j) d[(size_t)c * nj + j] = t[(size_t)j * ncol + c];
  return LH_OK;
}

int lh_set_state(lh_handle* h, const double* u1, const double* u2, const double* u3, const double* p) {
  if (!h) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  int r;
  if (u1 && (r = upload_field(h, h->u1, u1, h->dm.nyu))) return r;
  if (u2 && (r = upload_field(h, h->u2, u2, h->dm.nyv))) return r;
  if (u3 && (r = upload_field(h, h->u3, u3, h->dm.nyu))) return r;
  if (p && (r = upload_field(h, h->p, p, h->dm.nyp))) return r;
  return LH_OK;
}

int lh_get_state(lh_handle* h, double* u1, double* u2, double* u3, double* p) {
  if (!h) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  int r;
  if (u1 && (r = download_field(h, u1, h->u1, h->dm.nyu))) return r;
  if (u2 && (r = download_field(h, u2, h->u2, h->dm.nyv))) return r;
  if (u3 && (r = download_field(h, u3, h->u3, h->dm.nyu))) return r;
  if (p && (r = download_field(h, p, h->p, h->dm.nyp))) return r;
  return LH_OK;
}

// Nu in the reference's allocation range jlim(1)+1 : jlim(2)-1 (littleharsh.f90:122-124)
static int xfer_nl(lh_handle* h, double2* dev, double* host, int nj, bool up) {
  const int ncol = h->dm.ncol;
  std::vector<double2> t((size_t)nj * ncol);
  if (!up) HIPCHK(h, hipMemcpy(t.data(), dev, t.size() * 16, hipMemcpyDeviceToHost));
  double2* hh = reinterpret_cast<double2*>(host);
  for (int c = 0; c < ncol; ++c)
    for (int r = 0; r < nj - 2; ++r) {
      if (up) t[(size_t)(r + 1) * ncol + c] = hh[(size_t)c * (nj - 2) + r];
      else hh[(size_t)c * (nj - 2) + r] = t[(size_t)(r + 1) * ncol + c];
    }
  if (up) HIPCHK(h, hipMemcpy(dev, t.data(), t.size() * 16, hipMemcpyHostToDevice));
  return LH_OK;
}

int lh_set_nonlinear(lh_handle* h, const double* n1, const double* n2, const double* n3) {
  if (!h) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  int r;
  // both buffers: the next substep reads whichever its parity selects
  for (int b = 0; b < 2; ++b) {
    if (n1 && (r = xfer_nl(h, b ? h->nb1 : h->nu1, (double*)n1, h->dm.nyu, true))) return r;
    if (n2 && (r = xfer_nl(h, b ? h->nb2 : h->nu2, (double*)n2, h->dm.nyv, true))) return r;
    if (n3 && (r = xfer_nl(h, b ? h->nb3 : h->nu3, (double*)n3, h->dm.nyu, true))) return r;
  }
  return LH_OK;
}

int lh_get_nonlinear(lh_handle* h, double* n1, double* n2, double* n3) {
  if (!h) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  int r;
  const bool b = h->nu_last == 1;
  if (n1 && (r = xfer_nl(h, b ? h->nb1 : h->nu1, n1, h->dm.nyu, false))) return r;
  if (n2 && (r = xfer_nl(h, b ? h->nb2 : h->nu2, n2, h->dm.nyv, false))) return r;
  if (n3 && (r = xfer_nl(h, b ? h->nb3 : h->nu3, n3, h->dm.nyu, false))) return r;
  return LH_OK;
}

// one stage of substep kRK on the handle's stream, synchronous at return
static int run_stage(lh_handle* h, int krk, double dt, int stage) {
  if (!h) return LH_ERR_ARG;
  if (krk < 1 || krk > 3) return fail(h, LH_ERR_ARG, "kRK must be 1, 2 or 3");
  if (stage != kStageNL && !(dt > 0.0)) return fail(h, LH_ERR_ARG, "dt must be positive");
  if (dt > 0.0) hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_DT, dt);
  issue_substep(h, krk - 1, false, stage);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  tflush(h);
  return LH_OK;
}
int lh_nonlinear(lh_handle* h, int krk) { return run_stage(h, krk, 0.0, kStageNL); }
int lh_solve_u(lh_handle* h, int krk, double dt) { return run_stage(h, krk, dt, kStageVel); }
int lh_solve_p(lh_handle* h, int krk, double dt) { return run_stage(h, krk, dt, kStagePres); }

int lh_substep(lh_handle* h, int krk, double dt) {
  if (!h) return LH_ERR_ARG;
  if (krk < 1 || krk > 3) return fail(h, LH_ERR_ARG, "kRK must be 1, 2 or 3");
  if (dt > 0.0) hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_DT, dt);
  issue_substep(h, krk - 1, false);
  // meanflow_ctP follows the last substep of a step (littleharsh.f90:193-197)
  if (krk == 3 && h->flag_ctpress == 1)
    hipLaunchKernelGGL(k_flowrate, dim3(1), dim3(256), 0, h->st, (const ColJob*)h->coljobs);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  tflush(h);
  if (krk == 1) {   // the first substep takes the CFL maximum of the state
    double s[S_NSCAL];
    return numeric_check(h, false, s);
  }
  return LH_OK;
}

static int run_graph(lh_handle* h, int mode) {
  if (h->timing) {   // per-launch events cannot live in a graph: issue directly
    issue_step(h, mode == 1);
    return LH_OK;
  }
  if (!h->gexec[mode]) {
    hipGraph_t g;
    HIPCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    issue_step(h, mode == 1);
    HIPCHK(h, hipStreamEndCapture(h->st, &g));
    HIPCHK(h, hipGraphInstantiate(&h->gexec[mode], g, nullptr, nullptr, 0));
    HIPCHK(h, hipGraphDestroy(g));
  }
  const int ti = tbeg(h, LH_K_STEP, 0.0, true);
  HIPCHK(h, hipGraphLaunch(h->gexec[mode], h->st));
  tend(h, ti);
  return LH_OK;
}

int lh_step(lh_handle* h, double dt) {
  if (!h) return LH_ERR_ARG;
  if (!(dt > 0.0)) return fail(h, LH_ERR_ARG, "dt must be positive");
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_DT, dt);
  int r = run_graph(h, 0);
  if (r) return r;
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  tflush(h);
  double s[S_NSCAL];
  return numeric_check(h, false, s);
}

int lh_step_cfl(lh_handle* h, double cfl, double* dt_used) {
  if (!h) return LH_ERR_ARG;
  if (!(cfl > 0.0)) return fail(h, LH_ERR_ARG, "cfl must be positive");
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_CFL, cfl);
  int r = run_graph(h, 1);
  if (r) return r;
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  tflush(h);
  double s[S_NSCAL];
  r = numeric_check(h, true, s);
  if (dt_used) *dt_used = s[S_DT];
  return r;
}

int lh_get_scalars(lh_handle* h, double* mpgx, double* qx, double* dt, double* cflmax) {
  if (!h) return LH_ERR_ARG;
  double s[S_NSCAL];
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(s, h->scal, sizeof(s), hipMemcpyDeviceToHost));
  if (mpgx) *mpgx = s[S_MPGX];
  if (qx) *qx = s[S_QX];
  if (dt) *dt = s[S_DT];
  if (cflmax) *cflmax = s[S_CFLMAX];
  return LH_OK;
}

int lh_max_div(lh_handle* h, double* err) {
  if (!h || !err) return LH_ERR_ARG;
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_MAXDIV, 0.0);
  hipLaunchKernelGGL(k_maxdiv, dim3((h->dm.ncol + 255) / 256), dim3(256), 0, h->st, h->u1, h->u2,
                     h->u3, h->dthdyu, h->k1x, h->k1z, h->scal, h->dm);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, hipMemcpy(err, h->scal + S_MAXDIV, 8, hipMemcpyDeviceToHost));
  return LH_OK;
}

int lh_cfl_dt(lh_handle* h, double cfl, double* dt) {
  if (!h || !dt) return LH_ERR_ARG;
  if (!(cfl > 0.0)) return fail(h, LH_ERR_ARG, "cfl must be positive");
  const Dims& dm = h->dm;
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_CFL, cfl);
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_CFLMAX, 0.0);
  ZJobs jobs{};
  jobs.j[0] = ZJob{h->u1, h->zu, nullptr, 0, dm.nyu, 0};
  jobs.j[1] = ZJob{h->u3, h->zw, nullptr, 0, dm.nyu, 0};
  jobs.j[2] = ZJob{h->u2, h->zv, nullptr, 0, dm.nyv, 0};
  launch_zinv(h, jobs, 3, dm.nyu, false);
  XRAW a{h->zu, h->zw, h->zv, h->zp, h->dthdyu, h->scal, 1};
  if (use_xw(h)) {
    launch_xw(h, 1);
  } else if (h->prec == LH_PREC_FP64) {
    LH_DISPATCH_PLAN(dm.px, h->xplan, launch_xraw_t, double, h, a, h->twx)
  } else {
    LH_DISPATCH_PLAN(dm.px, h->xplan, launch_xraw_t, float, h, a, h->twx)
  }
  hipLaunchKernelGGL(k_cfl_dt, dim3(1), dim3(1), 0, h->st, h->scal);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  double s[S_NSCAL];
  HIPCHK(h, hipMemcpy(s, h->scal, sizeof(s), hipMemcpyDeviceToHost));
  if (!(s[S_CFLMAX] > 0.0) || !std::isfinite(s[S_DT]))
    return fail(h, LH_ERR_NUMERIC, "CFL maximum is zero or not finite");
  *dt = s[S_DT];
  return LH_OK;
}

int lh_timing_enable(lh_handle* h, int enable) {
  if (!h) return LH_ERR_ARG;
  if (enable < 0 || enable > 2) return fail(h, LH_ERR_ARG, "lh_timing_enable: mode 0, 1 or 2");
  HIPCHK(h, hipStreamSynchronize(h->st));
  tflush(h);
  h->timing = enable == 1;
  h->timing_steps = enable == 2;
  return LH_OK;
}

int lh_timing_reset(lh_handle* h) {
  if (!h) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  tflush(h);
  for (int k = 0; k < LH_K_COUNT; ++k) { h->tms[k] = 0; h->tcnt[k] = 0; h->tbytes[k] = 0; }
  return LH_OK;
}

int lh_timing_read(lh_handle* h, int kernel, double* total_ms, long long* launches,
                   double* bytes_per_launch) {
  if (!h || kernel < 0 || kernel >= LH_K_COUNT) return LH_ERR_ARG;
  if (total_ms) *total_ms = h->tms[kernel];
  if (launches) *launches = h->tcnt[kernel];
  if (bytes_per_launch) *bytes_per_launch = h->tcnt[kernel] ? h->tbytes[kernel] / h->tcnt[kernel] : 0.0;
  return LH_OK;
}

// ---- launch parameters and the autotuner
static void drop_graphs(lh_handle* h) {
  for (auto& g : h->gexec)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  h->gen += 1;
}

int lh_set_launch(lh_handle* h, int xraw_lc, int zasm_lc, int col_bs) {
  if (!h) return LH_ERR_ARG;
  if (xraw_lc < 1 || zasm_lc < 1 || col_bs < 32 || col_bs > 256 || col_bs % 32)
    return fail(h, LH_ERR_ARG, "lh_set_launch: xraw_lc, zasm_lc >= 1, col_bs in 32..256 (multiple of 32)");
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->xraw_lc = xraw_lc;
  h->zasm_lc = zasm_lc;
  h->col_bs = col_bs;
  drop_graphs(h);
  return LH_OK;
}

int lh_set_forcing(lh_handle* h, const double* f1, const double* f2, const double* f3) {
  if (!h) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  const bool on = f1 || f2 || f3;
  if (on != h->forcing) drop_graphs(h);   // the captured substeps gain or lose the add
  h->forcing = on;
  if (!on) return LH_OK;
  const size_t nu = (size_t)h->dm.nyu * h->dm.ncol, nv = (size_t)h->dm.nyv * h->dm.ncol;
  if (!h->fo1) {
    HIPCHK(h, dalloc(&h->fo1, nu)); HIPCHK(h, dalloc(&h->fo3, nu)); HIPCHK(h, dalloc(&h->fo2, nv));
  }
  // a null field is zero; the ghost planes of every field stay zero
  std::vector<double> zero(2 * nu, 0.0);
  int r;
  if ((r = xfer_nl(h, h->fo1, (double*)(f1 ? f1 : zero.data()), h->dm.nyu, true))) return r;
  if ((r = xfer_nl(h, h->fo2, (double*)(f2 ? f2 : zero.data()), h->dm.nyv, true))) return r;
  if ((r = xfer_nl(h, h->fo3, (double*)(f3 ? f3 : zero.data()), h->dm.nyu, true))) return r;
  return LH_OK;
}

}  // extern "C"

int lh_div_view(lh_handle* h, LhDivView* v) {
  if (!h || !v) return LH_ERR_ARG;
  *v = LhDivView{h->u1, h->u2, h->u3, h->dthdyu, h->k1x, h->k1z, h->dm.ddthetavi,
                 h->dm.ncol, h->dm.mx, h->dm.nyp, h->dev, h->st};
  return LH_OK;
}
int lh_fail(lh_handle* h, int code, const char* msg) { return fail(h, code, msg); }

int lh_solver_view(lh_handle* h, LhSolverView* v) {
  if (!h || !v) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  const size_t nu = (size_t)h->dm.nyu * h->dm.ncol, nv = (size_t)h->dm.nyv * h->dm.ncol;
  if (!h->fo1) {
    HIPCHK(h, dalloc(&h->fo1, nu)); HIPCHK(h, dalloc(&h->fo3, nu)); HIPCHK(h, dalloc(&h->fo2, nv));
    HIPCHK(h, hipMemset(h->fo1, 0, nu * 16)); HIPCHK(h, hipMemset(h->fo3, 0, nu * 16));
    HIPCHK(h, hipMemset(h->fo2, 0, nv * 16));
  }
  if (!h->forcing) drop_graphs(h);
  h->forcing = true;
  *v = LhSolverView{h->u1, h->u2, h->u3, h->fo1, h->fo2, h->fo3, h->k1x, h->k1z,
                    h->dm.ncol, h->dm.nyu, h->dm.nyv, h->dm.nz, h->dm.mx, h->st};
  return LH_OK;
}

extern "C" {

int lh_get_forcing(lh_handle* h, double* f1, double* f2, double* f3) {
  if (!h) return LH_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (!h->fo1) return fail(h, LH_ERR_ARG, "lh_get_forcing: no forcing has been set");
  int r;
  if (f1 && (r = xfer_nl(h, h->fo1, f1, h->dm.nyu, false))) return r;
  if (f2 && (r = xfer_nl(h, h->fo2, f2, h->dm.nyv, false))) return r;
  if (f3 && (r = xfer_nl(h, h->fo3, f3, h->dm.nyu, false))) return r;
  return LH_OK;
}

int lh_get_launch(lh_handle* h, int* xraw_lc, int* zasm_lc, int* col_bs) {
  if (!h) return LH_ERR_ARG;
  if (xraw_lc) *xraw_lc = h->xraw_lc;
  if (zasm_lc) *zasm_lc = h->zasm_lc;
  if (col_bs) *col_bs = h->col_bs;
  return LH_OK;
}

int lh_set_plans(lh_handle* h, int xplan, int zplan) {
  if (!h) return LH_ERR_ARG;
  if (xplan < 0 || xplan >= plans_of_len(h->dm.px) || zplan < 0 || zplan >= plans_of_len(h->dm.pz))
    return fail(h, LH_ERR_ARG, "lh_set_plans: plan out of range for this length (lh_get_plans gives the counts)");
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->xplan = xplan;
  h->zplan = zplan;
  HIPCHK(h, refresh_ib(h));
  drop_graphs(h);
  return LH_OK;
}

int lh_get_plans(lh_handle* h, int* xplan, int* zplan, int* xcount, int* zcount) {
  if (!h) return LH_ERR_ARG;
  if (xplan) *xplan = h->xplan;
  if (zplan) *zplan = h->zplan;
  if (xcount) *xcount = plans_of_len(h->dm.px);
  if (zcount) *zcount = plans_of_len(h->dm.pz);
  return LH_OK;
}

// Search over the FFT plans (every x/z pair), then coordinate search over the launch
// parameters: every candidate runs `steps` CFL steps (after one warm-up step) timed with
// events on the handle's stream; the state, the nonlinear terms and the device scalars
// are saved first and restored afterwards.
int lh_autotune(lh_handle* h, int steps) {
  if (!h) return LH_ERR_ARG;
  if (steps < 1) steps = 2;
  if (h->timing) return fail(h, LH_ERR_ARG, "lh_autotune: disable timing first");
  const Dims& dm = h->dm;
  HIPCHK(h, hipStreamSynchronize(h->st));
  const size_t nu = (size_t)dm.nyu * dm.ncol, nv = (size_t)dm.nyv * dm.ncol,
               np = (size_t)dm.nyp * dm.ncol;
  double2* bufs[10] = {h->u1, h->u2, h->u3, h->p, h->nu1, h->nu2, h->nu3, h->nb1, h->nb2, h->nb3};
  const size_t lens[10] = {nu, nv, nu, np, nu, nv, nu, nu, nv, nu};
  size_t tot = 0;
  for (size_t l : lens) tot += l;
  double2* save = nullptr;
  HIPCHK(h, dalloc(&save, tot + S_NSCAL));
  size_t off = 0;
  for (int q = 0; q < 10; ++q) {
    HIPCHK(h, hipMemcpyAsync(save + off, bufs[q], lens[q] * 16, hipMemcpyDeviceToDevice, h->st));
    off += lens[q];
  }
  HIPCHK(h, hipMemcpyAsync(save + off, h->scal, S_NSCAL * 8, hipMemcpyDeviceToDevice, h->st));
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->st, h->scal, (int)S_CFL, 0.5);
  auto trial = [&](float* ms) -> int {
    issue_step(h, true);
    HIPCHK(h, hipEventRecord(e0, h->st));
    for (int s = 0; s < steps; ++s) issue_step(h, true);
    HIPCHK(h, hipEventRecord(e1, h->st));
    HIPCHK(h, hipEventSynchronize(e1));
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventElapsedTime(ms, e0, e1));
    return LH_OK;
  };
  int r = LH_OK;
  // every pair of FFT plans (at most 2 x 2), the current pair first
  {
    const int nxp = plans_of_len(dm.px), nzp = plans_of_len(dm.pz);
    const int x00 = h->xplan, z00 = h->zplan;
    int bx = x00, bz = z00;
    float best_ms = 1e30f;
    for (int c = 0; c < nxp * nzp && r == LH_OK; ++c) {
      h->xplan = (x00 + c / nzp) % nxp;
      h->zplan = (z00 + c % nzp) % nzp;
      if (refresh_ib(h) != hipSuccess) { r = fail(h, LH_ERR_HIP, "lh_autotune: job table"); break; }
      float ms = 0.f;
      r = trial(&ms);
      if (ms < best_ms) { best_ms = ms; bx = h->xplan; bz = h->zplan; }
    }
    h->xplan = bx;
    h->zplan = bz;
    if (refresh_ib(h) != hipSuccess && r == LH_OK) r = fail(h, LH_ERR_HIP, "lh_autotune: job table");
  }
  // candidates of each knob, the current value first (the x pass default comes from the
  // dispatch-round model, its neighbours bracket it)
  const int x0 = h->xraw_lc, z0 = h->zasm_lc;
  const std::vector<int> cands[3] = {{h->col_bs, 64, 128, 256},
                                     {x0, x0 / 2 > 0 ? x0 / 2 : 1, 2 * x0, 16},
                                     {z0, 1, 2, 4, 8}};
  for (int pass = 0; pass < 3 && r == LH_OK; ++pass) {
    if (pass == 0 && (h->nibent == 0 || ib_jobs(h))) continue;   // no by-value canopy rhs launch
    int* knob = pass == 0 ? &h->col_bs : (pass == 1 ? &h->xraw_lc : &h->zasm_lc);
    int best = *knob;
    float best_ms = 1e30f;
    for (size_t c = 0; c < cands[pass].size() && r == LH_OK; ++c) {
      const int v = cands[pass][c];
      if (c > 0 && std::find(cands[pass].begin(), cands[pass].begin() + c, v) != cands[pass].begin() + c)
        continue;                               // already tried
      *knob = v;
      float ms = 0.f;
      r = trial(&ms);
      if (ms < best_ms) { best_ms = ms; best = v; }
    }
    *knob = best;
  }
  off = 0;
  for (int q = 0; q < 10; ++q) {
    (void)hipMemcpyAsync(bufs[q], save + off, lens[q] * 16, hipMemcpyDeviceToDevice, h->st);
    off += lens[q];
  }
  (void)hipMemcpyAsync(h->scal, save + off, S_NSCAL * 8, hipMemcpyDeviceToDevice, h->st);
  (void)hipStreamSynchronize(h->st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(save);
  drop_graphs(h);
  return r;
}

// ---- many independent channels advanced by one graph launch per step
struct lh_batch {
  std::vector<lh_handle*> hs;
  std::vector<unsigned> gens;   // member generations the graph was captured at
  hipStream_t master = nullptr;
  hipGraphExec_t g = nullptr;
  double cfl = -1.0;
  std::vector<hipEvent_t> ev;   // 3 forks (one per substep), then 3 x n joins
  ColJob* jobs = nullptr;       // [3][n] column-stage jobs of the captured graph
  int* colmap = nullptr;        // block -> member of the column launches
  int nblk = 0;
  IbSets ib;                    // canopy-stage job tables of the members with small grids
  double* dscal = nullptr;      // [n][S_NSCAL] the members' scalars after a step
  bool joined = false;          // the handles count this batch (in_batch)
  std::string err;
};

int lh_batch_create(int n, lh_handle* const* hs, lh_batch** out) {
  if (!out || n < 1 || !hs) return LH_ERR_ARG;
  *out = nullptr;
  lh_batch* b = new lh_batch();
  for (int i = 0; i < n; ++i) {
    if (!hs[i]) { delete b; return LH_ERR_ARG; }
    // a handle listed twice would be stepped twice per call on its one stream
    if (std::find(b->hs.begin(), b->hs.end(), hs[i]) != b->hs.end()) { delete b; return LH_ERR_ARG; }
    b->hs.push_back(hs[i]);
  }
  if (hipStreamCreate(&b->master) != hipSuccess) { delete b; return LH_ERR_HIP; }
  b->ev.assign(3 * (n + 1), nullptr);
  for (auto& e : b->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      lh_batch_destroy(b);
      return LH_ERR_HIP;
    }
  for (auto* h : b->hs) {
    h->in_batch += 1;
    drop_graphs(h);                                 // captured with the side branch
  }
  b->joined = true;
  *out = b;
  return LH_OK;
}

const char* lh_batch_last_error(const lh_batch* b) { return b ? b->err.c_str() : "null batch"; }

int lh_batch_step_cfl(lh_batch* b, double cfl) {
  if (!b) return LH_ERR_ARG;
  if (!(cfl > 0.0)) { b->err = "cfl must be positive"; return LH_ERR_ARG; }
  for (auto* h : b->hs)
    if (h->timing) { b->err = "disable timing on every handle of a batch"; return LH_ERR_ARG; }
  bool stale = b->gens.size() != b->hs.size();
  for (size_t i = 0; i < b->hs.size() && !stale; ++i) stale = b->gens[i] != b->hs[i]->gen;
  if (!b->g || b->cfl != cfl || stale) {
    if (b->g) { (void)hipGraphExecDestroy(b->g); b->g = nullptr; }
    hipGraph_t g;
    auto chk = [&](hipError_t e, const char* what) {
      if (e != hipSuccess) b->err = std::string(what) + ": " + hipGetErrorString(e);
      return e == hipSuccess;
    };
    // the column-stage jobs of the three substeps, before the capture (they are copied to
    // the device, where the graph's launches read them)
    const int n = (int)b->hs.size();
    std::vector<ColJob> jobs(3 * (size_t)n);
    // blocks of the members with the small-grid sweep (kVelUSmall rows per step) first
    std::vector<int> colmap;
    int nblk = 0, nsmall = 0;
    for (int pass = 0; pass < 2; ++pass)
      for (int i = 0; i < n; ++i) {
        lh_handle* h = b->hs[i];
        if ((h->vel_u == kVelUSmall) != (pass == 0)) continue;
        for (int k = 0; k < 3; ++k) jobs[(size_t)k * n + i] = make_job(h, k, nblk);
        const int nb = col_blocks(h);
        colmap.insert(colmap.end(), nb, i);
        nblk += nb;
        if (pass == 0) nsmall = nblk;
      }
    if (b->jobs) { (void)hipFree(b->jobs); b->jobs = nullptr; }
    if (b->colmap) { (void)hipFree(b->colmap); b->colmap = nullptr; }
    if (!chk(hipMalloc((void**)&b->jobs, jobs.size() * sizeof(ColJob)), "job table") ||
        !chk(hipMalloc((void**)&b->colmap, colmap.size() * sizeof(int)), "block map") ||
        !chk(hipMemcpy(b->jobs, jobs.data(), jobs.size() * sizeof(ColJob), hipMemcpyHostToDevice), "job upload") ||
        !chk(hipMemcpy(b->colmap, colmap.data(), colmap.size() * sizeof(int), hipMemcpyHostToDevice), "map upload"))
      return LH_ERR_HIP;
    b->nblk = nblk;
    if (!chk(b->ib.build(b->hs), "canopy job tables")) return LH_ERR_HIP;
    if (!b->dscal && !chk(hipMalloc((void**)&b->dscal, (size_t)n * S_NSCAL * sizeof(double)), "scalars"))
      return LH_ERR_HIP;
    // Per substep: each member's nonlinear stage on its own stream (parallel branches), with
    // the canopy stage of the members with large grids; joined on the master stream, the
    // canopy stage of the members with small grids in one launch per kernel and plan, then the
    // column stage of all members in one launch each of k_prep, the velocity sweeps and the
    // pressure step.
    if (!chk(hipStreamBeginCapture(b->master, hipStreamCaptureModeGlobal), "capture")) return LH_ERR_HIP;
    // Within a substep the members' launches are issued launch by launch (each launch for
    // every member, then the next), so that consecutive nodes of the graph are independent.
    static const int phases[] = {kStageNLz, kStageNLx, kStageNLa, kStageIBr, kStageIBz, kStageIBx, kStageIBf};
    bool ok = true;
    const int sb = (n + 63) / 64;
    for (int k = 0; k < 3 && ok; ++k) {
      // the CFL number and the cleared maximum of every member, before the fork
      if (k == 0) hipLaunchKernelGGL(k_jobs_cfl_begin, dim3(sb), dim3(64), 0, b->master, (const ColJob*)b->jobs, n, cfl);
      ok = chk(hipEventRecord(b->ev[k], b->master), "fork");
      for (int i = 0; i < n && ok; ++i)
        ok = chk(hipStreamWaitEvent(b->hs[i]->st, b->ev[k], 0), "fork wait");
      // members with small grids leave the time step to the launch after the join (nothing
      // of theirs before it needs dt); the others form it on their branch, before their
      // canopy stage
      for (int ph : phases)
        for (int i = 0; i < n && ok; ++i) {
          lh_handle* h = b->hs[i];
          if (!((ph & kStageIB) && ib_jobs(h))) issue_substep(h, k, true, ph | (small_channel(h) ? kCallerCfl : 0));
        }
      for (int i = 0; i < n && ok; ++i) {
        lh_handle* h = b->hs[i];
        hipEvent_t join = b->ev[3 + (size_t)k * n + i];
        ok = chk(hipEventRecord(join, h->st), "join");
        ok = ok && chk(hipStreamWaitEvent(b->master, join, 0), "join wait");
      }
      if (k == 0) hipLaunchKernelGGL(k_jobs_cfl_dt, dim3(sb), dim3(64), 0, b->master, (const ColJob*)b->jobs, n);
      b->ib.launch(k, b->master);
      const ColJob* jk = b->jobs + (size_t)k * n;
      hipLaunchKernelGGL(k_prep, dim3(n), dim3(64), 0, b->master, jk);
      if (nsmall > 0)
        hipLaunchKernelGGL(k_vel_rhs<kVelUSmall>, dim3(nsmall, 3), dim3(kMultiCols), 0, b->master, jk,
                           (const int*)b->colmap, 0);
      if (nblk > nsmall)
        hipLaunchKernelGGL(k_vel_rhs<LH_VEL_U>, dim3(nblk - nsmall, 3), dim3(kMultiCols), 0, b->master,
                           jk, (const int*)b->colmap, nsmall);
      hipLaunchKernelGGL(k_pres, dim3(nblk), dim3(kMultiCols), 0, b->master, jk,
                         (const int*)b->colmap);
    }
    hipLaunchKernelGGL(k_flowrate, dim3(n), dim3(256), 0, b->master, (const ColJob*)b->jobs);
    hipLaunchKernelGGL(k_jobs_scalars, dim3(sb), dim3(64), 0, b->master, (const ColJob*)b->jobs, n, b->dscal);
    const bool ended = chk(hipStreamEndCapture(b->master, &g), "end capture");
    if (!ok || !ended) return LH_ERR_HIP;
    if (!chk(hipGraphInstantiate(&b->g, g, nullptr, nullptr, 0), "instantiate")) {
      (void)hipGraphDestroy(g);
      return LH_ERR_HIP;
    }
    (void)hipGraphDestroy(g);
    b->cfl = cfl;
    b->gens.clear();
    for (auto* h : b->hs) b->gens.push_back(h->gen);
  }
  if (hipGraphLaunch(b->g, b->master) != hipSuccess || hipStreamSynchronize(b->master) != hipSuccess) {
    b->err = "batch graph launch failed";
    return LH_ERR_HIP;
  }
  // every member's scalars in one copy
  std::vector<double> sc(b->hs.size() * S_NSCAL);
  if (hipMemcpy(sc.data(), b->dscal, sc.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
    b->err = "batch scalars: copy failed";
    return LH_ERR_HIP;
  }
  for (size_t i = 0; i < b->hs.size(); ++i) {
    const int r = check_scalars(b->hs[i], true, sc.data() + i * S_NSCAL);
    if (r != LH_OK) {
      b->err = "instance " + std::to_string(i) + ": " + b->hs[i]->err;
      return r;
    }
  }
  return LH_OK;
}

int lh_batch_destroy(lh_batch* b) {
  if (!b) return LH_OK;
  if (b->master) (void)hipStreamSynchronize(b->master);
  if (b->joined)
    for (auto* h : b->hs) {
      h->in_batch -= 1;
      drop_graphs(h);                               // captured with k_prep in line
    }
  if (b->g) (void)hipGraphExecDestroy(b->g);
  if (b->jobs) (void)hipFree(b->jobs);
  if (b->colmap) (void)hipFree(b->colmap);
  if (b->dscal) (void)hipFree(b->dscal);
  b->ib.release();
  for (auto e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->master) (void)hipStreamDestroy(b->master);
  delete b;
  return LH_OK;
}

// ---- statistics and spectra
static const int kStatsGrid[13] = {0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1};

struct StatsGeom {
  int dnx, dnz, nb[3], toplo;
  long long qoff[13], clen;
};

static int stats_geom(lh_handle* h, StatsGeom& g) {
  const Dims& dm = h->dm;
  if (dm.px % h->ntilex || dm.pz % h->ntilez)
    return fail(h, LH_ERR_ARG, "statistics: px, pz must be multiples of ntilex, ntilez");
  g.dnx = dm.px / h->ntilex;
  g.dnz = dm.pz / h->ntilez;
  const int n31 = h->band_v[0], n32 = h->band_v[1];
  if (!(n31 >= -dm.d && n31 < n32 && n32 <= dm.ny + dm.d - 2))
    return fail(h, LH_ERR_ARG, "statistics: band limits band_v must satisfy -dsty <= N(3,1) < N(3,2)");
  g.nb[0] = n31 + dm.d + 2;   // u planes N(4,0) .. N(4,1)
  g.nb[1] = n31 + dm.d + 1;   // v planes N(3,0) .. N(3,1)
  g.nb[2] = n31 + dm.d;       // p planes N(4,0)+1 .. N(4,1)-1
  g.toplo = n32 + dm.d + 1;
  long long a = 0;
  for (int q = 0; q < 13; ++q) {
    g.qoff[q] = a;
    a += (long long)g.dnx * g.dnz * g.nb[kStatsGrid[q]];
  }
  g.clen = a;
  return LH_OK;
}

int lh_stats_accumulate(lh_handle* h) {
  if (!h) return LH_ERR_ARG;
  const Dims& dm = h->dm;
  StatsGeom g;
  int r = stats_geom(h, g);
  if (r) return r;
  if (!h->st_acc) {
    const size_t nmax = dm.nyu;
    HIPCHK(h, dalloc(&h->st_psum, nmax * g.dnz * 6));
    HIPCHK(h, dalloc(&h->st_pfold, nmax * g.dnz * 6 * g.dnx));
    HIPCHK(h, dalloc(&h->st_acc, 13 * nmax));
    HIPCHK(h, dalloc(&h->cst_acc, (size_t)g.clen));
    HIPCHK(h, hipMemset(h->st_acc, 0, 13 * nmax * sizeof(double)));
    HIPCHK(h, hipMemset(h->cst_acc, 0, (size_t)g.clen * sizeof(double)));
  }
  // spectral v_itp -> r1, w_itp -> r2, wx -> r3 (the rhs buffers are free between substeps)
  hipLaunchKernelGGL(k_stats_prep, dim3((dm.ncol + 255) / 256, dm.nyu), dim3(256), 0, h->st, h->u2, h->u3,
                     h->dthdyv, h->k1z, h->r1, h->r2, h->r3, dm);
  static const int qmap[3][6] = {{0, 1, 4, 5, 8, 9}, {2, 3, 10, 11, 12, 0}, {6, 7, 0, 0, 0, 0}};
  static const double tsg[3][6] = {{1, 1, 1, 1, -1, 1}, {-1, 1, -1, 1, 1, 1}, {1, 1, 1, 1, 1, 1}};
  static const int nloc[3] = {6, 5, 2};
  for (int type = 0; type < 3; ++type) {
    // FP64 Z slots in zp: field f of plane p at slot f*n + p
    const int n = type == 0 ? dm.nyu : (type == 1 ? dm.nyv : dm.nyp);
    ZJobs jobs{};
    int nf;
    if (type == 0) {
      jobs.j[0] = ZJob{h->u1, h->zp, nullptr, 0, n, 0};
      jobs.j[1] = ZJob{h->u3, h->zp, nullptr, 0, n, n};
      jobs.j[2] = ZJob{h->r1, h->zp, nullptr, 0, n, 2 * n};
      nf = 3;
    } else if (type == 1) {
      jobs.j[0] = ZJob{h->u2, h->zp, nullptr, 0, n, 0};
      jobs.j[1] = ZJob{h->r2, h->zp, nullptr, 0, n, n};
      jobs.j[2] = ZJob{h->r3, h->zp, nullptr, 0, n, 2 * n};
      nf = 3;
    } else {
      jobs.j[0] = ZJob{h->p, h->zp, nullptr, 0, n, 0};
      nf = 1;
    }
    launch_zinv(h, jobs, nf, n, true);
    STATS a{(const double2*)h->zp, nf, n, type, h->ntilex, h->ntilez, g.dnx, g.dnz,
            h->st_psum, h->st_pfold};
    LH_DISPATCH_LEN(dm.px, launch_stats_t, double, h, a)
    STFIN f{};
    f.psum = h->st_psum; f.pfold = h->st_pfold; f.st = h->st_acc; f.cst = h->cst_acc;
    for (int q = 0; q < 13; ++q) f.qoff[q] = g.qoff[q];
    for (int l = 0; l < 6; ++l) { f.qmap[l] = qmap[type][l]; f.topsign[l] = tsg[type][l]; }
    f.nloc = nloc[type]; f.n = n; f.type = type; f.dnx = g.dnx; f.dnz = g.dnz; f.nyu = dm.nyu;
    f.nb = g.nb[type]; f.toplo = g.toplo;
    hipLaunchKernelGGL(k_stats_fin_plain, dim3((n * f.nloc + 255) / 256), dim3(256), 0, h->st, f);
    const long long nc = (long long)f.nloc * f.nb * g.dnx * g.dnz;
    hipLaunchKernelGGL(k_stats_fin_cond, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, h->st, f);
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->nstat += 1;
  return LH_OK;
}

int lh_stats_read(lh_handle* h, double* out, int* nstat, int reset) {
  if (!h || !out) return LH_ERR_ARG;
  const size_t n = 13 * (size_t)h->dm.nyu;
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (h->st_acc) HIPCHK(h, hipMemcpy(out, h->st_acc, n * sizeof(double), hipMemcpyDeviceToHost));
  else memset(out, 0, n * sizeof(double));
  if (nstat) *nstat = h->nstat;
  if (reset && h->st_acc) {
    StatsGeom g;
    int r = stats_geom(h, g);
    if (r) return r;
    HIPCHK(h, hipMemset(h->st_acc, 0, n * sizeof(double)));
    HIPCHK(h, hipMemset(h->cst_acc, 0, (size_t)g.clen * sizeof(double)));
    h->nstat = 0;
  }
  return LH_OK;
}

int lh_cond_stats_size(lh_handle* h, long long* len) {
  if (!h || !len) return LH_ERR_ARG;
  StatsGeom g;
  int r = stats_geom(h, g);
  if (r) return r;
  *len = g.clen;
  return LH_OK;
}

int lh_cond_stats_read(lh_handle* h, double* out, long long out_len, int reset) {
  if (!h || !out) return LH_ERR_ARG;
  StatsGeom g;
  int r = stats_geom(h, g);
  if (r) return r;
  if (out_len < g.clen) return fail(h, LH_ERR_ARG, "lh_cond_stats_read: output too short");
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (h->cst_acc)
    HIPCHK(h, hipMemcpy(out, h->cst_acc, (size_t)g.clen * sizeof(double), hipMemcpyDeviceToHost));
  else
    memset(out, 0, (size_t)g.clen * sizeof(double));
  if (reset && h->cst_acc) {
    HIPCHK(h, hipMemset(h->cst_acc, 0, (size_t)g.clen * sizeof(double)));
    HIPCHK(h, hipMemset(h->st_acc, 0, 13 * (size_t)h->dm.nyu * sizeof(double)));
    h->nstat = 0;
  }
  return LH_OK;
}

int lh_spectra_accumulate(lh_handle* h) {
  if (!h) return LH_ERR_ARG;
  const Dims& dm = h->dm;
  const size_t n = 5 * (size_t)dm.nyu * dm.ncol;
  if (!h->sp_acc) {
    HIPCHK(h, dalloc(&h->sp_acc, n));
    HIPCHK(h, hipMemset(h->sp_acc, 0, n * sizeof(double)));
  }
  hipLaunchKernelGGL(k_spectra, dim3((dm.ncol + 255) / 256, dm.nyu), dim3(256), 0, h->st, h->u1,
                     h->u2, h->u3, h->p, h->sp_acc, dm);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->st));
  return LH_OK;
}

int lh_spectra_read(lh_handle* h, double* out, int reset) {
  if (!h || !out) return LH_ERR_ARG;
  const Dims& dm = h->dm;
  const size_t pl = (size_t)dm.nyu * dm.ncol;
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (!h->sp_acc) {
    memset(out, 0, 5 * pl * sizeof(double));
    return LH_OK;
  }
  std::vector<double> t(5 * pl);
  HIPCHK(h, hipMemcpy(t.data(), h->sp_acc, t.size() * sizeof(double), hipMemcpyDeviceToHost));
  // device [f][slot][column] -> sp(j, column), j fastest, each array ncol x nyu
  for (int f = 0; f < 5; ++f)
    for (int c = 0; c < dm.ncol; ++c)
      for (int s = 0; s < dm.nyu; ++s)
        out[f * pl + (size_t)c * dm.nyu + s] = t[f * pl + (size_t)s * dm.ncol + c];
  if (reset) HIPCHK(h, hipMemset(h->sp_acc, 0, 5 * pl * sizeof(double)));
  return LH_OK;
}

}  // extern "C"
