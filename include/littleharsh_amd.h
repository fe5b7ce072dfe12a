/*
 * littleharsh_amd.h -- C-ABI of the MI355X (gfx950) solver for the Littleharsh channel-DNS
 * time-step hot path.  Plain pointers and sizes only; every entry point returns an int
 * status (LH_OK = 0) and, on failure, leaves a message for lh_last_error().
 *
 * The boundary sits at the RK3 substep: one lh_substep call replaces the body of the
 * reference's `do kRK = 1,3` loop, littleharsh.f90:166-188, and lh_step also covers the
 * constant-pressure-gradient flow-rate bookkeeping after it, littleharsh.f90:193-197.
 *
 * Conventions (single resolution, every column in the reference's middle band):
 *   ny    reference input ny (input.in:19), dsty canopy planes per wall (input.in:62)
 *   u,w   planes j = -dsty .. ny+dsty       (jlim(:,ugrid,2), start.f90:1158-1159)
 *   v     planes j = -dsty .. ny+dsty-1     (jlim(:,vgrid,2), start.f90:1165-1166)
 *   p     planes j = -dsty+1 .. ny+dsty-1   (jlim(:,pgrid,2), start.f90:1172-1173)
 *   column c = k*(nx/2+1) + i, i = 0..nx/2, k = 0..nz-1 in FFT order (columns_i/k at one
 *   rank, start.f90:1105-1119).
 * Host state arrays are in the reference's own layout: complex(8) f(j, column), i.e.
 * interleaved (re, im) doubles with j fastest (type cfield, declaration.f90:133-135).
 *
 * One handle per process and device; calls on one handle must not overlap.  Work is
 * issued on one HIP stream owned by the handle and every call is synchronous at return.
 */
#ifndef LITTLEHARSH_AMD_H
#define LITTLEHARSH_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define LH_OK 0
#define LH_ERR_ARG 1      /* invalid argument or unsupported size                      */
#define LH_ERR_HIP 2      /* HIP runtime error                                         */
#define LH_ERR_NUMERIC 3  /* non-finite value detected                                 */

/* Precision of the nonlinear (FFT + product) stage.  The tridiagonal solves, the pressure
 * Poisson problem, the mean-flow mode and the statistics always run in FP64. */
#define LH_PREC_FP64 0
#define LH_PREC_MIXED 1

typedef struct lh_config {
  /* sizes: Fourier modes N(1,1), N(2,1) and collocation points Ngal(1,1), Ngal(2,1)
   * (start.f90:97-118); the solver requires px = 3*nx/2, pz = 3*nz/2 */
  int nx, nz, px, pz;
  int ny, dsty;
  /* physics (start.f90:69-76, 140-151) */
  double re, alp, bet;
  int flag_ctpress;   /* 1: constant mean pressure gradient mpgx; 0: constant flow rate */
  double mpgx;        /* initial mean pressure gradient */
  double qxt;         /* target flow rate QxT for flag_ctpress == 0 (start.f90:1546-1548) */
  double ddthetavi;   /* 1/dtheta (y_grid_canopy.f90:43) */
  /* metric arrays over the full reference index ranges (y_grid_canopy.f90:46-52):
   * yu, dthdyu: ny+2*dsty+1 values (j = -dsty..ny+dsty);
   * yv, dthdyv: ny+2*dsty values   (j = -dsty..ny+dsty-1);
   * dyu2i, dyv2i: Fortran dyu2i(3, j) -- three weights per plane, weight index fastest */
  const double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i;
  const double *gridweighting;  /* gridweighting(midband,1:2) (start.f90:500-501) */
  const double *rk;             /* aRK(1:3), bRK(1:3), gRK(1:3), cRK(1:3), dRK(1:3) */
  /* wavenumbers from def_k (start.f90:688-701): k1x = Im k1F_x (nx/2+1), k2x = k2F_x,
   * k1z = Im k1F_z (nz, FFT order), k2z = k2F_z */
  const double *k1x, *k2x, *k1z, *k2z;
  /* immersed-boundary points list_ib(3, nlist, grid): 1-based i in 1..px, k in 1..pz,
   * plane j (geom_sym_canopy.f90:98-234 after the getbounds plane filter) */
  int nlist_u, nlist_v;
  const int *list_u, *list_v;
  int tip_u[2], tip_v[2];  /* tip planes (tridLU_3D_IB.f90:237, 317) */
  /* canopy tiling for the conditional statistics (start.f90:456-469) */
  int ntilex, ntilez;
  int band_v[2];           /* N(3,1), N(3,2): v-grid band limits (y_grid_canopy.f90:152-163) */
  int precision;           /* LH_PREC_FP64 or LH_PREC_MIXED */
} lh_config;

typedef struct lh_handle lh_handle;

/* Allocate device state and upload the set-up; replaces the solver part of start()
 * (start.f90:392-652: FFT initialisation, RK tables, LU_buildP).  All pointers are copied
 * and may be freed on return. */
int lh_create(const lh_config* cfg, lh_handle** out);
/* Fails with LH_ERR_ARG while the handle belongs to a batch: destroy the batch first. */
int lh_destroy(lh_handle* h);
const char* lh_last_error(const lh_handle* h);

/* Upload / download u1, u2, u3, p (reference layout, see above).  Nu arrays may be NULL;
 * they carry the previous substep's nonlinear term (only read when cRK(kRK) != 0). */
int lh_set_state(lh_handle* h, const double* u1, const double* u2, const double* u3,
                 const double* p);
int lh_get_state(lh_handle* h, double* u1, double* u2, double* u3, double* p);
int lh_set_nonlinear(lh_handle* h, const double* nu1, const double* nu2, const double* nu3);
int lh_get_nonlinear(lh_handle* h, double* nu1, double* nu2, double* nu3);

/* Forcing added to the nonlinear term Nu_new of every substep, in the layout of
 * lh_set_nonlinear (interior planes of u, v, w); the stored Nu (Nu_old of the next substep)
 * includes it.  A NULL field is zero; all three NULL remove the forcing.  The learned
 * closure's subfilter-stress divergence d_j tau_ij enters the coarse channel this way.  The
 * reference has no such term: it is the coupling point of the closure. */
int lh_set_forcing(lh_handle* h, const double* f1, const double* f2, const double* f3);
/* The forcing last set (lh_set_forcing or lh_closure_apply), layout of lh_get_nonlinear. */
int lh_get_forcing(lh_handle* h, double* f1, double* f2, double* f3);

/* One RK3 substep kRK in 1..3 with time step dt: RHS0_u1/u2/u3, nonlinear, solveU/V/U,
 * meanflow_ctU (flag_ctpress == 0), solveP, v_corr (littleharsh.f90:166-188).
 * dt <= 0 uses the device CFL time step of the last lh_cfl_dt / lh_step_cfl.  With
 * kRK = 3 and flag_ctpress == 1 it also evaluates the flow rate (meanflow_ctP,
 * littleharsh.f90:193-197), so three lh_substep calls do what one lh_step does.
 * Non-finite values: the first substep of a step (kRK = 1), lh_step, lh_step_cfl and
 * lh_batch_step_cfl take the CFL maximum of the state at the start of the step with a
 * reduction that keeps NaN, and return LH_ERR_NUMERIC when it is not finite (or, for the
 * CFL-controlled steps, when the time step is not finite and positive).  The kernels run to
 * completion either way; the state is then not usable. */
int lh_substep(lh_handle* h, int krk, double dt);

/* Stage-level hooks for parity bisection: the three stages of substep kRK that lh_substep
 * runs, one at a time, on the same kernels (synchronous at return).  lh_substep equals
 * lh_nonlinear, lh_solve_u, lh_solve_p in turn.
 *   lh_nonlinear: Nu_new of the current state (call nonlinear, littleharsh.f90:171; the
 *     routine is absent from the reference), into the buffer of this substep, which
 *     lh_get_nonlinear then returns; the CFL maximum at kRK = 1.
 *   lh_solve_u: RHS0_u1/u2/u3 (littleharsh.f90:166-168, with Nu_old from the other buffer
 *     and Nu_new from lh_nonlinear), the immersed boundaries and solveU/V/U
 *     (littleharsh.f90:173-175; tridLU_3D_IB.f90:50-359): u, v, w before the pressure step.
 *     RHS0 is formed inside the sweep, so it has no hook of its own.
 *   lh_solve_p: meanflow_ctU when flag_ctpress = 0, solveP and v_corr
 *     (littleharsh.f90:178-188): v_corr is fused into the Poisson sweep. */
int lh_nonlinear(lh_handle* h, int krk);
int lh_solve_u(lh_handle* h, int krk, double dt);
int lh_solve_p(lh_handle* h, int krk, double dt);

/* One full RK3 step (three substeps) followed by meanflow_ctP when flag_ctpress == 1
 * (littleharsh.f90:164-197).  Launches are replayed from a HIP graph per dt mode. */
int lh_step(lh_handle* h, double dt);

/* CFL-controlled step: dt = cfl / max(|u|/dx + |v|/dy + |w|/dz) over the physical points
 * of the state at the start of the step, evaluated on the device (no host round trip);
 * *dt_used (may be NULL) receives the step taken. */
int lh_step_cfl(lh_handle* h, double cfl, double* dt_used);
/* The CFL time step of the current state, dt = cfl / max(...) as above, without stepping;
 * it also becomes the device time step used by lh_substep(h, krk, dt <= 0).  The CFL
 * formula is project-defined: dtc/dti are declared (declaration.f90:51) but set by the
 * absent nonlinear routine. */
int lh_cfl_dt(lh_handle* h, double cfl, double* dt);

/* Scalars: mean pressure gradient mpgx, flow rate Qx (flowrateIm, littleharsh.f90:412-431),
 * the last dt and the last CFL maximum. */
int lh_get_scalars(lh_handle* h, double* mpgx, double* qx, double* dt, double* cflmax);

/* max |div u| over the p grid (error, littleharsh.f90:323-352); NaN when the velocity holds
 * a NaN (the reduction keeps it). */
int lh_max_div(lh_handle* h, double* err);

/* lh_max_div without any device allocation (no counterpart in the reference): for callers
 * that cannot allocate device memory, such as code under HIP graph capture, tight serving
 * loops and harnesses that forbid hipMalloc.  Writes max |div u| (NaN kept) to the device
 * double err_dev, in order on the handle's stream, and returns without waiting for it.  Its
 * partial maxima live in a static device array of 4096 doubles per GPU, so the grid may hold
 * at most 1,048,576 spectral columns (LH_ERR_ARG beyond).  Calls that overlap on one GPU
 * (two streams or two host threads) share that array and race. */
int lh_max_div_noalloc(lh_handle* h, double* err_dev);

/* Statistics (stats.f90:1-108) of the current state: per-plane sums over the physical
 * collocation points of 0 U, 1 U2, 2 V, 3 V2, 4 W, 5 W2, 6 P, 7 P2, 8 U*V_itp, 9 U*W,
 * 10 V*W_itp, 11 wx, 12 wx^2 (the order of write_stats, stats.f90:337-349).  Rows 0-1, 4-5,
 * 8-9 are on the u grid, 2-3 and 10-12 on the v grid, 6-7 on the p grid.  The planes the
 * reference sums are made by its absent nonlinear routine; here v_itp is v averaged to the
 * u centres (0 on the two ghost planes), w_itp is w averaged to the v faces and wx =
 * dw/dy - dv/dz at the v faces, all on the 3/2 collocation grid.  The wx rows keep the
 * reference's loop bounds (stats.f90:54-56: x < nx, z < nz of that grid).
 * Each call adds one sample; lh_stats_read returns the sums. */
int lh_stats_accumulate(lh_handle* h);
/* out: 13 * (ny + 2*dsty + 1) doubles; row q, plane slot s (0 = first plane of the row's
 * grid) at out[q*(ny + 2*dsty + 1) + s].  *nstat (may be NULL) gets the number of samples.
 * reset != 0 clears the plain and conditional accumulators afterwards. */
int lh_stats_read(lh_handle* h, double* out, int* nstat, int reset);
/* Conditional statistics (conddown1/2, condup1/2, stats.f90:110-253): the planes of the
 * bottom canopy band and the mirrored planes of the top band (v sign flipped) folded onto
 * one dnx x dnz tile (dnx = px/ntilex, dnz = pz/ntilez), bottom and top summed as recvXC
 * writes them (stats.f90:474-508).  out holds the 13 quantities in the order above, each a
 * Fortran array (dnx, dnz, nb) with nb = N(4,1)-N(4,0)+1 (u grid), N(3,1)-N(3,0)+1 (v grid)
 * or N(4,1)-N(4,0)-1 (p grid); lh_cond_stats_size gives the total length. */
int lh_cond_stats_size(lh_handle* h, long long* len);
int lh_cond_stats_read(lh_handle* h, double* out, long long out_len, int reset);

/* Spectra (spectra.f90:1-56): sp(j, column) += |f|^2 for U, V, W, then Re(u) Re(v_itp) +
 * Im(u) Im(v_itp) (UV), then P.  out: 5 consecutive arrays of (ny+2*dsty+1) x ncol doubles,
 * sp(j, column) with j fastest (v and p fill the first nyv / nyp rows of each column). */
int lh_spectra_accumulate(lh_handle* h);
int lh_spectra_read(lh_handle* h, double* out, int reset);

/* Launch parameters and the autotuner (no counterpart in the reference).  The results do
 * not depend on them: each only changes how work is split into blocks.
 *   xraw_lc  levels per block of the sliding x pass        (>= 1)
 *   zasm_lc  levels per block of the z assembly            (>= 1)
 *   col_bs   threads per block of the canopy rhs kernel    (32..256, multiple of 32; used
 *            by channels with an x or z transform length above 144 only: the others, and
 *            the column stage, run 64 columns per block)
 * lh_create sets the two level counts from the grid and the device: each block of those
 * kernels holds a CU's LDS, so the launch runs in ceil(blocks / resident slots) rounds,
 * and the count minimising rounds x (levels + start-up) is taken.
 * lh_autotune times `steps` CFL steps per candidate (the current values, their halves
 * and doubles, a few fixed sizes) and keeps the fastest value of each parameter in turn;
 * the state is saved before and restored after. */
int lh_set_launch(lh_handle* h, int xraw_lc, int zasm_lc, int col_bs);
int lh_get_launch(lh_handle* h, int* xraw_lc, int* zasm_lc, int* col_bs);

/* FFT plans (no counterpart in the reference, whose cfftf/rfftf factor each length one
 * way, cft_buff.f and rft_buff.f).  Each transform length has plan 0, its default radix
 * sequence, and the lengths 48, 72, 96 and 144 also plan 1:
 *   length  plan 0 (radices, threads per line)   plan 1
 *   48      12 x 4, 4                           6 x 8, 8
 *   72      12 x 6, 6                           8 x 9, 9
 *   96      4 x 4 x 6, 8                        8 x 12, 12
 *   144     12 x 12, 12                         9 x 16, 16
 * xplan selects the plan of the x transforms (length px), zplan that of the z transforms
 * (length pz).  The plans agree to rounding (not bit for bit).  lh_get_plans also returns
 * how many plans each length has (xcount, zcount; 1 or 2). */
int lh_set_plans(lh_handle* h, int xplan, int zplan);
int lh_get_plans(lh_handle* h, int* xplan, int* zplan, int* xcount, int* zcount);

/* lh_autotune first times every (xplan, zplan) pair, then the launch parameters. */
int lh_autotune(lh_handle* h, int steps);

/* Many independent channels (handles of any sizes on one device) advanced together: one
 * CFL step of every instance per call, issued as one captured HIP graph (no counterpart in
 * the reference; the instances share nothing).  Per substep, the nonlinear stages of the
 * instances are parallel branches, each on its handle's stream (with the canopy stage of
 * instances whose x or z length exceeds 144); after them the canopy stage of the other
 * instances is one launch per kernel and (length, FFT plan) group, and the column stage
 * (LU of the velocity matrices, velocity sweeps, pressure step) of all instances one launch
 * of each kernel.  A member computes exactly what it computes stepped alone.  After the
 * step every member's scalars are checked as by lh_step_cfl; the error names the first
 * member that failed. */
/* A batch re-captures its graph when the CFL number changes or when any member's own graphs
 * were dropped (lh_set_launch, lh_set_plans, lh_autotune, lh_set_forcing switching the forcing on or off,
 * lh_closure_create); each handle may appear once.  Destroy the batch before its members. */
typedef struct lh_batch lh_batch;
int lh_batch_create(int n, lh_handle* const* handles, lh_batch** out);
int lh_batch_step_cfl(lh_batch* b, double cfl);
const char* lh_batch_last_error(const lh_batch* b);
int lh_batch_destroy(lh_batch* b);

/* Learned subfilter closure (no counterpart in the reference: the canopy-layer closure of
 * the coarse-grid model).  A multilayer perceptron nin -> nhid (tanh) x nlayers -> nout
 * (linear), trained and evaluated on the GPU by BF16 matrix-core kernels accumulating in
 * FP32, with FP32 master weights and Adam (beta 0.9, 0.999, eps 1e-8); loss: mean squared
 * error over the nout outputs.  Arrays are row-major host arrays of floats:
 *   parameters  per layer W (out x in) then b (out), lh_mlp_param_count values
 *   x           rows of nin inputs, y rows of nout targets / outputs
 * lh_mlp_loss_grad evaluates one batch of `batch` rows (loss and the FP32 gradient of every
 * parameter); lh_mlp_train takes `steps` Adam steps, each on `batch` rows drawn with
 * replacement from the data set given to lh_mlp_set_data (kept in device memory). */
typedef struct lh_mlp lh_mlp;
int lh_mlp_create(int nin, int nhid, int nlayers, int nout, int batch, unsigned long long seed,
                  lh_mlp** out);
int lh_mlp_destroy(lh_mlp* m);
const char* lh_mlp_last_error(lh_mlp* m);
long long lh_mlp_param_count(lh_mlp* m);
int lh_mlp_get_params(lh_mlp* m, float* params);
int lh_mlp_set_params(lh_mlp* m, const float* params);
int lh_mlp_set_data(lh_mlp* m, const float* x, const float* y, long long n);
int lh_mlp_loss_grad(lh_mlp* m, const float* x, const float* y, float* loss, float* grad);
int lh_mlp_train(lh_mlp* m, int steps, float lr, unsigned long long seed, float* loss_last);
/* The closure's shape (244 or up to 256 inputs, four hidden layers of 64, up to 16 outputs)
 * trains by one fused pass per 16 batch rows plus one launch for every layer's weight
 * gradients; lh_mlp_set_fused(m, 0) selects the per-layer kernels instead (1 restores the
 * default).  Both give bit-identical gradients and parameters; the switch is for tests and
 * timing.  Other shapes always use the per-layer kernels. */
int lh_mlp_set_fused(lh_mlp* m, int on);
int lh_mlp_predict(lh_mlp* m, const float* x, long long n, float* y);
/* Outputs at every point (j, k, i) of an ny x nz x nx grid, row (j nz + k) nx + i of y, from
 * inputs gathered on the device: the nch fields g (nch x ny x nz x nx) on the 3x3x3
 * neighbourhood (feature ((dj+1) 9 + (dk+1) 3 + (di+1)) nch + c, j clamped, k and i
 * periodic), then the extra field ex (ny x nz x nx) at the point when nin = 27 nch + 1
 * (ex NULL: nin = 27 nch).  Input f is standardised as (v - mean[f]) * (1 / scale[f]), the
 * reciprocal rounded to float (as numpy's float32 (v - mean) * (1 / scale)).  The coarse
 * channel evaluates the closure this way from its nine velocity gradients and the solid
 * fraction. */
int lh_mlp_predict_grid(lh_mlp* h, const float* g, int nch, const float* ex, int ny, int nz, int nx,
                        const float* mean, const float* scale, float* y);

/* ---- the learned closure inside a coarse channel (no counterpart in the reference) ----
 * lh_closure_create binds a network (244 inputs: the nine velocity gradients on the 3x3x3
 * stencil and the solid fraction; 6 outputs: the stresses uu, vv, ww, uv, uw, vw) to a
 * channel h whose interior u planes are the ny = nyu - 2 coarse centres at heights yc, with
 * nz x nx coarse points (nx = 2 (mx - 1)) spaced dxc, dzc, shifted by (shift_x, shift_z) to
 * the box-filter centres the network was trained at.  sf: solid fraction (ny x nz x nx);
 * xmean / xscale: input standardisation (244); ymean / yscale: output scaling (6).  Creating
 * it switches the channel's forcing on (zero until the first apply).
 * lh_closure_apply: from the current state, the forcing d_j tau_ij for the following
 * substeps, computed on the channel's stream (no host round trip); call it before each
 * lh_step. */
typedef struct lh_closure lh_closure;
int lh_closure_create(lh_handle* h, lh_mlp* net, const double* yc, double dxc, double dzc, double shift_x,
                      double shift_z, const float* sf, const float* xmean, const float* xscale,
                      const float* ymean, const float* yscale, lh_closure** out);
/* The same coupling with the Smagorinsky stresses cs2 (-2 Delta^2 |S| S_ij) in place of the
 * network, Delta^3 = dxc dzc dy (dy: centred spacing of the centres): the classical closure
 * the learned one is measured against. */
int lh_closure_create_smagorinsky(lh_handle* h, const double* yc, double dxc, double dzc, double shift_x,
                                  double shift_z, double cs2, lh_closure** out);
int lh_closure_apply(lh_closure* c);
/* on != 0: keep only the plane-mean column (kx = kz = 0) of the closure's forcing, the
 * stresses' effect on the mean profile (default 0: all columns). */
int lh_closure_set_mean_only(lh_closure* c, int on);
/* Destroy the closure before its channel and its network: it holds their device buffers and
 * the channel's stream.  The channel's forcing stays as last applied. */
int lh_closure_destroy(lh_closure* c);
const char* lh_closure_last_error(const lh_closure* c);

/* Timing (no counterpart in the reference; measurement support).  Mode 1: lh_step and
 * lh_step_cfl issue their kernels directly (no graph) with a HIP event pair around every
 * launch on the handle's stream; durations accumulate per kernel class.  Mode 2: the
 * captured step graph is replayed as usual, with one event pair around each replay on the
 * handle's stream (class LH_K_STEP, one launch per RK3 step).  Mode 0: off.
 * lh_timing_read returns the accumulated milliseconds, the launch count and the algorithmic
 * bytes one launch of that class moves (each input read once, each output written once;
 * 0 for LH_K_STEP). */
#define LH_K_ZINV 0   /* inverse z transforms of u, w, v                  */
#define LH_K_XNL 1    /* x transforms + products                          */
#define LH_K_ZFWD 2   /* forward z transforms of the products             */
#define LH_K_ASM 3    /* forward z transforms of the products -> Nu       */
#define LH_K_IBZ 4    /* canopy-plane z transforms                         */
#define LH_K_XIB 5    /* canopy-plane x transforms + forcing               */
#define LH_K_VEL 6    /* velocity Thomas sweeps                            */
#define LH_K_PRES 7   /* pressure solve + v_corr                          */
#define LH_K_MISC 8   /* small kernels: LU factors, mean flow, scalars     */
#define LH_K_RHS 9    /* RHS0 + nonlinear + rhs of the solves, per column  */
#define LH_K_STEP 10  /* one replay of the captured RK3 step graph (mode 2) */
#define LH_K_COUNT 11
int lh_timing_enable(lh_handle* h, int enable);
int lh_timing_read(lh_handle* h, int kernel, double* total_ms, long long* launches,
                   double* bytes_per_launch);
int lh_timing_reset(lh_handle* h);

#ifdef __cplusplus
}
#endif
#endif /* LITTLEHARSH_AMD_H */
