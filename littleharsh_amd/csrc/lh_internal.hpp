// lh_internal.hpp -- views across the library's translation units (not part of the C-ABI).
//
// The closure coupling (lh_closure.hip) reads the coarse channel's state and writes its
// forcing buffers (lh_api.hip), and evaluates the network (lh_mlp.hip) on device-resident
// inputs, all on the channel's stream.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/littleharsh_amd.h"

struct LhSolverView {
  const double2 *u, *v, *w;        // state, slot-major: [slot][col]
  double2 *f1, *f2, *f3;           // forcing added to Nu_new (u, v, w grids), same layout
  const double *k1x, *k1z;         // wavenumbers (device): kx of col % mx, kz of col / mx
  int ncol, nyu, nyv, nz, mx;
  hipStream_t st;
};

// Fills v; allocates the forcing buffers (zeroed) on first use and switches the forcing on,
// so that every following substep adds them.
int lh_solver_view(lh_handle* h, LhSolverView* v);

// Outputs (n x nout, device) of the network at every point of an ny x nz x nx grid from the
// nch device fields g and the device extra field ex (or null), standardised as
// (v - mean[f]) * rscale[f] with the device arrays mean and rscale (the reciprocals of the
// scales, rounded to float); launched on stream st.  nin must be 27 nch (+1 with ex).
int lh_mlp_eval_grid_dev(lh_mlp* h, const float* g, int nch, const float* ex, int ny, int nz, int nx,
                         const float* mean, const float* rscale, float* y, hipStream_t st);
int lh_mlp_dims(const lh_mlp* h, int* nin, int* nout);

// What the divergence monitor reads (lh_noalloc.hip): the handle's state and metrics on the
// device, its device and its stream.
struct LhDivView {
  const double2 *u1, *u2, *u3;     // state, slot-major: [slot][col]
  const double *dthdyu, *k1x, *k1z;
  double ddthetavi;
  int ncol, mx, nyp;
  int dev;
  hipStream_t st;
};
int lh_div_view(lh_handle* h, LhDivView* v);
// records msg as the handle's last error and returns code
int lh_fail(lh_handle* h, int code, const char* msg);
