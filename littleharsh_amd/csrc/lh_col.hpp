// lh_col.hpp -- column stage of the RK3 substep: one thread per spectral column (kx, kz)
// builds the rhs of solveU/solveV and runs the velocity Thomas sweeps of LUsolU/LUsolV
// (tridLU_3D_IB.f90:106-116, 179-189),
// the constant-flow-rate correction of the (0,0) column when enabled (flowrate_corr,
// littleharsh.f90:453-485), solveP with LUsolP (littleharsh.f90:557-612,
// tridLU_3D_IB.f90:362-396) and v_corr (littleharsh.f90:826-929).  Consecutive columns
// sit on consecutive lanes, so every row access is coalesced; rows are processed U at a
// time with all loads issued before the dependent recurrence.
#pragma once
#include "lh_kernels.hpp"

namespace lh {

struct COL {
  double2 *r1, *r2, *r3;
  double2 *u1, *u2, *u3, *p, *psi;
  const double *ayu, *ayv;
  const double* dg3;                       // modified uppers of the Poisson LU
  const double *dthdyu, *dthdyv, *k1x, *k1z, *k2x, *k2z, *yu;
  double* scal;
  double ab;            // aRK + bRK
  double gw0, gw1;
  int ctu;              // flag_ctpress == 0
  CTU ct;
};

// flowrate_corr body on the (0,0) column (thread of column 0 only)
__device__ void ctu_column(const CTU& a) {
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
  // flowrateRe(Qx1, u11) declares its dummy u(1:nn+2) while u11 is (0:nn+2)
  // (littleharsh.f90:443 vs 472-474): inside it u(j) is u11(j-1)
  double qx1 = 0.0, qx0 = 0.0;
  for (int j = 1; j <= n - 1; ++j) {
    const double dy = a.yv[j + a.d] - a.yv[j - 1 + a.d];
    qx1 += dy * u11[j - 1];
    qx0 += dy * a.u1[(size_t)(j + a.d) * a.ncol].x;
  }
  const double g = (a.scal[S_QXT] - qx0) / qx1;
  a.scal[S_MPGX] = a.scal[S_MPGX] + g;
  a.scal[S_DGX] = g;
  for (int j = 0; j <= n; ++j) a.u1[(size_t)(j + a.d) * a.ncol].x += g * u11[j];
}

// RHS0 + nonlinear + rhs of solveU/solveV (littleharsh.f90:614-731, 733-824) at plane s of
// component f for column c:
//   Lu  = dy2i(1,s) u(s-1) + (k2x + k2z + dy2i(2,s)) u(s) + dy2i(3,s) u(s+1)
//   du  = aRK/Re Lu - cRK Nu_old - dRK grad p [- dRK mpgx on the (0,0) u mode] - gRK Nu_new
//   rhs = (u + dt du) / axz
// The velocity sweeps (k_vel) evaluate it on the fly; the canopy planes, which the
// immersed-boundary stage rewrites in physical space first, get it from k_rhs_ib together
// with Lu_y on the tip planes.
struct RHS {
  double2 *u1, *u2, *u3;
  const double2* p;
  const double2 *no1, *no2, *no3;            // Nu_old
  const double2 *nn1, *nn2, *nn3;            // Nu_new
  double2 *r1, *r2, *r3, *tipl;
  const double *dyu2i, *dyv2i, *dthdyv, *k1x, *k2x, *k1z, *k2z, *scal;
  const int *planes_u, *planes_v;            // canopy planes (slots) of the u and v grids
  const unsigned char *ib_u, *ib_v;          // per slot: 1 on canopy planes
  int np_u, np_v;
  int tip_u0, tip_u1, tip_v0, tip_v1;        // slots
  double C1, C2, dR, g, bre;
};

struct RhsCol {
  double kx, kz, k2, axzi, kp, mp, dt;      // axzi = 1/axz
};

__device__ __forceinline__ RhsCol rhs_col(const RHS& a, const Dims& dm, int c, int f) {
  RhsCol r;
  const int i = c % dm.mx, k = c / dm.mx;
  const gptr<const double> scal = gp(a.scal);
  r.dt = scal[S_DT];
  const double beta = a.bre * r.dt;
  r.kx = gp(a.k1x)[i];
  r.kz = gp(a.k1z)[k];
  const double k2x = gp(a.k2x)[i], k2z = gp(a.k2z)[k];
  r.k2 = k2x + k2z;
  r.axzi = 1.0 / (1.0 - beta * (k2x + k2z));
  r.kp = -a.dR * (f == 0 ? r.kx : r.kz);
  r.mp = (f == 0 && c == 0) ? -a.dR * scal[S_MPGX] : 0.0;
  return r;
}

// rhs at slot s from u(s-1), u(s), u(s+1), the pressure rows (p(s-1) for u/w; p(s-1), p(s)
// for v), Nu_old(s) and Nu_new(s)
__device__ __forceinline__ double2 rhs_at(const RHS& a, const Dims& dm, const RhsCol& k, int f,
                                          int s, double2 um, double2 u0, double2 up, double2 pl,
                                          double2 pr, double2 vo, double2 vn) {
  const gptr<const double> wy = gp(f < 2 ? a.dyu2i : a.dyv2i) + 3 * s;
  const double2 Lu = dadd(dadd(drm(wy[0], um), drm(k.k2 + wy[1], u0)), drm(wy[2], up));
  double2 du;
  if (f < 2) {
    du = dadd(dadd(drm(a.C1, Lu), drm(a.C2, vo)), dim(k.kp, pl));
    du.x += k.mp;
  } else {
    const double C3v = -a.dR * dm.ddthetavi * gp(a.dthdyv)[s];
    du = dadd(dadd(drm(a.C1, Lu), drm(a.C2, vo)), drm(C3v, dsub(pr, pl)));
  }
  du = dsub(du, drm(a.g, vn));
  return drm(k.axzi, dadd(u0, drm(k.dt, du)));
}

// rhs on the canopy planes (grid: columns x planes x component) and Lu_y on the tip planes
// The component is a template argument: picked at run time among the fields of the RHS
// argument, the buffers would keep the whole argument in per-lane scratch memory.
template <int f>
__device__ __forceinline__ void rhs_ib_f(const RHS& a, const Dims& dm, int c, int q) {
  const int ncol = dm.ncol;
  if (c >= ncol || q >= (f < 2 ? a.np_u : a.np_v)) return;
  const int s = gp(f < 2 ? a.planes_u : a.planes_v)[q];
  const G2 uu = g2(f == 0 ? a.u1 : (f == 1 ? a.u3 : a.u2));
  const G2 no = g2(f == 0 ? a.no1 : (f == 1 ? a.no3 : a.no2));
  const G2 nn = g2(f == 0 ? a.nn1 : (f == 1 ? a.nn3 : a.nn2));
  const G2 rr = g2(f == 0 ? a.r1 : (f == 1 ? a.r3 : a.r2));
  const G2 pp = g2(a.p);
  const RhsCol k = rhs_col(a, dm, c, f);
  const size_t o = (size_t)s * ncol + c;
  const double2 um = uu[o - ncol], u0 = uu[o], up = uu[o + ncol];
  const double2 pl = pp[o - ncol], pr = f == 2 ? (double2)pp[o] : d2(0.0, 0.0);
  const double2 vo = a.C2 != 0.0 ? (double2)no[o] : d2(0.0, 0.0);   // C2 = -cRK is 0 in substep 1
  rr[o] = rhs_at(a, dm, k, f, s, um, u0, up, pl, pr, vo, nn[o]);
  const bool tip = f < 2 ? (s == a.tip_u0 || s == a.tip_u1) : (s == a.tip_v0 || s == a.tip_v1);
  if (tip) {
    const gptr<const double> wy = gp(f < 2 ? a.dyu2i : a.dyv2i) + 3 * s;
    const int tipr = f < 2 ? 2 * f + (s == a.tip_u0 ? 0 : 1) : 4 + (s == a.tip_v0 ? 0 : 1);
    g2(a.tipl)[(size_t)tipr * ncol + c] = dadd(dadd(drm(wy[0], um), drm(wy[1], u0)), drm(wy[2], up));
  }
}
// f: 0 u1, 1 u3, 2 u2 (uniform over the block)
__device__ __forceinline__ void rhs_ib_body(const RHS& a, const Dims& dm, int c, int q, int f) {
  if (f == 0) rhs_ib_f<0>(a, dm, c, q);
  else if (f == 1) rhs_ib_f<1>(a, dm, c, q);
  else rhs_ib_f<2>(a, dm, c, q);
}
__global__ void __launch_bounds__(256) k_rhs_ib(RHS a, Dims dm) {
  rhs_ib_body(a, dm, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, blockIdx.z);
}
// the canopy rhs from a job table: 64 columns per block, a channel's grid is gx column blocks
// x npl planes x 3 components
struct RhsIbJob {
  RHS a;
  Dims dm;
  int blk0, gx, gy;
};
__global__ void __launch_bounds__(64) k_rhs_ib_jobs(const RhsIbJob* __restrict__ jt, const int* __restrict__ map) {
  const RhsIbJob& J = jt[map[blockIdx.x]];
  const int lb = blockIdx.x - J.blk0;
  const RHS a = J.a;
  const Dims dm = J.dm;
  rhs_ib_body(a, dm, (lb % J.gx) * 64 + threadIdx.x, (lb / J.gx) % J.gy, lb / (J.gx * J.gy));
}

// solveU/solveV with the rhs built on the fly: forward elimination of LUsolU/LUsolV
// (tridLU_3D_IB.f90:106-116) walking up the planes with the rhs of every non-canopy plane
// computed from u(s-1..s+1), p, Nu_old, Nu_new in registers (canopy planes read the rhs the
// immersed-boundary stage left), the eliminated rows parked in the rhs array, then the
// back substitution into u.  F: 0 u1, 1 u3, 2 u2.
template <int U, int F>
__device__ __forceinline__ void solve_rhs(const RHS& a, const Dims& dm, int c,
                                          const double* __restrict__ ayp) {
  const int ncol = dm.ncol;
  const gptr<const double> ay = gp(ayp);
  const G2 no = g2(F == 0 ? a.no1 : (F == 1 ? a.no3 : a.no2));
  const G2 nn = g2(F == 0 ? a.nn1 : (F == 1 ? a.nn3 : a.nn2));
  const G2 uu = g2(F == 0 ? a.u1 : (F == 1 ? a.u3 : a.u2));
  const G2 x = g2(F == 0 ? a.r1 : (F == 1 ? a.r3 : a.r2));
  const gptr<const unsigned char> ib = gp(F < 2 ? a.ib_u : a.ib_v);
  const G2 pp = g2(a.p);
  const int nj = F < 2 ? dm.nyu : dm.nyv;
  const RhsCol k = rhs_col(a, dm, c, F);
  const double2 zero = d2(0.0, 0.0);
  // Nu_old enters with C2 = -cRK, which is 0 in the first substep: its rows are not read then
  const bool use_old = a.C2 != 0.0;
  // row 0: rhs = 0 (boundary row)
  double2 prev = drm(ay[1], zero);
  x[c] = prev;
  double2 um = uu[c], u0 = uu[(size_t)ncol + c];
  double2 pl = pp[c];                          // v: p(s-1) of plane s = 1
  int r = 1;
  for (; r < nj - 1; r += U) {
    double2 up[U], pr[U], vo[U], vn[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int s = r + q;
      if (s < nj - 1) {
        const size_t o = (size_t)s * ncol + c;
        up[q] = uu[o + ncol];
        pr[q] = F == 2 ? pp[o] : pp[o - ncol];
        if (use_old) vo[q] = no[o];
        else vo[q] = zero;
        vn[q] = nn[o];
      }
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int s = r + q;
      if (s < nj - 1) {
        const size_t o = (size_t)s * ncol + c;
        double2 rhs;
        if (ib[s]) rhs = x[o];
        else if (F == 2) rhs = rhs_at(a, dm, k, F, s, um, u0, up[q], pl, pr[q], vo[q], vn[q]);
        else rhs = rhs_at(a, dm, k, F, s, um, u0, up[q], pr[q], zero, vo[q], vn[q]);
        if (F == 2) pl = pr[q];
        prev = drm(ay[3 * s + 1], dsub(rhs, drm(ay[3 * s], prev)));
        x[o] = prev;
        um = u0;
        u0 = up[q];
      }
    }
  }
  // last row: rhs = 0
  prev = drm(ay[3 * (nj - 1) + 1], dsub(zero, drm(ay[3 * (nj - 1)], prev)));
  uu[(size_t)(nj - 1) * ncol + c] = prev;
  r = nj - 2;
  for (; r - U + 1 >= 0; r -= U) {
    double2 v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) v[q] = x[(size_t)(r - q) * ncol + c];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      prev = dsub(v[q], drm(ay[3 * (r - q) + 2], prev));
      uu[(size_t)(r - q) * ncol + c] = prev;
    }
  }
  for (; r >= 0; --r) {
    prev = dsub(x[(size_t)r * ncol + c], drm(ay[3 * r + 2], prev));
    uu[(size_t)r * ncol + c] = prev;
  }
}

template <int U>
__device__ __forceinline__ void vel_column(const RHS& a, const Dims& dm, int c, int f,
                                           const double* __restrict__ ayu, const double* __restrict__ ayv) {
  if (f == 0) solve_rhs<U, 0>(a, dm, c, ayu);
  else if (f == 1) solve_rhs<U, 1>(a, dm, c, ayu);
  else solve_rhs<U, 2>(a, dm, c, ayv);
}



// constant-flow-rate correction (column 0), solveP and v_corr, one column per thread
template <int U>
__device__ __forceinline__ void pres_column(const COL& a, const Dims& dm, int c) {
  const int ncol = dm.ncol;
  if (a.ctu && c == 0) ctu_column(a.ct);
  // ---- solveP: divergence, psi = dtirk*div, forward elimination of LUsolP
  const int i = c % dm.mx, k = c / dm.mx;
  const double kx = gp(a.k1x)[i], kz = gp(a.k1z)[k];
  const double dt = gp(a.scal)[S_DT];
  const double dtirk = (1.0 / dt) / a.ab;
  const int np = dm.nyp;
  const double dv = dm.ddthetavi;
  const G2 u1 = g2(a.u1);
  const G2 u2 = g2(a.u2);
  const G2 u3 = g2(a.u3);
  const G2 psi = g2(a.psi);
  const G2 pp = g2(a.p);
  const gptr<const double> dthdyu = gp(a.dthdyu), dthdyv = gp(a.dthdyv), dg3 = gp(a.dg3);
  double2 prev = d2(0.0, 0.0);
  double2 vlow = u2[c];   // v face below p row 0
  // LU_decP on the fly: the lower and the inverse pivot of each row come from the operator
  // row and the running modified upper a3p (the same operations k_buildP used for dg3)
  const double k2x = gp(a.k2x)[i], k2z = gp(a.k2z)[k];
  double a3p = 0.0;
  int r = 0;
  // rows in chunks of U; every load of a chunk is issued (rows past the last one read the
  // last one again), so that a chunk waits for memory once
  const double dypin = gp(a.yu)[2 + dm.d] - gp(a.yu)[1 + dm.d];
  for (; r < np; r += U) {
    double2 a1[U], a3[U], a2[U];
    double hu[U], hv[U + 1];              // dthdyu(rr+1), dthdyv(rr) of the chunk's rows
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int rq = min(r + q, np - 1);
      const size_t ou = (size_t)(rq + 1) * ncol + c;
      a1[q] = u1[ou];
      a3[q] = u3[ou];
      a2[q] = u2[ou];
      hu[q] = dthdyu[rq + 1];
    }
#pragma unroll
    for (int q = 0; q <= U; ++q) hv[q] = dthdyv[min(r + q, np)];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int rr = r + q;
      if (rr < np) {
        const PoisRow w = poisson_row_v(rr, c == 0, k2x, k2z, np, dv, hu[q], hv[q], hv[q + 1], dypin);
        const double g1 = w.lo;
        const double g2 = 1.0 / (w.dg - w.lo * a3p);
        a3p = w.up * g2;
        const double2 dvg = dadd(dadd(dim(kx, a1[q]), dim(kz, a3[q])),
                                 drm(dv * hu[q], dsub(a2[q], vlow)));
        vlow = a2[q];
        double2 ps = drm(dtirk, dvg);
        if (c == 0 && rr == 0) ps = d2(0.0, 0.0);
        if (rr == 0) prev = drm(g2, ps);
        else prev = drm(g2, dsub(ps, drm(g1, prev)));
        psi[(size_t)rr * ncol + c] = prev;
      }
    }
  }
  // ---- back substitution fused with p += psi and v_corr
  const double dtrk = dt * a.ab;
  const double kxr = dtrk * kx, kzr = dtrk * kz, dtrk_u2 = dtrk * dv;
  double2 pnext = prev;
  {
    const size_t o = (size_t)(np - 1) * ncol + c;
    pp[o] = dadd(pp[o], pnext);
    const size_t ou = (size_t)np * ncol + c;
    u1[ou] = dsub(u1[ou], dim(kxr, pnext));
    u3[ou] = dsub(u3[ou], dim(kzr, pnext));
  }
  r = np - 2;
  for (; r >= 0; r -= U) {
    double2 sv[U], pv[U], a1[U], a3[U], a2[U];
    double g3[U], hv[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int rr = max(r - q, 0);
      const size_t o = (size_t)rr * ncol + c, ou = (size_t)(rr + 1) * ncol + c;
      sv[q] = psi[o];
      g3[q] = dg3[o];
      pv[q] = pp[o];
      a1[q] = u1[ou];
      a3[q] = u3[ou];
      a2[q] = u2[ou];
      hv[q] = dthdyv[rr + 1];
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int rr = r - q;
      if (rr >= 0) {
        const size_t o = (size_t)rr * ncol + c, ou = (size_t)(rr + 1) * ncol + c;
        const double2 ps = dsub(sv[q], drm(g3[q], pnext));
        pp[o] = dadd(pv[q], ps);
        u1[ou] = dsub(a1[q], dim(kxr, ps));
        u3[ou] = dsub(a3[q], dim(kzr, ps));
        u2[ou] = dsub(a2[q], drm(dtrk_u2 * hv[q], dsub(pnext, ps)));
        pnext = ps;
      }
    }
  }
  const size_t top = (size_t)(dm.nyu - 1) * ncol + c;
  u1[c] = drm(-a.gw0, u1[c + ncol]);
  u1[top] = drm(-a.gw1, u1[top - ncol]);
  u3[c] = drm(-a.gw0, u3[c + ncol]);
  u3[top] = drm(-a.gw1, u3[top - ncol]);
  u2[c] = d2(0.0, 0.0);
  u2[(size_t)(dm.nyv - 1) * ncol + c] = d2(0.0, 0.0);
}

}  // namespace lh
