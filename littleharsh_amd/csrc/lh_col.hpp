// lh_col.hpp -- column stage of the RK3 substep: one thread per spectral column (kx, kz)
// runs the velocity Thomas sweeps of LUsolU/LUsolV (tridLU_3D_IB.f90:106-116, 179-189),
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
  const double *dg1, *dg2, *dg3;
  const double *dthdyu, *dthdyv, *k1x, *k1z;
  double* scal;
  double ab;            // aRK + bRK
  double gw0, gw1;
  int ctu;              // flag_ctpress == 0
  CTU ct;
};

template <int U>
__device__ __forceinline__ void thomas_u(double2* __restrict__ x, double2* __restrict__ out,
                                         const double* __restrict__ ay, int nj, int ncol, int c) {
  double2 prev = drm(ay[1], x[c]);
  x[c] = prev;
  int r = 1;
  for (; r + U <= nj; r += U) {
    double2 v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) v[q] = x[(size_t)(r + q) * ncol + c];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      prev = drm(ay[3 * (r + q) + 1], dsub(v[q], drm(ay[3 * (r + q)], prev)));
      x[(size_t)(r + q) * ncol + c] = prev;
    }
  }
  for (; r < nj; ++r) {
    prev = drm(ay[3 * r + 1], dsub(x[(size_t)r * ncol + c], drm(ay[3 * r], prev)));
    x[(size_t)r * ncol + c] = prev;
  }
  out[(size_t)(nj - 1) * ncol + c] = prev;
  r = nj - 2;
  for (; r - U + 1 >= 0; r -= U) {
    double2 v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) v[q] = x[(size_t)(r - q) * ncol + c];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      prev = dsub(v[q], drm(ay[3 * (r - q) + 2], prev));
      out[(size_t)(r - q) * ncol + c] = prev;
    }
  }
  for (; r >= 0; --r) {
    prev = dsub(x[(size_t)r * ncol + c], drm(ay[3 * r + 2], prev));
    out[(size_t)r * ncol + c] = prev;
  }
}

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

// velocity Thomas sweeps, one (column, component) per thread
// (blockIdx.y + f0 selects u1, u3 or u2)
template <int U>
__global__ void __launch_bounds__(256) k_vel(COL a, Dims dm, int f0) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dm.ncol) return;
  const int f = blockIdx.y + f0;
  if (f == 0) thomas_u<U>(a.r1, a.u1, a.ayu, dm.nyu, dm.ncol, c);
  else if (f == 1) thomas_u<U>(a.r3, a.u3, a.ayu, dm.nyu, dm.ncol, c);
  else thomas_u<U>(a.r2, a.u2, a.ayv, dm.nyv, dm.ncol, c);
}

// constant-flow-rate correction (column 0), solveP and v_corr, one column per thread
template <int U>
__global__ void __launch_bounds__(256) k_pres(COL a, Dims dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncol = dm.ncol;
  if (c >= ncol) return;
  if (a.ctu && c == 0) ctu_column(a.ct);
  // ---- solveP: divergence, psi = dtirk*div, forward elimination of LUsolP
  const int i = c % dm.mx, k = c / dm.mx;
  const double kx = a.k1x[i], kz = a.k1z[k];
  const double dt = a.scal[S_DT];
  const double dtirk = (1.0 / dt) / a.ab;
  const int np = dm.nyp;
  const double dv = dm.ddthetavi;
  double2* __restrict__ u1 = a.u1;
  double2* __restrict__ u2 = a.u2;
  double2* __restrict__ u3 = a.u3;
  double2* __restrict__ psi = a.psi;
  double2* __restrict__ pp = a.p;
  double2 prev = d2(0.0, 0.0);
  double2 vlow = u2[c];   // v face below p row 0
  int r = 0;
  for (; r < np; r += U) {
    double2 a1[U], a3[U], a2[U];
    double g1[U], g2[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      if (r + q < np) {
        const size_t ou = (size_t)(r + q + 1) * ncol + c;
        a1[q] = u1[ou];
        a3[q] = u3[ou];
        a2[q] = u2[ou];
        g1[q] = a.dg1[(size_t)(r + q) * ncol + c];
        g2[q] = a.dg2[(size_t)(r + q) * ncol + c];
      }
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int rr = r + q;
      if (rr < np) {
        const double2 dvg = dadd(dadd(dim(kx, a1[q]), dim(kz, a3[q])),
                                 drm(dv * a.dthdyu[rr + 1], dsub(a2[q], vlow)));
        vlow = a2[q];
        double2 ps = drm(dtirk, dvg);
        if (c == 0 && rr == 0) ps = d2(0.0, 0.0);
        if (rr == 0) prev = drm(g2[q], ps);
        else prev = drm(g2[q], dsub(ps, drm(g1[q], prev)));
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
    double g3[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int rr = r - q;
      if (rr >= 0) {
        const size_t o = (size_t)rr * ncol + c, ou = (size_t)(rr + 1) * ncol + c;
        sv[q] = psi[o];
        g3[q] = a.dg3[o];
        pv[q] = pp[o];
        a1[q] = u1[ou];
        a3[q] = u3[ou];
        a2[q] = u2[ou];
      }
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
        u2[ou] = dsub(a2[q], drm(dtrk_u2 * a.dthdyv[rr + 1], dsub(pnext, ps)));
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
