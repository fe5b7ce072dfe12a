/*
 * lh_oracle.c -- CPU restatement of the Littleharsh RK3 substep, used ONLY as the checker
 * of the HIP library (tests, smoke(), and the CPU-baseline leg of bench.py).  It is test
 * infrastructure: nothing in the product path links or calls it.
 *
 * It keeps the reference's own data layout -- complex(8) f(j, column), j fastest, columns
 * in the np=1 order of proc_lims_columns (start.f90:1105-1119) -- so that the device
 * layout [plane][column] of the HIP library is checked against an independent indexing.
 *
 * Routines follow, one to one:
 *   RHS0_u1/u2/u3, laplacian_U/V     littleharsh.f90:614-731, 247-321
 *   nonlinear                         (FOU3D.f90 is absent from the reference; restated
 *                                      here as the divergence-form, 3/2-dealiased product
 *                                      term on the staggered grid -- parity unpinned by
 *                                      reference code, pinned by the closed-form tests)
 *   solveU/solveV, LUsolU/LUsolV      littleharsh.f90:733-824, tridLU_3D_IB.f90:50-194
 *   LU_build, LU_dec                  tridLU_3D_IB.f90:427-493, 601-640
 *   immersed_boundaries_U/V           tridLU_3D_IB.f90:198-359
 *   flowrate*, meanflow_ct*, LUsol0   littleharsh.f90:412-538, tridLU_3D_IB.f90:398-424,
 *                                      573-599, 686-703
 *   solveP, LU_buildP/decP, LUsolP    littleharsh.f90:557-612, tridLU_3D_IB.f90:362-396,
 *                                      497-571, 643-684
 *   v_corr, divergence, error         littleharsh.f90:826-937, 210-245, 323-352
 *   stats / spectra accumulation      stats.f90:1-108, spectra.f90:1-56
 *   four_to_phys / phys_to_four       (absent; restated with the rft/cft conventions of
 *                                      rft_buff.f:24-86 and cft_buff.f:1-65)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/littleharsh_amd.h"

typedef struct { double re, im; } cplx;

static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx cmul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r;
}
static inline cplx rmul(double s, cplx a) { cplx r = {s * a.re, s * a.im}; return r; }
/* (0 + i k) * a, exactly as complex(8) (0,k)*(a.re,a.im) */
static inline cplx imul(double k, cplx a) { cplx r = {-k * a.im, k * a.re}; return r; }

/* ------------------------------------------------------------------------------------ */
/* Mixed-radix complex FFT (decimation in time, radix 4/2/3/generic).  Unnormalised;     */
/* sign -1: sum x_n e^{-2 pi i nk/N} (cft/rft forward), +1: e^{+...} (inverse).          */
/* ------------------------------------------------------------------------------------ */
typedef struct {
  int n;
  int nfac;
  int fac[64];      /* pairs (p, m) */
  cplx* tw_f;       /* e^{-2 pi i m/n} */
  cplx* tw_b;       /* e^{+2 pi i m/n} */
} fftplan;

static int plan_init(fftplan* pl, int n) {
  pl->n = n;
  int m = n, k = 0;
  while (m > 1) {
    int p;
    if (m % 4 == 0) p = 4;
    else if (m % 2 == 0) p = 2;
    else if (m % 3 == 0) p = 3;
    else if (m % 5 == 0) p = 5;
    else return -1;
    m /= p;
    pl->fac[k++] = p;
    pl->fac[k++] = m;
  }
  pl->nfac = k / 2;
  pl->tw_f = (cplx*)malloc(sizeof(cplx) * n);
  pl->tw_b = (cplx*)malloc(sizeof(cplx) * n);
  for (int i = 0; i < n; ++i) {
    double a = 2.0 * M_PI * (double)i / (double)n;
    pl->tw_f[i].re = cos(a); pl->tw_f[i].im = -sin(a);
    pl->tw_b[i].re = cos(a); pl->tw_b[i].im = sin(a);
  }
  return 0;
}

static void plan_free(fftplan* pl) { free(pl->tw_f); free(pl->tw_b); }

static void fft_rec(cplx* out, const cplx* in, int fstride, int istride, const int* fac,
                    const cplx* tw, int n, int sign) {
  const int p = fac[0], m = fac[1];
  if (m == 1) {
    for (int j = 0; j < p; ++j) out[j] = in[(size_t)j * fstride * istride];
  } else {
    for (int j = 0; j < p; ++j)
      fft_rec(out + j * m, in + (size_t)j * fstride * istride, fstride * p, istride, fac + 2,
              tw, n, sign);
  }
  if (p == 2) {
    for (int u = 0; u < m; ++u) {
      cplx t = cmul(out[u + m], tw[u * fstride]);
      out[u + m] = csub(out[u], t);
      out[u] = cadd(out[u], t);
    }
  } else if (p == 4) {
    for (int u = 0; u < m; ++u) {
      cplx a0 = out[u];
      cplx a1 = cmul(out[u + m], tw[u * fstride]);
      cplx a2 = cmul(out[u + 2 * m], tw[2 * u * fstride]);
      cplx a3 = cmul(out[u + 3 * m], tw[3 * u * fstride]);
      cplx s0 = cadd(a0, a2), s1 = csub(a0, a2), s2 = cadd(a1, a3), s3 = csub(a1, a3);
      /* multiply s3 by -i (forward) or +i (inverse) */
      cplx r3 = sign < 0 ? (cplx){s3.im, -s3.re} : (cplx){-s3.im, s3.re};
      out[u] = cadd(s0, s2);
      out[u + 2 * m] = csub(s0, s2);
      out[u + m] = cadd(s1, r3);
      out[u + 3 * m] = csub(s1, r3);
    }
  } else {
    cplx scratch[8];
    for (int u = 0; u < m; ++u) {
      for (int q = 0; q < p; ++q) scratch[q] = cmul(out[u + q * m], tw[(q * u * fstride) % n]);
      for (int q1 = 0; q1 < p; ++q1) {
        cplx acc = scratch[0];
        for (int q = 1; q < p; ++q) {
          int idx = (int)(((long long)q * q1 * m * fstride) % n);
          acc = cadd(acc, cmul(scratch[q], tw[idx]));
        }
        out[u + q1 * m] = acc;
      }
    }
  }
}

/* out[k] = sum_n in[n*istride] w^{nk}, n,k in 0..N-1 */
static void fft_exec(const fftplan* pl, const cplx* in, int istride, cplx* out, int sign) {
  fft_rec(out, in, 1, istride, pl->fac, sign < 0 ? pl->tw_f : pl->tw_b, pl->n, sign);
}

/* ------------------------------------------------------------------------------------ */
/* Oracle state                                                                          */
/* ------------------------------------------------------------------------------------ */
typedef struct lho {
  int nx, nz, px, pz, mx, ncol, ny, d, nyu, nyv, nyp;
  double re, alp, bet, ddthetavi;
  int flag_ctpress;
  double mpgx, qx, qxt, dgx;
  double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i;   /* indexed by j+d */
  double gw[2];
  double aRK[3], bRK[3], gRK[3], cRK[3], dRK[3];
  double *k1x, *k2x, *k1z, *k2z;
  /* canopy masks: per plane (j+d), a px*pz byte mask or NULL; tips */
  unsigned char **mask_u, **mask_v;
  int tip_u[2], tip_v[2];
  int ntilex, ntilez, band_v[2];
  /* state, reference layout f[c*nj + (j - jlo)] */
  cplx *u1, *u2, *u3, *p, *psi, *div;
  cplx *du1, *du2, *du3, *Lu1, *Lu2, *Lu3;
  cplx *Nu1, *Nu2, *Nu3;
  cplx *F1, *F2, *F3;          /* optional forcing added to Nu_new (NULL: none) */
  double* DG;  /* DG(3, j, column) as f_dg(3, nyp, ncol) */
  fftplan plx, plz, plnx, plnz;
  double cflmax;
  /* statistics */
  double *st;  /* 13 x nyu */
  int nstat;
  double *sp;  /* 5 x (nyu x ncol) */
  double *cst; /* conditional stats */
  long long cst_len;
} lho;

#define UJ(o, j) ((j) + (o)->d)               /* array index of plane j (u and v grids) */
#define FU(o, f, c, j) (f)[(size_t)(c) * (o)->nyu + ((j) + (o)->d)]
#define FV(o, f, c, j) (f)[(size_t)(c) * (o)->nyv + ((j) + (o)->d)]
#define FP(o, f, c, j) (f)[(size_t)(c) * (o)->nyp + ((j) + (o)->d - 1)]

static double* dupd(const double* a, size_t n) {
  double* r = (double*)malloc(n * sizeof(double));
  memcpy(r, a, n * sizeof(double));
  return r;
}

/* ------------------------------------------------------------------------------------ */
/* Plane transforms.  A spectral plane is sp[k*mx + i] (i = 0..nx/2, k in FFT order),    */
/* a physical plane is ph[z*px + x].                                                     */
/* four_to_phys: zero-pad to (px/2+1) x pz, inverse cft along z (unscaled), inverse rft  */
/* along x (unscaled, Im of the kx=0 and Nyquist coefficients ignored, rft_buff.f:77-84) */
/* phys_to_four: forward rft along x (x 1/px, rft_buff.f:63-73), forward cft along z    */
/* (x 1/pz, cft_buff.f:37-46), truncation to the nx x nz modes.                          */
/* ------------------------------------------------------------------------------------ */
static void four_to_phys(const lho* o, const cplx* sp, double* ph, cplx* work /* >= pz*mx + 2*max(px,pz) */) {
  const int mx = o->mx, nz = o->nz, pz = o->pz, px = o->px;
  cplx* hz = work;                 /* [z][i], i < mx */
  cplx* line = work + (size_t)pz * mx;
  /* line holds a z line (pz) or an x line (px): the output starts after the longer one */
  cplx* lout = line + (px > pz ? px : pz);
  cplx* col = lout;                /* reuse */
  for (int i = 0; i < mx; ++i) {
    for (int z = 0; z < pz; ++z) line[z] = (cplx){0.0, 0.0};
    for (int k = 0; k < nz; ++k) {
      int kk = k < nz / 2 ? k : k + (pz - nz);
      line[kk] = sp[k * mx + i];
    }
    fft_exec(&o->plz, line, 1, col, +1);
    for (int z = 0; z < pz; ++z) hz[z * mx + i] = col[z];
  }
  /* two x rows per complex transform: c = A + iB */
  for (int z = 0; z < pz; z += 2) {
    const cplx* A = hz + (size_t)z * mx;
    const cplx* B = hz + (size_t)(z + 1) * mx;
    for (int n = 0; n < px; ++n) line[n] = (cplx){0.0, 0.0};
    line[0] = (cplx){A[0].re, B[0].re};
    for (int n = 1; n < mx; ++n) {
      line[n] = (cplx){A[n].re - B[n].im, A[n].im + B[n].re};
      line[px - n] = (cplx){A[n].re + B[n].im, -A[n].im + B[n].re};
    }
    fft_exec(&o->plx, line, 1, lout, +1);
    for (int n = 0; n < px; ++n) {
      ph[(size_t)z * px + n] = lout[n].re;
      ph[(size_t)(z + 1) * px + n] = lout[n].im;
    }
  }
}

static void phys_to_four(const lho* o, const double* ph, cplx* sp, cplx* work) {
  const int mx = o->mx, nz = o->nz, pz = o->pz, px = o->px;
  cplx* hz = work;
  cplx* line = work + (size_t)pz * mx;
  cplx* lout = line + px;
  const double sx = 1.0 / px, sz = 1.0 / pz;
  for (int z = 0; z < pz; z += 2) {
    for (int n = 0; n < px; ++n)
      line[n] = (cplx){ph[(size_t)z * px + n], ph[(size_t)(z + 1) * px + n]};
    fft_exec(&o->plx, line, 1, lout, -1);
    cplx* A = hz + (size_t)z * mx;
    cplx* B = hz + (size_t)(z + 1) * mx;
    for (int k = 0; k < mx; ++k) {
      cplx zk = lout[k], zn = lout[k == 0 ? 0 : px - k];
      cplx a = {0.5 * (zk.re + zn.re), 0.5 * (zk.im - zn.im)};
      cplx b = {0.5 * (zk.im + zn.im), -0.5 * (zk.re - zn.re)};
      A[k] = rmul(sx, a);
      B[k] = rmul(sx, b);
    }
    A[0].im = 0.0;
    B[0].im = 0.0;
  }
  for (int i = 0; i < mx; ++i) {
    fft_exec(&o->plz, hz + i, mx, line, -1);
    for (int k = 0; k < nz; ++k) {
      int kk = k < nz / 2 ? k : k + (pz - nz);
      sp[k * mx + i] = rmul(sz, line[kk]);
    }
  }
}

static size_t work_len(const lho* o) { return (size_t)o->pz * o->mx + 2 * (size_t)(o->px > o->pz ? o->px : o->pz) + 8; }

/* gather spectral plane j of a field in reference layout */
static void gather_plane(const lho* o, const cplx* f, int nj, int joff, cplx* sp) {
  for (int c = 0; c < o->ncol; ++c) sp[c] = f[(size_t)c * nj + joff];
}
static void scatter_plane(const lho* o, cplx* f, int nj, int joff, const cplx* sp) {
  for (int c = 0; c < o->ncol; ++c) f[(size_t)c * nj + joff] = sp[c];
}

/* ------------------------------------------------------------------------------------ */
/* Creation                                                                              */
/* ------------------------------------------------------------------------------------ */
static void lu_buildP(lho* o);

lho* lho_create(const lh_config* c) {
  lho* o = (lho*)calloc(1, sizeof(lho));
  o->nx = c->nx; o->nz = c->nz; o->px = c->px; o->pz = c->pz;
  o->mx = c->nx / 2 + 1; o->ncol = o->mx * c->nz;
  o->ny = c->ny; o->d = c->dsty;
  o->nyu = c->ny + 2 * c->dsty + 1; o->nyv = c->ny + 2 * c->dsty; o->nyp = c->ny + 2 * c->dsty - 1;
  o->re = c->re; o->alp = c->alp; o->bet = c->bet; o->ddthetavi = c->ddthetavi;
  o->flag_ctpress = c->flag_ctpress; o->mpgx = c->mpgx; o->qxt = c->qxt;
  o->yu = dupd(c->yu, o->nyu); o->dthdyu = dupd(c->dthdyu, o->nyu);
  o->yv = dupd(c->yv, o->nyv); o->dthdyv = dupd(c->dthdyv, o->nyv);
  o->dyu2i = dupd(c->dyu2i, 3 * (size_t)o->nyu); o->dyv2i = dupd(c->dyv2i, 3 * (size_t)o->nyv);
  o->gw[0] = c->gridweighting[0]; o->gw[1] = c->gridweighting[1];
  for (int k = 0; k < 3; ++k) {
    o->aRK[k] = c->rk[k]; o->bRK[k] = c->rk[3 + k]; o->gRK[k] = c->rk[6 + k];
    o->cRK[k] = c->rk[9 + k]; o->dRK[k] = c->rk[12 + k];
  }
  o->k1x = dupd(c->k1x, o->mx); o->k2x = dupd(c->k2x, o->mx);
  o->k1z = dupd(c->k1z, o->nz); o->k2z = dupd(c->k2z, o->nz);
  o->tip_u[0] = c->tip_u[0]; o->tip_u[1] = c->tip_u[1];
  o->tip_v[0] = c->tip_v[0]; o->tip_v[1] = c->tip_v[1];
  o->ntilex = c->ntilex; o->ntilez = c->ntilez;
  o->band_v[0] = c->band_v[0]; o->band_v[1] = c->band_v[1];
  size_t pp = (size_t)o->px * o->pz;
  o->mask_u = (unsigned char**)calloc(o->nyu, sizeof(unsigned char*));
  o->mask_v = (unsigned char**)calloc(o->nyv, sizeof(unsigned char*));
  for (int n = 0; n < c->nlist_u; ++n) {
    int i = c->list_u[3 * n], k = c->list_u[3 * n + 1], j = c->list_u[3 * n + 2];
    int jj = j + o->d;
    if (!o->mask_u[jj]) o->mask_u[jj] = (unsigned char*)calloc(pp, 1);
    o->mask_u[jj][(size_t)(k - 1) * o->px + (i - 1)] = 1;
  }
  for (int n = 0; n < c->nlist_v; ++n) {
    int i = c->list_v[3 * n], k = c->list_v[3 * n + 1], j = c->list_v[3 * n + 2];
    int jj = j + o->d;
    if (!o->mask_v[jj]) o->mask_v[jj] = (unsigned char*)calloc(pp, 1);
    o->mask_v[jj][(size_t)(k - 1) * o->px + (i - 1)] = 1;
  }
  size_t nu = (size_t)o->ncol * o->nyu, nv = (size_t)o->ncol * o->nyv, np = (size_t)o->ncol * o->nyp;
  o->u1 = calloc(nu, sizeof(cplx)); o->u3 = calloc(nu, sizeof(cplx)); o->u2 = calloc(nv, sizeof(cplx));
  o->p = calloc(np, sizeof(cplx)); o->psi = calloc(np, sizeof(cplx)); o->div = calloc(np, sizeof(cplx));
  o->du1 = calloc(nu, sizeof(cplx)); o->du3 = calloc(nu, sizeof(cplx)); o->du2 = calloc(nv, sizeof(cplx));
  o->Lu1 = calloc(nu, sizeof(cplx)); o->Lu3 = calloc(nu, sizeof(cplx)); o->Lu2 = calloc(nv, sizeof(cplx));
  o->Nu1 = calloc(nu, sizeof(cplx)); o->Nu3 = calloc(nu, sizeof(cplx)); o->Nu2 = calloc(nv, sizeof(cplx));
  o->DG = calloc(3 * np, sizeof(double));
  if (plan_init(&o->plx, o->px) || plan_init(&o->plz, o->pz) || plan_init(&o->plnx, o->nx) ||
      plan_init(&o->plnz, o->nz)) {
    return NULL;
  }
  o->st = calloc(13 * (size_t)o->nyu, sizeof(double));
  o->sp = calloc(5 * (size_t)o->nyu * o->ncol, sizeof(double));
  lu_buildP(o);
  return o;
}

void lho_destroy(lho* o) {
  if (!o) return;
  free(o->yu); free(o->yv); free(o->dthdyu); free(o->dthdyv); free(o->dyu2i); free(o->dyv2i);
  free(o->k1x); free(o->k2x); free(o->k1z); free(o->k2z);
  for (int j = 0; j < o->nyu; ++j) free(o->mask_u[j]);
  for (int j = 0; j < o->nyv; ++j) free(o->mask_v[j]);
  free(o->mask_u); free(o->mask_v);
  free(o->u1); free(o->u2); free(o->u3); free(o->p); free(o->psi); free(o->div);
  free(o->du1); free(o->du2); free(o->du3); free(o->Lu1); free(o->Lu2); free(o->Lu3);
  free(o->Nu1); free(o->Nu2); free(o->Nu3); free(o->F1); free(o->F2); free(o->F3); free(o->DG); free(o->st); free(o->sp); free(o->cst);
  plan_free(&o->plx); plan_free(&o->plz); plan_free(&o->plnx); plan_free(&o->plnz);
  free(o);
}

void lho_set_state(lho* o, const double* u1, const double* u2, const double* u3, const double* p) {
  size_t nu = (size_t)o->ncol * o->nyu, nv = (size_t)o->ncol * o->nyv, np = (size_t)o->ncol * o->nyp;
  if (u1) memcpy(o->u1, u1, nu * sizeof(cplx));
  if (u2) memcpy(o->u2, u2, nv * sizeof(cplx));
  if (u3) memcpy(o->u3, u3, nu * sizeof(cplx));
  if (p) memcpy(o->p, p, np * sizeof(cplx));
}
void lho_get_state(const lho* o, double* u1, double* u2, double* u3, double* p) {
  size_t nu = (size_t)o->ncol * o->nyu, nv = (size_t)o->ncol * o->nyv, np = (size_t)o->ncol * o->nyp;
  if (u1) memcpy(u1, o->u1, nu * sizeof(cplx));
  if (u2) memcpy(u2, o->u2, nv * sizeof(cplx));
  if (u3) memcpy(u3, o->u3, nu * sizeof(cplx));
  if (p) memcpy(p, o->p, np * sizeof(cplx));
}
/* Nu in the reference's allocation range jlim(1)+1 : jlim(2)-1 (littleharsh.f90:122-124) */
void lho_get_nonlinear(const lho* o, double* n1, double* n2, double* n3) {
  for (int c = 0; c < o->ncol; ++c) {
    for (int r = 0; r < o->nyu - 2; ++r) {
      ((cplx*)n1)[(size_t)c * (o->nyu - 2) + r] = o->Nu1[(size_t)c * o->nyu + r + 1];
      ((cplx*)n3)[(size_t)c * (o->nyu - 2) + r] = o->Nu3[(size_t)c * o->nyu + r + 1];
    }
    for (int r = 0; r < o->nyv - 2; ++r)
      ((cplx*)n2)[(size_t)c * (o->nyv - 2) + r] = o->Nu2[(size_t)c * o->nyv + r + 1];
  }
}
/* Forcing added to Nu_new in every substep, in the layout of lho_set_nonlinear; all three
   NULL removes it.  A NULL field among non-NULL ones is zero. */
void lho_set_forcing(lho* o, const double* f1, const double* f2, const double* f3) {
  free(o->F1); free(o->F2); free(o->F3);
  o->F1 = o->F2 = o->F3 = NULL;
  if (!f1 && !f2 && !f3) return;
  o->F1 = calloc((size_t)o->ncol * o->nyu, sizeof(cplx));
  o->F3 = calloc((size_t)o->ncol * o->nyu, sizeof(cplx));
  o->F2 = calloc((size_t)o->ncol * o->nyv, sizeof(cplx));
  for (int c = 0; c < o->ncol; ++c) {
    for (int r = 0; r < o->nyu - 2; ++r) {
      if (f1) o->F1[(size_t)c * o->nyu + r + 1] = ((const cplx*)f1)[(size_t)c * (o->nyu - 2) + r];
      if (f3) o->F3[(size_t)c * o->nyu + r + 1] = ((const cplx*)f3)[(size_t)c * (o->nyu - 2) + r];
    }
    if (f2)
      for (int r = 0; r < o->nyv - 2; ++r)
        o->F2[(size_t)c * o->nyv + r + 1] = ((const cplx*)f2)[(size_t)c * (o->nyv - 2) + r];
  }
}
void lho_set_nonlinear(lho* o, const double* n1, const double* n2, const double* n3) {
  for (int c = 0; c < o->ncol; ++c) {
    for (int r = 0; r < o->nyu - 2; ++r) {
      o->Nu1[(size_t)c * o->nyu + r + 1] = ((const cplx*)n1)[(size_t)c * (o->nyu - 2) + r];
      o->Nu3[(size_t)c * o->nyu + r + 1] = ((const cplx*)n3)[(size_t)c * (o->nyu - 2) + r];
    }
    for (int r = 0; r < o->nyv - 2; ++r)
      o->Nu2[(size_t)c * o->nyv + r + 1] = ((const cplx*)n2)[(size_t)c * (o->nyv - 2) + r];
  }
}
void lho_get_scalars(const lho* o, double* mpgx, double* qx, double* cflmax) {
  if (mpgx) *mpgx = o->mpgx;
  if (qx) *qx = o->qx;
  if (cflmax) *cflmax = o->cflmax;
}

/* ------------------------------------------------------------------------------------ */
/* laplacian_U / laplacian_V, littleharsh.f90:247-321                                    */
/* RHS0_u1/u2/u3, littleharsh.f90:614-731                                                */
/* ------------------------------------------------------------------------------------ */
static void rhs0(lho* o, int krk) {
  const int d = o->d, nyu = o->nyu, nyv = o->nyv;
  const double C1 = o->aRK[krk] / o->re, C2 = -o->cRK[krk], dR = o->dRK[krk];
#pragma omp parallel for schedule(static)
  for (int c = 0; c < o->ncol; ++c) {
    const int i = c % o->mx, k = c / o->mx;
    const double k2x = o->k2x[i], k2z = o->k2z[k];
    const double kx = o->k1x[i], kz = o->k1z[k];
    const double C3x = -dR * kx, C3z = -dR * kz;   /* -dRK*k1F: complex (0, -dRK*k) */
    for (int j = -d + 1; j <= o->ny + d - 1; ++j) {
      const double* w = o->dyu2i + 3 * (size_t)UJ(o, j);
      /* u1 */
      cplx um = FU(o, o->u1, c, j - 1), u0 = FU(o, o->u1, c, j), up = FU(o, o->u1, c, j + 1);
      cplx Lu = cadd(cadd(rmul(w[0], um), rmul(k2x + k2z + w[1], u0)), rmul(w[2], up));
      cplx Luy = cadd(cadd(rmul(w[0], um), rmul(w[1], u0)), rmul(w[2], up));
      FU(o, o->Lu1, c, j) = Luy;
      FU(o, o->du1, c, j) = cadd(cadd(rmul(C1, Lu), rmul(C2, FU(o, o->Nu1, c, j))),
                                 imul(C3x, FP(o, o->p, c, j)));
      /* u3 */
      um = FU(o, o->u3, c, j - 1); u0 = FU(o, o->u3, c, j); up = FU(o, o->u3, c, j + 1);
      Lu = cadd(cadd(rmul(w[0], um), rmul(k2x + k2z + w[1], u0)), rmul(w[2], up));
      Luy = cadd(cadd(rmul(w[0], um), rmul(w[1], u0)), rmul(w[2], up));
      FU(o, o->Lu3, c, j) = Luy;
      FU(o, o->du3, c, j) = cadd(cadd(rmul(C1, Lu), rmul(C2, FU(o, o->Nu3, c, j))),
                                 imul(C3z, FP(o, o->p, c, j)));
    }
    for (int j = -d + 1; j <= o->ny + d - 2; ++j) {
      const double* w = o->dyv2i + 3 * (size_t)UJ(o, j);
      cplx um = FV(o, o->u2, c, j - 1), u0 = FV(o, o->u2, c, j), up = FV(o, o->u2, c, j + 1);
      cplx Lu = cadd(cadd(rmul(w[0], um), rmul(k2x + k2z + w[1], u0)), rmul(w[2], up));
      cplx Luy = cadd(cadd(rmul(w[0], um), rmul(w[1], u0)), rmul(w[2], up));
      FV(o, o->Lu2, c, j) = Luy;
      const double C3 = -dR * o->ddthetavi * o->dthdyv[UJ(o, j)];
      FV(o, o->du2, c, j) = cadd(cadd(rmul(C1, Lu), rmul(C2, FV(o, o->Nu2, c, j))),
                                 rmul(C3, csub(FP(o, o->p, c, j + 1), FP(o, o->p, c, j))));
    }
  }
  (void)nyu; (void)nyv;
  /* mean pressure gradient on mode (0,0): littleharsh.f90:651-657 */
  const double C1m = -dR * o->mpgx;
  for (int j = -d + 1; j <= o->ny - 1 + d; ++j) FU(o, o->du1, 0, j).re += C1m;
}

/* ------------------------------------------------------------------------------------ */
/* nonlinear (restatement; FOU3D.f90 absent).  Divergence form on the staggered grid:     */
/*   Nu1 = d/dx(uu) + d/dz(uw) + dy(uv_f)   at u centres                                */
/*   Nu2 = d/dx(uv_f) + d/dz(vw_f) + dy(vv_c) at v faces                                */
/*   Nu3 = d/dx(uw) + d/dz(ww) + dy(vw_f)   at u centres                                */
/* products of the 3/2 zero-padded physical fields, v interpolated to centres and u, w    */
/* to faces by the theta-midpoint average, wall-face products zero; du -= gRK*Nu_new.    */
/* During substep 1 the CFL maximum max(|u|/dx + |v_c| dthdyu + |w|/dz) is recorded.    */
/* ------------------------------------------------------------------------------------ */
static void nonlinear(lho* o, int krk) {
  const int d = o->d, ny = o->ny, nyu = o->nyu, nyv = o->nyv, ncol = o->ncol;
  const size_t pp = (size_t)o->px * o->pz;
  double* U = (double*)calloc(pp * nyu, sizeof(double));
  double* W = (double*)calloc(pp * nyu, sizeof(double));
  double* V = (double*)calloc(pp * nyv, sizeof(double));
  cplx* UU = calloc((size_t)ncol * nyu, sizeof(cplx));
  cplx* UW = calloc((size_t)ncol * nyu, sizeof(cplx));
  cplx* WW = calloc((size_t)ncol * nyu, sizeof(cplx));
  cplx* VV = calloc((size_t)ncol * nyu, sizeof(cplx));
  cplx* UV = calloc((size_t)ncol * nyv, sizeof(cplx));
  cplx* VW = calloc((size_t)ncol * nyv, sizeof(cplx));
  const size_t wl = work_len(o);
  /* physical velocities: u, w at interior centres, v at interior faces (walls are 0) */
#pragma omp parallel
  {
    cplx* work = malloc(wl * sizeof(cplx));
    cplx* sp = malloc(ncol * sizeof(cplx));
#pragma omp for schedule(dynamic)
    for (int j = -d + 1; j <= ny + d - 1; ++j) {
      gather_plane(o, o->u1, nyu, UJ(o, j), sp);
      four_to_phys(o, sp, U + pp * UJ(o, j), work);
      gather_plane(o, o->u3, nyu, UJ(o, j), sp);
      four_to_phys(o, sp, W + pp * UJ(o, j), work);
      if (j <= ny + d - 2) {
        gather_plane(o, o->u2, nyv, UJ(o, j), sp);
        four_to_phys(o, sp, V + pp * UJ(o, j), work);
      }
    }
    free(work); free(sp);
  }
  /* CFL estimate at centres (substep 1 only) */
  if (krk == 0) {
    double mx = 0.0;
    const double dxi = o->px * o->alp / (2.0 * M_PI), dzi = o->pz * o->bet / (2.0 * M_PI);
    for (int j = -d + 1; j <= ny + d - 1; ++j) {
      const double* u = U + pp * UJ(o, j);
      const double* w = W + pp * UJ(o, j);
      const double* vb = V + pp * UJ(o, j - 1);
      const double* vt = V + pp * UJ(o, j);
      const double dyi = o->dthdyu[UJ(o, j)];
      for (size_t n = 0; n < pp; ++n) {
        double vc = 0.5 * (vb[n] + vt[n]);
        double s = fabs(u[n]) * dxi + fabs(vc) * dyi + fabs(w[n]) * dzi;
        if (s > mx) mx = s;
      }
    }
    o->cflmax = mx;
  }
  /* products and forward transforms */
#pragma omp parallel
  {
    cplx* work = malloc(wl * sizeof(cplx));
    cplx* sp = malloc(ncol * sizeof(cplx));
    double* ph = malloc(pp * sizeof(double));
#pragma omp for schedule(dynamic)
    for (int j = -d + 1; j <= ny + d - 1; ++j) {
      const double* u = U + pp * UJ(o, j);
      const double* w = W + pp * UJ(o, j);
      const double* vb = V + pp * UJ(o, j - 1);
      const double* vt = V + pp * UJ(o, j);
      for (size_t n = 0; n < pp; ++n) ph[n] = u[n] * u[n];
      phys_to_four(o, ph, sp, work); scatter_plane(o, UU, nyu, UJ(o, j), sp);
      for (size_t n = 0; n < pp; ++n) ph[n] = u[n] * w[n];
      phys_to_four(o, ph, sp, work); scatter_plane(o, UW, nyu, UJ(o, j), sp);
      for (size_t n = 0; n < pp; ++n) ph[n] = w[n] * w[n];
      phys_to_four(o, ph, sp, work); scatter_plane(o, WW, nyu, UJ(o, j), sp);
      for (size_t n = 0; n < pp; ++n) { double vc = 0.5 * (vb[n] + vt[n]); ph[n] = vc * vc; }
      phys_to_four(o, ph, sp, work); scatter_plane(o, VV, nyu, UJ(o, j), sp);
      if (j <= ny + d - 2) {   /* interior face j between centres j and j+1 */
        const double* ua = U + pp * UJ(o, j + 1);
        const double* wa = W + pp * UJ(o, j + 1);
        for (size_t n = 0; n < pp; ++n) { double uf = 0.5 * (u[n] + ua[n]); ph[n] = uf * vt[n]; }
        phys_to_four(o, ph, sp, work); scatter_plane(o, UV, nyv, UJ(o, j), sp);
        for (size_t n = 0; n < pp; ++n) { double wf = 0.5 * (w[n] + wa[n]); ph[n] = vt[n] * wf; }
        phys_to_four(o, ph, sp, work); scatter_plane(o, VW, nyv, UJ(o, j), sp);
      }
    }
    free(work); free(sp); free(ph);
  }
  /* assemble Nu_new (plus the optional forcing), du -= gRK*Nu_new, keep Nu_new for the next
     substep */
  const double g = o->gRK[krk], dv = o->ddthetavi;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < ncol; ++c) {
    const int i = c % o->mx, k = c / o->mx;
    const double kx = o->k1x[i], kz = o->k1z[k];
    for (int j = -d + 1; j <= ny + d - 1; ++j) {
      const double m = dv * o->dthdyu[UJ(o, j)];
      cplx n1 = cadd(cadd(imul(kx, FU(o, UU, c, j)), imul(kz, FU(o, UW, c, j))),
                     rmul(m, csub(FV(o, UV, c, j), FV(o, UV, c, j - 1))));
      cplx n3 = cadd(cadd(imul(kx, FU(o, UW, c, j)), imul(kz, FU(o, WW, c, j))),
                     rmul(m, csub(FV(o, VW, c, j), FV(o, VW, c, j - 1))));
      if (o->F1) { n1 = cadd(n1, FU(o, o->F1, c, j)); n3 = cadd(n3, FU(o, o->F3, c, j)); }
      FU(o, o->Nu1, c, j) = n1;
      FU(o, o->Nu3, c, j) = n3;
      FU(o, o->du1, c, j) = csub(FU(o, o->du1, c, j), rmul(g, n1));
      FU(o, o->du3, c, j) = csub(FU(o, o->du3, c, j), rmul(g, n3));
    }
    for (int j = -d + 1; j <= ny + d - 2; ++j) {
      const double m = dv * o->dthdyv[UJ(o, j)];
      cplx n2 = cadd(cadd(imul(kx, FV(o, UV, c, j)), imul(kz, FV(o, VW, c, j))),
                     rmul(m, csub(FU(o, VV, c, j + 1), FU(o, VV, c, j))));
      if (o->F2) n2 = cadd(n2, FV(o, o->F2, c, j));
      FV(o, o->Nu2, c, j) = n2;
      FV(o, o->du2, c, j) = csub(FV(o, o->du2, c, j), rmul(g, n2));
    }
  }
  free(U); free(W); free(V); free(UU); free(UW); free(WW); free(VV); free(UV); free(VW);
}

/* ------------------------------------------------------------------------------------ */
/* LU_build + LU_dec for the velocity grids (tridLU_3D_IB.f90:427-493, 601-640).  The    */
/* matrix is shared by every column; axz(c) = 1 - beta*(k2x+k2z) scales the rhs.         */
/* ay layout: ay[3*r + m], r = row 0..nj-1.                                              */
/* ------------------------------------------------------------------------------------ */
static void lu_build_dec(const lho* o, int grid_u, double beta, double* ay) {
  const int nj = grid_u ? o->nyu : o->nyv;
  const double* dd = grid_u ? o->dyu2i : o->dyv2i;
  ay[0] = 0.0; ay[1] = 1.0; ay[2] = grid_u ? o->gw[0] : 0.0;
  for (int r = 1; r < nj - 1; ++r) {
    ay[3 * r + 0] = -beta * dd[3 * r + 0];
    ay[3 * r + 1] = 1.0 - beta * (dd[3 * r + 1]);
    ay[3 * r + 2] = -beta * dd[3 * r + 2];
  }
  const int r = nj - 1;
  ay[3 * r + 0] = grid_u ? o->gw[1] : 0.0; ay[3 * r + 1] = 1.0; ay[3 * r + 2] = 0.0;
  /* LU_dec */
  ay[1] = 1.0 / ay[1];
  for (int q = 1; q < nj; ++q) {
    ay[3 * (q - 1) + 2] = ay[3 * (q - 1) + 2] * ay[3 * (q - 1) + 1];
    ay[3 * q + 1] = 1.0 / (ay[3 * q + 1] - ay[3 * q + 0] * ay[3 * (q - 1) + 2]);
  }
}

/* immersed_boundaries_U/V (tridLU_3D_IB.f90:198-359): on every plane that holds canopy  */
/* points, rhs -> physical, rhs = 0 at listed points, rhs = -beta*Lu at listed points of */
/* the tip planes, rhs -> Fourier.                                                       */
static void immersed_boundaries(lho* o, cplx* x, const cplx* Lux, int grid_u, double beta) {
  const int nj = grid_u ? o->nyu : o->nyv;
  unsigned char** mask = grid_u ? o->mask_u : o->mask_v;
  const int* tip = grid_u ? o->tip_u : o->tip_v;
  const size_t pp = (size_t)o->px * o->pz;
  const size_t wl = work_len(o);
#pragma omp parallel
  {
    cplx* work = malloc(wl * sizeof(cplx));
    cplx* sp = malloc(o->ncol * sizeof(cplx));
    double* ph = malloc(pp * sizeof(double));
    double* pl = malloc(pp * sizeof(double));
#pragma omp for schedule(dynamic)
    for (int jj = 0; jj < nj; ++jj) {
      if (!mask[jj]) continue;
      const int j = jj - o->d;
      const int is_tip = (j == tip[0] || j == tip[1]);
      gather_plane(o, x, nj, jj, sp);
      four_to_phys(o, sp, ph, work);
      if (is_tip) {
        gather_plane(o, Lux, nj, jj, sp);
        four_to_phys(o, sp, pl, work);
      }
      for (size_t n = 0; n < pp; ++n)
        if (mask[jj][n]) ph[n] = is_tip ? -beta * pl[n] : 0.0;
      phys_to_four(o, ph, sp, work);
      scatter_plane(o, x, nj, jj, sp);
    }
    free(work); free(sp); free(ph); free(pl);
  }
}

/* solveU / solveV (littleharsh.f90:733-824) with LUsolU / LUsolV (tridLU_3D_IB.f90:50-194) */
static void solve_vel(lho* o, cplx* u, cplx* du, const cplx* Lu, int grid_u, int krk, double dt) {
  const int nj = grid_u ? o->nyu : o->nyv;
  const double beta = o->bRK[krk] * dt / o->re;
  double* ay = malloc(3 * (size_t)nj * sizeof(double));
  lu_build_dec(o, grid_u, beta, ay);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < o->ncol; ++c) {
    cplx* x = du + (size_t)c * nj;
    const cplx* uc = u + (size_t)c * nj;
    x[0] = (cplx){0.0, 0.0};
    for (int r = 1; r < nj - 1; ++r) x[r] = cadd(uc[r], rmul(dt, x[r]));
    x[nj - 1] = (cplx){0.0, 0.0};
    const int i = c % o->mx, k = c / o->mx;
    const double axz = 1.0 - beta * (o->k2x[i] + o->k2z[k]);
    for (int r = 1; r < nj - 1; ++r) { x[r].re /= axz; x[r].im /= axz; }
  }
  immersed_boundaries(o, du, Lu, grid_u, beta);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < o->ncol; ++c) {
    cplx* x = du + (size_t)c * nj;
    x[0] = rmul(ay[1], x[0]);
    for (int r = 1; r < nj; ++r) x[r] = rmul(ay[3 * r + 1], csub(x[r], rmul(ay[3 * r], x[r - 1])));
    for (int r = nj - 2; r >= 0; --r) x[r] = csub(x[r], rmul(ay[3 * r + 2], x[r + 1]));
    memcpy(u + (size_t)c * nj, x, nj * sizeof(cplx));
  }
  free(ay);
}

/* flowrateIm / flowrateRe (littleharsh.f90:412-451): between the tips, j = 1..nn+1 */
static double flowrate_col(const lho* o, const cplx* ucol) {
  double q = 0.0;
  for (int j = 1; j <= o->ny - 1; ++j) q += (o->yv[UJ(o, j)] - o->yv[UJ(o, j - 1)]) * ucol[UJ(o, j)].re;
  return q;
}

/* flowrate_corr + LUsol0 (littleharsh.f90:453-485, tridLU_3D_IB.f90:398-424, 573-599) */
static void meanflow_ctU(lho* o, int krk, double dt) {
  const int n = o->ny;        /* rows 0..nn+2 = 0..ny */
  double* u11 = malloc((n + 1) * sizeof(double));
  double* a = malloc(3 * (size_t)(n + 1) * sizeof(double));
  const double bdt = -o->dRK[krk] * dt;
  u11[0] = 0.0;
  for (int j = 1; j <= n - 1; ++j) u11[j] = bdt;
  u11[n] = 0.0;
  const double beta = o->bRK[krk] * dt / o->re;
  a[0] = 0.0; a[1] = 1.0; a[2] = o->gw[0];
  for (int j = 1; j <= n - 1; ++j) {
    const double* w = o->dyu2i + 3 * (size_t)UJ(o, j);
    a[3 * j] = -beta * w[0]; a[3 * j + 1] = 1.0 - beta * w[1]; a[3 * j + 2] = -beta * w[2];
  }
  a[3 * n] = o->gw[1]; a[3 * n + 1] = 1.0; a[3 * n + 2] = 0.0;
  a[1] = 1.0 / a[1];
  for (int j = 1; j <= n; ++j) {
    a[3 * (j - 1) + 2] = a[3 * (j - 1) + 2] * a[3 * (j - 1) + 1];
    a[3 * j + 1] = 1.0 / (a[3 * j + 1] - a[3 * j] * a[3 * (j - 1) + 2]);
  }
  u11[0] = u11[0] * a[1];
  for (int j = 1; j <= n; ++j) u11[j] = (u11[j] - a[3 * j] * u11[j - 1]) * a[3 * j + 1];
  for (int j = n - 1; j >= 0; --j) u11[j] = u11[j] - a[3 * j + 2] * u11[j + 1];
  /* flowrateRe(Qx1, u11) declares its dummy u(1:nn+2) while u11 is (0:nn+2)
     (littleharsh.f90:443 vs 472-474): inside it u(j) is u11(j-1) */
  double qx1 = 0.0;
  for (int j = 1; j <= n - 1; ++j) qx1 += (o->yv[UJ(o, j)] - o->yv[UJ(o, j - 1)]) * u11[j - 1];
  const double qx0 = flowrate_col(o, o->u1);
  const double g = (o->qxt - qx0) / qx1;
  o->mpgx = o->mpgx + g;
  o->dgx = g;
  for (int j = 0; j <= n; ++j) o->u1[UJ(o, j)].re += g * u11[j];
  free(u11); free(a);
}

/* divergence (littleharsh.f90:210-245) over the p grid */
static void divergence(const lho* o, cplx* div) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < o->ncol; ++c) {
    const int i = c % o->mx, k = c / o->mx;
    const double kx = o->k1x[i], kz = o->k1z[k];
    for (int j = -o->d + 1; j <= o->ny + o->d - 1; ++j)
      FP(o, div, c, j) = cadd(cadd(imul(kx, FU(o, o->u1, c, j)), imul(kz, FU(o, o->u3, c, j))),
                              rmul(o->ddthetavi * o->dthdyu[UJ(o, j)],
                                   csub(FV(o, o->u2, c, j), FV(o, o->u2, c, j - 1))));
  }
}

/* LU_buildP + LU_decP (tridLU_3D_IB.f90:497-571, 643-684), once at start */
static void lu_buildP(lho* o) {
  const int ns = -o->d + 1, ne = o->ny + o->d - 1, np = o->nyp;
  const double dv = o->ddthetavi;
  /* D_vec_b(j-1) = -dv*dthdyu(j); G_vec_b(j) = -dv*dthdyv(j); D_vec_t(j) = dv*dthdyu(j);
     G_vec_t(j+1) = dv*dthdyv(j), for j = ns..ne */
  double* Db = malloc((np + 2) * sizeof(double));
  double* Gb = malloc((np + 2) * sizeof(double));
  double* Dt = malloc((np + 2) * sizeof(double));
  double* Gt = malloc((np + 2) * sizeof(double));
#define IX(j) ((j) - (ns - 1))
  for (int j = ns; j <= ne; ++j) {
    Db[IX(j - 1)] = -dv * o->dthdyu[UJ(o, j)];
    Gb[IX(j)] = -dv * o->dthdyv[UJ(o, j)];
    Dt[IX(j)] = dv * o->dthdyu[UJ(o, j)];
    Gt[IX(j + 1)] = dv * o->dthdyv[UJ(o, j)];
  }
  for (int c = 0; c < o->ncol; ++c) {
    const int i = c % o->mx, k = c / o->mx;
    const double k2x = o->k2x[i], k2z = o->k2z[k];
    double* a = o->DG + (size_t)c * 3 * np;
#define A(m, j) a[3 * ((j) - ns) + (m)]
    A(1, ns) = (Dt[IX(ns)] * Gb[IX(ns)]) + k2x + k2z;
    A(2, ns) = (Dt[IX(ns)] * Gt[IX(ns + 1)]);
    A(1, ne) = (Db[IX(ne - 1)] * Gt[IX(ne)]) + k2x + k2z;
    A(0, ne) = (Db[IX(ne - 1)] * Gb[IX(ne - 1)]);
    for (int j = ns + 1; j <= ne - 1; ++j) {
      A(1, j) = (Db[IX(j - 1)] * Gt[IX(j)]) + (Dt[IX(j)] * Gb[IX(j)]) + k2x + k2z;
      A(0, j) = (Db[IX(j - 1)] * Gb[IX(j - 1)]);
      A(2, j) = (Dt[IX(j)] * Gt[IX(j + 1)]);
    }
    if (c == 0) {   /* tridLU_3D_IB.f90:548-554: the (0,0) pressure pin */
      const double dy = o->yu[UJ(o, 2)] - o->yu[UJ(o, 1)];
      A(0, ns) = 0.0;
      A(1, ns) = 1.0 / (dy * dy);
      A(2, ns) = 0.0;
    }
    /* LU_decP */
    A(1, ns) = 1.0 / A(1, ns);
    for (int j = ns + 1; j <= ne; ++j) {
      A(2, j - 1) = A(2, j - 1) * A(1, j - 1);
      A(1, j) = 1.0 / (A(1, j) - A(0, j) * A(2, j - 1));
    }
#undef A
  }
#undef IX
  free(Db); free(Gb); free(Dt); free(Gt);
}

/* solveP (littleharsh.f90:557-612) with LUsolP (tridLU_3D_IB.f90:362-396) */
static void solveP(lho* o, int krk, double dt) {
  const double dti = 1.0 / dt;
  const double dtirk = dti / (o->aRK[krk] + o->bRK[krk]);
  divergence(o, o->div);
  const int np = o->nyp;
  for (size_t n = 0; n < (size_t)o->ncol * np; ++n) o->psi[n] = rmul(dtirk, o->div[n]);
  o->psi[0] = (cplx){0.0, 0.0};
#pragma omp parallel for schedule(static)
  for (int c = 0; c < o->ncol; ++c) {
    cplx* x = o->psi + (size_t)c * np;
    const double* a = o->DG + (size_t)c * 3 * np;
    x[0] = rmul(a[1], x[0]);
    for (int r = 1; r < np; ++r) x[r] = rmul(a[3 * r + 1], csub(x[r], rmul(a[3 * r], x[r - 1])));
    for (int r = np - 2; r >= 0; --r) x[r] = csub(x[r], rmul(a[3 * r + 2], x[r + 1]));
    cplx* pc = o->p + (size_t)c * np;
    for (int r = 0; r < np; ++r) pc[r] = cadd(pc[r], x[r]);
  }
}

/* v_corr (littleharsh.f90:826-937) */
static void v_corr(lho* o, int krk, double dt) {
  const double dtrk = dt * (o->aRK[krk] + o->bRK[krk]);
  const int d = o->d, ny = o->ny;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < o->ncol; ++c) {
    const int i = c % o->mx, k = c / o->mx;
    const double kx = dtrk * o->k1x[i], kz = dtrk * o->k1z[k];
    for (int j = -d + 1; j <= ny + d - 1; ++j) {
      FU(o, o->u1, c, j) = csub(FU(o, o->u1, c, j), imul(kx, FP(o, o->psi, c, j)));
      FU(o, o->u3, c, j) = csub(FU(o, o->u3, c, j), imul(kz, FP(o, o->psi, c, j)));
    }
    FU(o, o->u1, c, -d) = rmul(-o->gw[0], FU(o, o->u1, c, -d + 1));
    FU(o, o->u1, c, ny + d) = rmul(-o->gw[1], FU(o, o->u1, c, ny + d - 1));
    FU(o, o->u3, c, -d) = rmul(-o->gw[0], FU(o, o->u3, c, -d + 1));
    FU(o, o->u3, c, ny + d) = rmul(-o->gw[1], FU(o, o->u3, c, ny + d - 1));
    const double dtrk_u2 = dtrk * o->ddthetavi;
    for (int j = -d + 1; j <= ny + d - 2; ++j)
      FV(o, o->u2, c, j) = csub(FV(o, o->u2, c, j),
                                rmul(dtrk_u2 * o->dthdyv[UJ(o, j)],
                                     csub(FP(o, o->psi, c, j + 1), FP(o, o->psi, c, j))));
    FV(o, o->u2, c, -d) = (cplx){0.0, 0.0};
    FV(o, o->u2, c, ny + d - 1) = (cplx){0.0, 0.0};
  }
  divergence(o, o->div);
}

/* One substep, littleharsh.f90:166-188; krk is 1-based */
int lho_substep(lho* o, int krk1, double dt) {
  const int krk = krk1 - 1;
  if (krk < 0 || krk > 2 || !(dt > 0.0)) return LH_ERR_ARG;
  rhs0(o, krk);
  nonlinear(o, krk);
  solve_vel(o, o->u1, o->du1, o->Lu1, 1, krk, dt);
  solve_vel(o, o->u2, o->du2, o->Lu2, 0, krk, dt);
  solve_vel(o, o->u3, o->du3, o->Lu3, 1, krk, dt);
  if (o->flag_ctpress == 0) meanflow_ctU(o, krk, dt);
  solveP(o, krk, dt);
  v_corr(o, krk, dt);
  return LH_OK;
}

/* Full step, littleharsh.f90:164-197 */
int lho_step(lho* o, double dt) {
  for (int k = 1; k <= 3; ++k) {
    int r = lho_substep(o, k, dt);
    if (r) return r;
  }
  if (o->flag_ctpress == 1) o->qx = flowrate_col(o, o->u1);   /* meanflow_ctP */
  return LH_OK;
}

double lho_flowrate(const lho* o) { return flowrate_col(o, o->u1); }

/* error (littleharsh.f90:323-352): max |div| over the p grid */
double lho_max_div(lho* o) {
  divergence(o, o->div);
  double e = 0.0;
  for (size_t n = 0; n < (size_t)o->ncol * o->nyp; ++n) {
    double a = hypot(o->div[n].re, o->div[n].im);
    if (a > e) e = a;
  }
  return e;
}

/* transform helpers exposed for the FFT checks */
void lho_four_to_phys(lho* o, const double* sp, double* ph) {
  cplx* work = malloc(work_len(o) * sizeof(cplx));
  four_to_phys(o, (const cplx*)sp, ph, work);
  free(work);
}
void lho_phys_to_four(lho* o, const double* ph, double* sp) {
  cplx* work = malloc(work_len(o) * sizeof(cplx));
  phys_to_four(o, ph, (cplx*)sp, work);
  free(work);
}
/* plain complex FFT of length n (checks against the reference FFTPACK) */
void lho_cfft(int n, const double* in, double* out, int sign) {
  fftplan pl;
  plan_init(&pl, n);
  fft_exec(&pl, (const cplx*)in, 1, (cplx*)out, sign);
  plan_free(&pl);
}

/* ------------------------------------------------------------------------------------ */
/* Routine-level hooks: used by the checks against vectors produced by the reference's   */
/* own routines (tests/golden).                                                          */
/* ------------------------------------------------------------------------------------ */
/* RHS0_u1/u2/u3 alone: du and the y-Laplacians Lu for the current state and Nu */
void lho_hook_rhs0(lho* o, int krk1) { rhs0(o, krk1 - 1); }

/* which: 0..2 du1..3, 3..5 Lu1..3, 6 psi, 7 div (reference layout, full plane ranges) */
void lho_get_work(const lho* o, int which, double* out) {
  const cplx* src[8] = {o->du1, o->du2, o->du3, o->Lu1, o->Lu2, o->Lu3, o->psi, o->div};
  const int nj[8] = {o->nyu, o->nyv, o->nyu, o->nyu, o->nyv, o->nyu, o->nyp, o->nyp};
  memcpy(out, src[which], sizeof(cplx) * (size_t)o->ncol * nj[which]);
}
void lho_set_nonlinear_full(lho* o, const double* n1, const double* n2, const double* n3) {
  memcpy(o->Nu1, n1, sizeof(cplx) * (size_t)o->ncol * o->nyu);
  memcpy(o->Nu2, n2, sizeof(cplx) * (size_t)o->ncol * o->nyv);
  memcpy(o->Nu3, n3, sizeof(cplx) * (size_t)o->ncol * o->nyu);
}
/* solveP followed by v_corr on the current state */
void lho_hook_solvep_vcorr(lho* o, int krk1, double dt) {
  solveP(o, krk1 - 1, dt);
  v_corr(o, krk1 - 1, dt);
}
/* the substep split where the GPU's stage hooks split it (littleharsh.f90:166-188):
 * the velocity part (RHS0, nonlinear, solveU/V/U) and the pressure part (meanflow_ctU when
 * flag_ctpress = 0, solveP, v_corr) */
void lho_hook_substep_vel(lho* o, int krk1, double dt) {
  const int krk = krk1 - 1;
  rhs0(o, krk);
  nonlinear(o, krk);
  solve_vel(o, o->u1, o->du1, o->Lu1, 1, krk, dt);
  solve_vel(o, o->u2, o->du2, o->Lu2, 0, krk, dt);
  solve_vel(o, o->u3, o->du3, o->Lu3, 1, krk, dt);
}
void lho_hook_substep_pres(lho* o, int krk1, double dt) {
  const int krk = krk1 - 1;
  if (o->flag_ctpress == 0) meanflow_ctU(o, krk, dt);
  solveP(o, krk, dt);
  v_corr(o, krk, dt);
}
/* LU_build + LU_dec of a velocity grid: ay (3 x rows) and axz per column */
void lho_hook_lu_vel(lho* o, int grid_u, int krk1, double dt, double* ay, double* axz) {
  const double beta = o->bRK[krk1 - 1] * dt / o->re;
  lu_build_dec(o, grid_u, beta, ay);
  for (int c = 0; c < o->ncol; ++c)
    axz[c] = 1.0 - beta * (o->k2x[c % o->mx] + o->k2z[c / o->mx]);
}
void lho_get_dg(const lho* o, double* out) {
  memcpy(out, o->DG, sizeof(double) * 3 * (size_t)o->nyp * o->ncol);
}
/* flowrate_corr on column 0 of u1 with the given mpgx and QxT; returns g */
double lho_hook_flowrate_corr(lho* o, int krk1, double dt, double qxt, double* mpgx) {
  o->qxt = qxt;
  o->mpgx = *mpgx;
  meanflow_ctU(o, krk1 - 1, dt);
  *mpgx = o->mpgx;
  return o->dgx;
}

/* ------------------------------------------------------------------------------------ */
/* spectra (spectra.f90:1-56): sp(j, column) += |f|^2 for u, v, w, p, and                */
/* Re(u) Re(v_c) + Im(u) Im(v_c) with v averaged to the u centres (u2_itp2)              */
/* Layout: 5 arrays of ncol x nyu (p in its first nyp rows, v in its first nyv rows).   */
/* ------------------------------------------------------------------------------------ */
void lho_spectra_accumulate(lho* o) {
  const int nu = o->nyu;
  double* spU = o->sp;
  double* spV = o->sp + (size_t)o->ncol * nu;
  double* spW = o->sp + 2 * (size_t)o->ncol * nu;
  double* spUV = o->sp + 3 * (size_t)o->ncol * nu;
  double* spP = o->sp + 4 * (size_t)o->ncol * nu;
  for (int c = 0; c < o->ncol; ++c) {
    for (int r = 0; r < o->nyu; ++r) {
      cplx a = o->u1[(size_t)c * nu + r], w = o->u3[(size_t)c * nu + r];
      spU[(size_t)c * nu + r] += a.re * a.re + a.im * a.im;
      spW[(size_t)c * nu + r] += w.re * w.re + w.im * w.im;
      /* v at the u centre r: average of the faces r-1 and r (interior centres) */
      cplx vc = {0.0, 0.0};
      if (r >= 1 && r <= o->nyu - 2) {
        cplx b0 = o->u2[(size_t)c * o->nyv + r - 1], b1 = o->u2[(size_t)c * o->nyv + r];
        vc.re = 0.5 * (b0.re + b1.re);
        vc.im = 0.5 * (b0.im + b1.im);
      }
      spUV[(size_t)c * nu + r] += a.re * vc.re + a.im * vc.im;
    }
    for (int r = 0; r < o->nyv; ++r) {
      cplx v = o->u2[(size_t)c * o->nyv + r];
      spV[(size_t)c * nu + r] += v.re * v.re + v.im * v.im;
    }
    for (int r = 0; r < o->nyp; ++r) {
      cplx p = o->p[(size_t)c * o->nyp + r];
      spP[(size_t)c * nu + r] += p.re * p.re + p.im * p.im;
    }
  }
}
void lho_spectra_read(const lho* o, double* out) {
  memcpy(out, o->sp, sizeof(double) * 5 * (size_t)o->ncol * o->nyu);
}

/* ------------------------------------------------------------------------------------ */
/* statistics (stats.f90:1-108): per-plane sums over the physical collocation points of  */
/*   0 U, 1 U2, 2 V, 3 V2, 4 W, 5 W2, 6 P, 7 P2, 8 UV, 9 UW, 10 VW, 11 wx, 12 wx2          */
/* (the order of write_stats, stats.f90:337-349).  Rows 0,1,4,5,8,9 live on the u grid,  */
/* 2,3,10,11,12 on the v grid, 6,7 on the p grid; row q, slot s at st[q*nyu + s].        */
/* Fields: u, v, w, p; v_itp = v averaged to the u centres (0 on the two ghost planes);  */
/* w_itp = w averaged to the v faces; wx = dw/dy - dv/dz at the v faces.  The producer   */
/* of these planes (the nonlinear routine) is absent from the reference, so the          */
/* interpolations and wx follow the stencils of v_corr (littleharsh.f90:880-893) on the  */
/* collocation grid.  The plain wx rows keep the loop bounds of stats.f90:54-56: the     */
/* points x < nx, z < nz of that px x pz plane; every other row sums all px x pz points. */
/* Conditional statistics (conddown1/2, condup1/2, stats.f90:110-253; recvXC sums the    */
/* two bands, stats.f90:474-508): the bottom band planes j in [N(g,0), N(g,1)] (p grid:  */
/* [N(4,0)+1, N(4,1)-1]) and the top band planes mirrored, j1 = N(g,3)+1-j (p: N(4,3)-j),*/
/* with the sign of v flipped, folded onto one dnx x dnz canopy tile:                    */
/*   cst_q(i, k, j1) += sum_{ix, iz} x(i + dnx*ix, k + dnz*iz, j)                        */
/* stored per quantity q as a Fortran (dnx, dnz, nb_q) array, quantities in order.       */
/* ------------------------------------------------------------------------------------ */
static const int st_grid[13] = {0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1};
static const double st_topsign[13] = {1, 1, -1, 1, 1, 1, 1, 1, -1, 1, -1, 1, 1};

/* planes per band on grid g (0 u, 1 v, 2 p) and the first top-band slot */
static void stats_bands(const lho* o, int nb[3], int* toplo) {
  const int n31 = o->band_v[0], n32 = o->band_v[1], d = o->d;
  nb[0] = n31 + 1 + d + 1;   /* u: slots 0 .. N(4,1)+d */
  nb[1] = n31 + d + 1;       /* v: slots 0 .. N(3,1)+d */
  nb[2] = n31 + d;           /* p: slots 0 .. N(4,1)-1+d-1 */
  *toplo = n32 + d + 1;      /* same slot on all three grids */
}

long long lho_cond_stats_len(const lho* o) {
  int nb[3], tl;
  stats_bands(o, nb, &tl);
  const long long tile = (long long)(o->px / o->ntilex) * (o->pz / o->ntilez);
  long long n = 0;
  for (int q = 0; q < 13; ++q) n += tile * nb[st_grid[q]];
  return n;
}

/* j1 of slot s on grid g, sign in *sg (0 = bottom, 1 = top); -1 outside both bands */
static int stats_j1(const lho* o, int g, int s, int* top) {
  int nb[3], tl;
  stats_bands(o, nb, &tl);
  const int n = g == 0 ? o->nyu : (g == 1 ? o->nyv : o->nyp);
  if (s < nb[g]) { *top = 0; return s; }
  if (s >= tl) {
    const int j1 = n - 1 - s;
    if (j1 >= 0 && j1 < nb[g]) { *top = 1; return j1; }
  }
  return -1;
}

/* spectral plane of field f at slot s, with v_itp / w_itp / wx built in Fourier space */
enum { SF_U, SF_V, SF_W, SF_P, SF_VI, SF_WI, SF_WX };
static void stats_plane(const lho* o, int f, int s, cplx* sp) {
  const int nu = o->nyu, nv = o->nyv;
  for (int c = 0; c < o->ncol; ++c) {
    cplx r = {0.0, 0.0};
    switch (f) {
      case SF_U: r = o->u1[(size_t)c * nu + s]; break;
      case SF_W: r = o->u3[(size_t)c * nu + s]; break;
      case SF_V: r = o->u2[(size_t)c * nv + s]; break;
      case SF_P: r = o->p[(size_t)c * o->nyp + s]; break;
      case SF_VI:
        if (s >= 1 && s <= nu - 2)
          r = rmul(0.5, cadd(o->u2[(size_t)c * nv + s - 1], o->u2[(size_t)c * nv + s]));
        break;
      case SF_WI:
        r = rmul(0.5, cadd(o->u3[(size_t)c * nu + s], o->u3[(size_t)c * nu + s + 1]));
        break;
      case SF_WX: {
        const int k = c / o->mx;
        const cplx dw = rmul(o->ddthetavi * o->dthdyv[s],
                             csub(o->u3[(size_t)c * nu + s + 1], o->u3[(size_t)c * nu + s]));
        r = csub(dw, imul(o->k1z[k], o->u2[(size_t)c * nv + s]));
      } break;
    }
    sp[c] = r;
  }
}

void lho_stats_accumulate(lho* o) {
  const int px = o->px, pz = o->pz;
  const int dnx = px / o->ntilex, dnz = pz / o->ntilez;
  const size_t pp = (size_t)px * pz, tile = (size_t)dnx * dnz;
  const long long clen = lho_cond_stats_len(o);
  if (!o->cst) {
    o->cst = (double*)calloc((size_t)clen, sizeof(double));
    o->cst_len = clen;
  }
  long long qoff[13];
  {
    int nb[3], tl;
    stats_bands(o, nb, &tl);
    long long a = 0;
    for (int q = 0; q < 13; ++q) { qoff[q] = a; a += (long long)tile * nb[st_grid[q]]; }
  }
  /* per plane task: grid, slot */
  const int ntask = o->nyu + o->nyv + o->nyp;
  double* psum = (double*)calloc((size_t)ntask * 13, sizeof(double));
  double* pfold = (double*)calloc((size_t)ntask * 6 * tile, sizeof(double));
#pragma omp parallel
  {
    cplx* sp = (cplx*)malloc(sizeof(cplx) * o->ncol);
    cplx* work = (cplx*)malloc(sizeof(cplx) * work_len(o));
    double* ph[3];
    for (int q = 0; q < 3; ++q) ph[q] = (double*)malloc(sizeof(double) * pp);
#pragma omp for schedule(dynamic)
    for (int t = 0; t < ntask; ++t) {
      int g, s, fl[3], nf;
      if (t < o->nyu) { g = 0; s = t; fl[0] = SF_U; fl[1] = SF_W; fl[2] = SF_VI; nf = 3; }
      else if (t < o->nyu + o->nyv) { g = 1; s = t - o->nyu; fl[0] = SF_V; fl[1] = SF_WI; fl[2] = SF_WX; nf = 3; }
      else { g = 2; s = t - o->nyu - o->nyv; fl[0] = SF_P; nf = 1; }
      for (int q = 0; q < nf; ++q) {
        stats_plane(o, fl[q], s, sp);
        four_to_phys(o, sp, ph[q], work);
      }
      /* local quantities: u grid (U, U2, W, W2, UV, UW), v grid (V, V2, VW, wx, wx2), p (P, P2) */
      double* S = psum + (size_t)t * 13;
      double* F = pfold + (size_t)t * 6 * tile;
      for (size_t e = 0; e < pp; ++e) {
        const size_t x = e % px, z = e / px;
        const size_t ti = (z % dnz) * dnx + (x % dnx);
        double v[6];
        int nv6;
        if (g == 0) {
          const double a = ph[0][e], b = ph[1][e], c = ph[2][e];
          v[0] = a; v[1] = a * a; v[2] = b; v[3] = b * b; v[4] = a * c; v[5] = a * b; nv6 = 6;
        } else if (g == 1) {
          const double a = ph[0][e], b = ph[1][e], c = ph[2][e];
          v[0] = a; v[1] = a * a; v[2] = a * b; v[3] = c; v[4] = c * c; nv6 = 5;
        } else {
          const double a = ph[0][e];
          v[0] = a; v[1] = a * a; nv6 = 2;
        }
        const int wxq = (g == 1 && ((int)x >= o->nx || (int)z >= o->nz)) ? 3 : nv6;
        for (int q = 0; q < nv6; ++q) {
          if (q < wxq) S[q] += v[q];
          F[(size_t)q * tile + ti] += v[q];
        }
      }
    }
    free(sp); free(work);
    for (int q = 0; q < 3; ++q) free(ph[q]);
  }
  /* local -> global quantity index per grid */
  static const int qmap[3][6] = {{0, 1, 4, 5, 8, 9}, {2, 3, 10, 11, 12, -1}, {6, 7, -1, -1, -1, -1}};
  static const int nloc[3] = {6, 5, 2};
  for (int t = 0; t < ntask; ++t) {
    int g, s;
    if (t < o->nyu) { g = 0; s = t; }
    else if (t < o->nyu + o->nyv) { g = 1; s = t - o->nyu; }
    else { g = 2; s = t - o->nyu - o->nyv; }
    int top;
    const int j1 = stats_j1(o, g, s, &top);
    for (int l = 0; l < nloc[g]; ++l) {
      const int q = qmap[g][l];
      o->st[(size_t)q * o->nyu + s] += psum[(size_t)t * 13 + l];
      if (j1 < 0) continue;
      const double sg = top ? st_topsign[q] : 1.0;
      double* dst = o->cst + qoff[q] + (size_t)j1 * tile;
      const double* src = pfold + ((size_t)t * 6 + l) * tile;
      for (size_t e = 0; e < tile; ++e) dst[e] += sg * src[e];
    }
  }
  free(psum);
  free(pfold);
  o->nstat += 1;
}

int lho_stats_read(lho* o, double* out) {
  memcpy(out, o->st, sizeof(double) * 13 * (size_t)o->nyu);
  return o->nstat;
}
long long lho_cond_stats_read(lho* o, double* out, long long len) {
  const long long n = lho_cond_stats_len(o);
  if (out && len >= n) {
    if (o->cst) memcpy(out, o->cst, sizeof(double) * (size_t)n);
    else memset(out, 0, sizeof(double) * (size_t)n);
  }
  return n;
}

/* nonlinear alone (du arrays are scratch here): Nu_new for the current state */
void lho_hook_nonlinear(lho* o, int krk1) {
  memset(o->du1, 0, sizeof(cplx) * (size_t)o->ncol * o->nyu);
  memset(o->du2, 0, sizeof(cplx) * (size_t)o->ncol * o->nyv);
  memset(o->du3, 0, sizeof(cplx) * (size_t)o->ncol * o->nyu);
  nonlinear(o, krk1 - 1);
}
void lho_get_nonlinear_full(const lho* o, double* n1, double* n2, double* n3) {
  memcpy(n1, o->Nu1, sizeof(cplx) * (size_t)o->ncol * o->nyu);
  memcpy(n2, o->Nu2, sizeof(cplx) * (size_t)o->ncol * o->nyv);
  memcpy(n3, o->Nu3, sizeof(cplx) * (size_t)o->ncol * o->nyu);
}
/* solveU / solveV alone (littleharsh.f90:733-824 with LUsolU / LUsolV): u (in/out) and du
   (in, overwritten by the solve's work) on the u grid (grid_u) or the v grid, with the
   y-Laplacian lu for the canopy tip planes; reference layout, full plane ranges */
void lho_hook_solve_vel(lho* o, int grid_u, int krk1, double dt, double* u, double* du, double* lu) {
  solve_vel(o, (cplx*)u, (cplx*)du, (const cplx*)lu, grid_u, krk1 - 1, dt);
}
/* immersed-boundary forcing alone on x (reference layout, grid_u: u grid else v grid)
   with the y-Laplacian Lu and beta */
void lho_hook_ib(lho* o, int grid_u, double beta, double* x, const double* lu) {
  immersed_boundaries(o, (cplx*)x, (const cplx*)lu, grid_u, beta);
}
