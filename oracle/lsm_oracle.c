/*
 * lsm_oracle.c -- CPU restatement of LSMLIB's 3D level-set update toolbox.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is never linked into, loaded by, or called from the
 * lsmlib_amd product.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Every routine restates one reference Fortran-77 subroutine loop for loop,
 * with the reference's operation order and constants, and is exported under
 * the reference's by-reference F77 signature with an "oracle_" prefix (e.g.
 * oracle_lsm3dhjweno5_ for lsm3dhjweno5_), so a test can call the product,
 * this restatement and the flang build of the reference with one argument
 * list.  File:line citations are relative to the LSMLIB source tree.
 *
 * Precision mirrors the reference's build switch (configure.ac:258-272):
 *   -DORACLE_REAL_DOUBLE  real = double, zero_tol = 1e-11, tiny = 1e-99
 *   (default)             real = float,  zero_tol = 1e-5,  tiny = 1e-35
 * Fortran "d0" literals are written as C double literals and default-real
 * literals as (real) casts, so the float build keeps the reference's mixed
 * precision promotion (a float times 0.25d0 is evaluated in double).
 *
 * Compiled with -ffp-contract=off: the reference objects contain no fused
 * multiply-adds.
 */
#include <math.h>
#include <stddef.h>

#ifdef ORACLE_REAL_DOUBLE
typedef double real;
#define ZERO_TOL ((real)1e-11)
#define TINY ((real)1e-99)
#define SIN sin
#define COS cos
#else
typedef float real;
#define ZERO_TOL ((real)1e-5f)
#define TINY ((real)1e-35)
#define SIN sinf
#define COS cosf
#endif

/* Fortran MAX/MIN of two reals: first argument wins ties and NaN handling
 * follows "a > b ? a : b" as flang emits it for real operands. */
static inline double fmax2(double a, double b) { return (a > b) ? a : b; }
static inline double fmin2(double a, double b) { return (a < b) ? a : b; }

/* ---------------------------------------------------------------------- */
/* Boxes: a column-major array a(ilo:ihi, jlo:jhi, klo:khi), i fastest.    */
/* ---------------------------------------------------------------------- */
typedef struct {
  int ilo, ihi, jlo, jhi, klo, khi;
  long nx, nxy;
} box_t;

#define BOXARGS(p)                                                        \
  const int *ilo_##p, const int *ihi_##p, const int *jlo_##p,            \
      const int *jhi_##p, const int *klo_##p, const int *khi_##p
#define MKBOX(p) mkbox(*ilo_##p, *ihi_##p, *jlo_##p, *jhi_##p, *klo_##p, *khi_##p)

static box_t mkbox(int ilo, int ihi, int jlo, int jhi, int klo, int khi) {
  box_t b;
  b.ilo = ilo; b.ihi = ihi; b.jlo = jlo; b.jhi = jhi; b.klo = klo; b.khi = khi;
  b.nx = (long)(ihi - ilo + 1);
  b.nxy = b.nx * (long)(jhi - jlo + 1);
  return b;
}

static inline long at(const box_t *b, int i, int j, int k) {
  return (long)(i - b->ilo) + b->nx * (long)(j - b->jlo) + b->nxy * (long)(k - b->klo);
}
#define A(arr, b, i, j, k) (arr)[at(&(b), (i), (j), (k))]

/* ---------------------------------------------------------------------- */
/* lsm3dComputeDn  lsm_spatial_derivatives3d.f.in:49-176                   */
/* fb[6] = {ilo,ihi,jlo,jhi,klo,khi} of the (possibly widened) fill box.   */
/* ---------------------------------------------------------------------- */
static void compute_dn(real *Dn, box_t bd, const real *Dm, box_t bm,
                       const int fb[6], int n, int dir) {
  const real big = (real)1.e10;            /* :87 parameter (big=1.d10) */
  int off[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
  real sgn;
  if (n % 2 == 1) {                        /* :99-103 odd: face-centred */
    off[dir] = 1; sgn = (real)1.0; sh[dir] = 1;
  } else {                                 /* :104-108 even: centred */
    off[dir] = -1; sgn = (real)-1.0; sh[dir] = 0;
  }
  for (int k = fb[4]; k <= fb[5] + sh[2]; k++)           /* :111-121 */
    for (int j = fb[2]; j <= fb[3] + sh[1]; j++)
      for (int i = fb[0]; i <= fb[1] + sh[0]; i++)
        A(Dn, bd, i, j, k) =
            sgn * (A(Dm, bm, i, j, k) - A(Dm, bm, i - off[0], j - off[1], k - off[2]));
  /* :124-170 sentinel outside the computed range */
  for (int k = bd.klo; k <= bd.khi; k++)
    for (int j = bd.jlo; j <= bd.jhi; j++) {
      for (int i = bd.ilo; i <= fb[0] - 1; i++) A(Dn, bd, i, j, k) = big;
      for (int i = fb[1] + sh[0] + 1; i <= bd.ihi; i++) A(Dn, bd, i, j, k) = big;
    }
  for (int k = bd.klo; k <= bd.khi; k++) {
    for (int j = bd.jlo; j <= fb[2] - 1; j++)
      for (int i = bd.ilo; i <= bd.ihi; i++) A(Dn, bd, i, j, k) = big;
    for (int j = fb[3] + sh[1] + 1; j <= bd.jhi; j++)
      for (int i = bd.ilo; i <= bd.ihi; i++) A(Dn, bd, i, j, k) = big;
  }
  for (int k = bd.klo; k <= fb[4] - 1; k++)
    for (int j = bd.jlo; j <= bd.jhi; j++)
      for (int i = bd.ilo; i <= bd.ihi; i++) A(Dn, bd, i, j, k) = big;
  for (int k = fb[5] + sh[2] + 1; k <= bd.khi; k++)
    for (int j = bd.jlo; j <= bd.jhi; j++)
      for (int i = bd.ilo; i <= bd.ihi; i++) A(Dn, bd, i, j, k) = big;
}

/* fill box widened by (lo, hi) cells along dir, as the callers of
 * lsm3dComputeDn pass it (e.g. ilo_fb-2, ihi_fb+2 at :1277-1279). */
static void widen(int out[6], const int fb[6], int dir, int lo, int hi) {
  for (int d = 0; d < 6; d++) out[d] = fb[d];
  out[2 * dir] -= lo;
  out[2 * dir + 1] += hi;
}

/* unit offset along dir */
#define DI(d) ((d) == 0)
#define DJ(d) ((d) == 1)
#define DK(d) ((d) == 2)

/* ---------------------------------------------------------------------- */
/* lsm3dHJENO1  lsm_spatial_derivatives3d.f.in:199-369                      */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dhjeno1_(real *px_p, real *py_p, real *pz_p, BOXARGS(gp),
                         real *px_m, real *py_m, real *pz_m, BOXARGS(gm),
                         const real *phi, BOXARGS(phi), real *D1, BOXARGS(d1),
                         BOXARGS(fb), const real *dx, const real *dy, const real *dz) {
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bphi = MKBOX(phi), bd1 = MKBOX(d1);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb};
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;   /* :276-278 */
  real *outp[3] = {px_p, py_p, pz_p}, *outm[3] = {px_m, py_m, pz_m};
  for (int d = 0; d < 3; d++) {
    compute_dn(D1, bd1, phi, bphi, fb, 1, d);                       /* :284-296 */
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {                      /* :299-308 */
          A(outp[d], bp, i, j, k) = A(D1, bd1, i + DI(d), j + DJ(d), k + DK(d)) * inv[d];
          A(outm[d], bm, i, j, k) = A(D1, bd1, i, j, k) * inv[d];
        }
  }
}

/* ---------------------------------------------------------------------- */
/* lsm3dHJENO2  lsm_spatial_derivatives3d.f.in:393-664                      */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dhjeno2_(real *px_p, real *py_p, real *pz_p, BOXARGS(gp),
                         real *px_m, real *py_m, real *pz_m, BOXARGS(gm),
                         const real *phi, BOXARGS(phi), real *D1, BOXARGS(d1),
                         real *D2, BOXARGS(d2), BOXARGS(fb),
                         const real *dx, const real *dy, const real *dz) {
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bphi = MKBOX(phi), bd1 = MKBOX(d1), bd2 = MKBOX(d2);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb}, w[6];
  const real half = (real)0.5;                                       /* :471 */
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *outp[3] = {px_p, py_p, pz_p}, *outm[3] = {px_m, py_m, pz_m};
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    widen(w, fb, d, 1, 1);                                           /* :485-511 */
    compute_dn(D1, bd1, phi, bphi, w, 1, d);
    compute_dn(D2, bd2, D1, bd1, w, 2, d);
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {                      /* :518-536 */
          real d2m = A(D2, bd2, i - a, j - b, k - c), d20 = A(D2, bd2, i, j, k);
          real d2p = A(D2, bd2, i + a, j + b, k + c);
          if (fabs(d20) < fabs(d2p))
            A(outp[d], bp, i, j, k) = (A(D1, bd1, i + a, j + b, k + c) - half * d20) * inv[d];
          else
            A(outp[d], bp, i, j, k) = (A(D1, bd1, i + a, j + b, k + c) - half * d2p) * inv[d];
          if (fabs(d2m) < fabs(d20))
            A(outm[d], bm, i, j, k) = (A(D1, bd1, i, j, k) + half * d2m) * inv[d];
          else
            A(outm[d], bm, i, j, k) = (A(D1, bd1, i, j, k) + half * d20) * inv[d];
        }
  }
}

/* ---------------------------------------------------------------------- */
/* lsm3dHJENO3  lsm_spatial_derivatives3d.f.in:689-1138                     */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dhjeno3_(real *px_p, real *py_p, real *pz_p, BOXARGS(gp),
                         real *px_m, real *py_m, real *pz_m, BOXARGS(gm),
                         const real *phi, BOXARGS(phi), real *D1, BOXARGS(d1),
                         real *D2, BOXARGS(d2), real *D3, BOXARGS(d3), BOXARGS(fb),
                         const real *dx, const real *dy, const real *dz) {
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bphi = MKBOX(phi);
  box_t bd1 = MKBOX(d1), bd2 = MKBOX(d2), bd3 = MKBOX(d3);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb}, w2[6], w1[6];
  const real half = (real)0.5, third = (real)(1.0 / 3.0), sixth = (real)(1.0 / 6.0); /* :776-778 */
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *outp[3] = {px_p, py_p, pz_p}, *outm[3] = {px_m, py_m, pz_m};
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    widen(w2, fb, d, 2, 2);                                          /* :793-832 */
    widen(w1, fb, d, 1, 1);
    compute_dn(D1, bd1, phi, bphi, w2, 1, d);
    compute_dn(D2, bd2, D1, bd1, w2, 2, d);
    compute_dn(D3, bd3, D2, bd2, w1, 3, d);
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {
#define D2AT(s) A(D2, bd2, i + (s) * a, j + (s) * b, k + (s) * c)
#define D3AT(s) A(D3, bd3, i + (s) * a, j + (s) * b, k + (s) * c)
          real v;                                                   /* :838-868 plus */
          v = A(D1, bd1, i + a, j + b, k + c);
          if (fabs(D2AT(0)) < fabs(D2AT(1))) {
            v = v - half * D2AT(0);
            if (fabs(D3AT(0)) < fabs(D3AT(1))) v = v - sixth * D3AT(0);
            else v = v - sixth * D3AT(1);
          } else {
            v = v - half * D2AT(1);
            if (fabs(D3AT(1)) < fabs(D3AT(2))) v = v + third * D3AT(1);
            else v = v + third * D3AT(2);
          }
          A(outp[d], bp, i, j, k) = v * inv[d];
          v = A(D1, bd1, i, j, k);                                  /* :872-896 minus */
          if (fabs(D2AT(-1)) < fabs(D2AT(0))) {
            v = v + half * D2AT(-1);
            if (fabs(D3AT(-1)) < fabs(D3AT(0))) v = v + third * D3AT(-1);
            else v = v + third * D3AT(0);
          } else {
            v = v + half * D2AT(0);
            if (fabs(D3AT(0)) < fabs(D3AT(1))) v = v - sixth * D3AT(0);
            else v = v - sixth * D3AT(1);
          }
          A(outm[d], bm, i, j, k) = v * inv[d];
#undef D2AT
#undef D3AT
        }
  }
}

/* ---------------------------------------------------------------------- */
/* The HJ-WENO5 formula for one side of one axis,                           */
/* lsm_spatial_derivatives3d.f.in:1292-1326 (identical in every variant).   */
/* ---------------------------------------------------------------------- */
static real weno5(real v1, real v2, real v3, real v4, real v5) {
  const real one_third = (real)(1.0 / 3.0), seven_sixths = (real)(7.0 / 6.0);
  const real eleven_sixths = (real)(11.0 / 6.0), one_sixth = (real)(1.0 / 6.0);
  const real five_sixths = (real)(5.0 / 6.0), thirteen_twelfths = (real)(13.0 / 12.0);
  const real one_fourth = (real)0.25;                                /* :1240-1247 */
  real eps, p1, p2, p3, S1, S2, S3, a1, a2, a3, inv_sum_a, m;
  /* :1296-1297  eps = 1e-6*max(v1*v1,...,v5*v5) + tiny  (MAX left to right) */
  m = v1 * v1;
  m = fmax2(m, v2 * v2);
  m = fmax2(m, v3 * v3);
  m = fmax2(m, v4 * v4);
  m = fmax2(m, v5 * v5);
  eps = (real)1e-6 * m + TINY;
  /* :1300-1303 */
  p1 = one_third * v1 - seven_sixths * v2 + eleven_sixths * v3;
  p2 = -one_sixth * v2 + five_sixths * v3 + one_third * v4;
  p3 = one_third * v3 + five_sixths * v4 - one_sixth * v5;
  /* :1306-1311  (x)**2 evaluated as x*x, "2.d0*" terms promote to double */
  {
    double t1 = v1 - 2.0 * v2 + v3, t2 = v1 - 4.0 * v2 + 3.0 * v3;
    S1 = thirteen_twelfths * (t1 * t1) + one_fourth * (t2 * t2);
    double t3 = v2 - 2.0 * v3 + v4;
    real t4 = v2 - v4;
    S2 = thirteen_twelfths * (t3 * t3) + one_fourth * (t4 * t4);
    double t5 = v3 - 2.0 * v4 + v5, t6 = 3.0 * v3 - 4.0 * v4 + v5;
    S3 = thirteen_twelfths * (t5 * t5) + one_fourth * (t6 * t6);
  }
  /* :1314-1320 */
  {
    real s;
    s = S1 + eps; a1 = 0.1 / (s * s);
    s = S2 + eps; a2 = 0.6 / (s * s);
    s = S3 + eps; a3 = 0.3 / (s * s);
  }
  inv_sum_a = 1.0 / (a1 + a2 + a3);
  a1 = a1 * inv_sum_a;
  a2 = a2 * inv_sum_a;
  a3 = a3 * inv_sum_a;
  return a1 * p1 + a2 * p2 + a3 * p3;                                /* :1323 */
}

/* ---------------------------------------------------------------------- */
/* lsm3dHJWENO5  lsm_spatial_derivatives3d.f.in:1161-1612                   */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dhjweno5_(real *px_p, real *py_p, real *pz_p, BOXARGS(gp),
                          real *px_m, real *py_m, real *pz_m, BOXARGS(gm),
                          const real *phi, BOXARGS(phi), real *D1, BOXARGS(d1),
                          BOXARGS(fb), const real *dx, const real *dy, const real *dz) {
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bphi = MKBOX(phi), bd1 = MKBOX(d1);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb}, w[6];
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;      /* :1256-1258 */
  real *outp[3] = {px_p, py_p, pz_p}, *outm[3] = {px_m, py_m, pz_m};
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    widen(w, fb, d, 2, 2);                                           /* :1265-1277 */
    compute_dn(D1, bd1, phi, bphi, w, 1, d);
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {
#define D1AT(s) (A(D1, bd1, i + (s) * a, j + (s) * b, k + (s) * c) * inv[d])
          real v1, v2, v3, v4, v5;
          v1 = D1AT(3); v2 = D1AT(2); v3 = D1AT(1); v4 = D1AT(0); v5 = D1AT(-1); /* :1286-1290 */
          A(outp[d], bp, i, j, k) = weno5(v1, v2, v3, v4, v5);
          v1 = D1AT(-2); v2 = D1AT(-1); v3 = D1AT(0); v4 = D1AT(1); v5 = D1AT(2); /* :1330-1334 */
          A(outm[d], bm, i, j, k) = weno5(v1, v2, v3, v4, v5);
#undef D1AT
        }
  }
}

/* ---------------------------------------------------------------------- */
/* Upwind variants: one side per axis chosen by the velocity sign;          */
/* |vel| < zero_tol gives 0  (lsm_spatial_derivatives3d.f.in:1735-1743,     */
/* 1990-2010, 2330-2380, 2756-2810).                                        */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dupwindhjeno1_(real *px, real *py, real *pz, BOXARGS(g),
                               const real *phi, BOXARGS(phi), const real *vx,
                               const real *vy, const real *vz, BOXARGS(vel), real *D1,
                               BOXARGS(d1), BOXARGS(fb), const real *dx, const real *dy,
                               const real *dz) {
  /* lsm_spatial_derivatives3d.f.in:1630-1835 */
  box_t bg = MKBOX(g), bphi = MKBOX(phi), bv = MKBOX(vel), bd1 = MKBOX(d1);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb};
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *out[3] = {px, py, pz};
  const real *vel[3] = {vx, vy, vz};
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    compute_dn(D1, bd1, phi, bphi, fb, 1, d);
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {
          real v = A(vel[d], bv, i, j, k);
          if (fabs(v) < ZERO_TOL) A(out[d], bg, i, j, k) = (real)0.0;
          else if (v > 0) A(out[d], bg, i, j, k) = A(D1, bd1, i, j, k) * inv[d];
          else A(out[d], bg, i, j, k) = A(D1, bd1, i + a, j + b, k + c) * inv[d];
        }
  }
}

void oracle_lsm3dupwindhjeno2_(real *px, real *py, real *pz, BOXARGS(g),
                               const real *phi, BOXARGS(phi), const real *vx,
                               const real *vy, const real *vz, BOXARGS(vel), real *D1,
                               BOXARGS(d1), real *D2, BOXARGS(d2), BOXARGS(fb),
                               const real *dx, const real *dy, const real *dz) {
  /* lsm_spatial_derivatives3d.f.in:1854-2153 */
  box_t bg = MKBOX(g), bphi = MKBOX(phi), bv = MKBOX(vel), bd1 = MKBOX(d1), bd2 = MKBOX(d2);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb}, w[6];
  const real half = (real)0.5;
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *out[3] = {px, py, pz};
  const real *vel[3] = {vx, vy, vz};
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    widen(w, fb, d, 1, 1);
    compute_dn(D1, bd1, phi, bphi, w, 1, d);
    compute_dn(D2, bd2, D1, bd1, w, 2, d);
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {
          real v = A(vel[d], bv, i, j, k);
          real d2m = A(D2, bd2, i - a, j - b, k - c), d20 = A(D2, bd2, i, j, k);
          real d2p = A(D2, bd2, i + a, j + b, k + c);
          real r;
          if (fabs(v) < ZERO_TOL) r = (real)0.0;
          else if (v > 0) {
            if (fabs(d2m) < fabs(d20)) r = (A(D1, bd1, i, j, k) + half * d2m) * inv[d];
            else r = (A(D1, bd1, i, j, k) + half * d20) * inv[d];
          } else {
            if (fabs(d20) < fabs(d2p)) r = (A(D1, bd1, i + a, j + b, k + c) - half * d20) * inv[d];
            else r = (A(D1, bd1, i + a, j + b, k + c) - half * d2p) * inv[d];
          }
          A(out[d], bg, i, j, k) = r;
        }
  }
}

void oracle_lsm3dupwindhjeno3_(real *px, real *py, real *pz, BOXARGS(g),
                               const real *phi, BOXARGS(phi), const real *vx,
                               const real *vy, const real *vz, BOXARGS(vel), real *D1,
                               BOXARGS(d1), real *D2, BOXARGS(d2), real *D3, BOXARGS(d3),
                               BOXARGS(fb), const real *dx, const real *dy, const real *dz) {
  /* lsm_spatial_derivatives3d.f.in:2174-2609 */
  box_t bg = MKBOX(g), bphi = MKBOX(phi), bv = MKBOX(vel);
  box_t bd1 = MKBOX(d1), bd2 = MKBOX(d2), bd3 = MKBOX(d3);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb}, w2[6], w1[6];
  const real half = (real)0.5, third = (real)(1.0 / 3.0), sixth = (real)(1.0 / 6.0);
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *out[3] = {px, py, pz};
  const real *vel[3] = {vx, vy, vz};
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    widen(w2, fb, d, 2, 2);
    widen(w1, fb, d, 1, 1);
    compute_dn(D1, bd1, phi, bphi, w2, 1, d);
    compute_dn(D2, bd2, D1, bd1, w2, 2, d);
    compute_dn(D3, bd3, D2, bd2, w1, 3, d);
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {
#define D2AT(s) A(D2, bd2, i + (s) * a, j + (s) * b, k + (s) * c)
#define D3AT(s) A(D3, bd3, i + (s) * a, j + (s) * b, k + (s) * c)
          real vv = A(vel[d], bv, i, j, k), r;
          if (fabs(vv) < ZERO_TOL) {
            r = (real)0.0;
          } else if (vv > 0) {
            r = A(D1, bd1, i, j, k);
            if (fabs(D2AT(-1)) < fabs(D2AT(0))) {
              r = r + half * D2AT(-1);
              if (fabs(D3AT(-1)) < fabs(D3AT(0))) r = r + third * D3AT(-1);
              else r = r + third * D3AT(0);
            } else {
              r = r + half * D2AT(0);
              if (fabs(D3AT(0)) < fabs(D3AT(1))) r = r - sixth * D3AT(0);
              else r = r - sixth * D3AT(1);
            }
            r = r * inv[d];
          } else {
            r = A(D1, bd1, i + a, j + b, k + c);
            if (fabs(D2AT(0)) < fabs(D2AT(1))) {
              r = r - half * D2AT(0);
              if (fabs(D3AT(0)) < fabs(D3AT(1))) r = r - sixth * D3AT(0);
              else r = r - sixth * D3AT(1);
            } else {
              r = r - half * D2AT(1);
              if (fabs(D3AT(1)) < fabs(D3AT(2))) r = r + third * D3AT(1);
              else r = r + third * D3AT(2);
            }
            r = r * inv[d];
          }
          A(out[d], bg, i, j, k) = r;
#undef D2AT
#undef D3AT
        }
  }
}

void oracle_lsm3dupwindhjweno5_(real *px, real *py, real *pz, BOXARGS(g),
                                const real *phi, BOXARGS(phi), const real *vx,
                                const real *vy, const real *vz, BOXARGS(vel), real *D1,
                                BOXARGS(d1), BOXARGS(fb), const real *dx, const real *dy,
                                const real *dz) {
  /* lsm_spatial_derivatives3d.f.in:2628-2999 */
  box_t bg = MKBOX(g), bphi = MKBOX(phi), bv = MKBOX(vel), bd1 = MKBOX(d1);
  int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb}, w[6];
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *out[3] = {px, py, pz};
  const real *vel[3] = {vx, vy, vz};
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    widen(w, fb, d, 2, 2);                                           /* :2735-2747 */
    compute_dn(D1, bd1, phi, bphi, w, 1, d);
    for (int k = fb[4]; k <= fb[5]; k++)
      for (int j = fb[2]; j <= fb[3]; j++)
        for (int i = fb[0]; i <= fb[1]; i++) {
#define D1AT(s) (A(D1, bd1, i + (s) * a, j + (s) * b, k + (s) * c) * inv[d])
          real vv = A(vel[d], bv, i, j, k);
          if (fabs(vv) < ZERO_TOL) {                                /* :2756-2757 */
            A(out[d], bg, i, j, k) = (real)0.0;
          } else if (vv > 0) {                                      /* :2759-2766 */
            A(out[d], bg, i, j, k) = weno5(D1AT(-2), D1AT(-1), D1AT(0), D1AT(1), D1AT(2));
          } else {                                                  /* :2768-2775 */
            A(out[d], bg, i, j, k) = weno5(D1AT(3), D1AT(2), D1AT(1), D1AT(0), D1AT(-1));
          }
#undef D1AT
        }
  }
}

/* ---------------------------------------------------------------------- */
/* Level set equation RHS  lsm_level_set_evolution3d.f.in                    */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dzerooutlevelseteqnrhs_(real *rhs, BOXARGS(r)) {
  /* :16-47 zero the whole ghost box */
  box_t br = MKBOX(r);
  for (int k = br.klo; k <= br.khi; k++)
    for (int j = br.jlo; j <= br.jhi; j++)
      for (int i = br.ilo; i <= br.ihi; i++) A(rhs, br, i, j, k) = (real)0.0;
}

void oracle_lsm3dsamraiutilitiescopydata_(real *dst, BOXARGS(d), const real *src, BOXARGS(s), BOXARGS(fb)) {
  /* lsm_samrai_f77_utilities.f:124-171 (src/parallel/fortran) */
  box_t bd = MKBOX(d), bs = MKBOX(s);
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) A(dst, bd, i, j, k) = A(src, bs, i, j, k);
}

void oracle_lsm3daddadvectiontermtolserhs_(real *rhs, BOXARGS(r), const real *px,
                                           const real *py, const real *pz, BOXARGS(g),
                                           const real *vx, const real *vy, const real *vz,
                                           BOXARGS(vel), BOXARGS(fb)) {
  /* :53-125; the update is at :113-116 */
  box_t br = MKBOX(r), bg = MKBOX(g), bv = MKBOX(vel);
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++)
        A(rhs, br, i, j, k) = A(rhs, br, i, j, k) -
                              (A(vx, bv, i, j, k) * A(px, bg, i, j, k) +
                               A(vy, bv, i, j, k) * A(py, bg, i, j, k) +
                               A(vz, bv, i, j, k) * A(pz, bg, i, j, k));
}

/* Godunov |grad phi|^2 for a normal speed of the given sign,
 * lsm_level_set_evolution3d.f.in:206-225 (and :308-327). */
static real godunov_norm_sq(real vn, real xp, real yp, real zp, real xm, real ym, real zm) {
  double r;
  if (vn > 0.0) {
    double a = fmax2(xm, 0.0), b = fmin2(xp, 0.0);
    double c = fmax2(ym, 0.0), e = fmin2(yp, 0.0);
    double f = fmax2(zm, 0.0), g = fmin2(zp, 0.0);
    r = fmax2(a * a, b * b) + fmax2(c * c, e * e) + fmax2(f * f, g * g);
  } else {
    double a = fmin2(xm, 0.0), b = fmax2(xp, 0.0);
    double c = fmin2(ym, 0.0), e = fmax2(yp, 0.0);
    double f = fmin2(zm, 0.0), g = fmax2(zp, 0.0);
    r = fmax2(a * a, b * b) + fmax2(c * c, e * e) + fmax2(f * f, g * g);
  }
  return (real)r;
}

void oracle_lsm3daddnormalveltermtolserhs_(real *rhs, BOXARGS(r), const real *xp,
                                           const real *yp, const real *zp, BOXARGS(gp),
                                           const real *xm, const real *ym, const real *zm,
                                           BOXARGS(gm), const real *vn, BOXARGS(vel),
                                           BOXARGS(fb)) {
  /* :131-242 */
  box_t br = MKBOX(r), bp = MKBOX(gp), bm = MKBOX(gm), bv = MKBOX(vel);
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real v = A(vn, bv, i, j, k);
        if (fabs(v) >= ZERO_TOL) {
          real n2 = godunov_norm_sq(v, A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k),
                                    A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k));
          A(rhs, br, i, j, k) = A(rhs, br, i, j, k) - v * (real)sqrt(n2);   /* :229-230 */
        }
      }
}

void oracle_lsm3daddconstnormalveltermtolserhs_(real *rhs, BOXARGS(r), const real *xp,
                                                const real *yp, const real *zp, BOXARGS(gp),
                                                const real *xm, const real *ym, const real *zm,
                                                BOXARGS(gm), const real *vn, BOXARGS(fb)) {
  /* :248-350 */
  box_t br = MKBOX(r), bp = MKBOX(gp), bm = MKBOX(gm);
  real v = *vn;
  if (fabs(v) >= ZERO_TOL)
    for (int k = *klo_fb; k <= *khi_fb; k++)
      for (int j = *jlo_fb; j <= *jhi_fb; j++)
        for (int i = *ilo_fb; i <= *ihi_fb; i++) {
          real n2 = godunov_norm_sq(v, A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k),
                                    A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k));
          A(rhs, br, i, j, k) = A(rhs, br, i, j, k) - v * (real)sqrt(n2);
        }
}

/* lsm3dAddConstCurvTermToLSERHS  lsm_level_set_evolution3d.f.in:356-460:
 * rhs += b * kappa |grad phi| with kappa |grad phi| formed from the central
 * first and second derivatives as  N / |grad phi|^2,  N the sum at :439-447
 * (left to right, products left-associated; 0 where |grad phi|^2 < zero_tol). */
void oracle_lsm3daddconstcurvtermtolserhs_(real *rhs, BOXARGS(r), const real *px,
                                           const real *py, const real *pz, BOXARGS(g1),
                                           const real *pxx, const real *pxy, const real *pxz,
                                           const real *pyy, const real *pyz, const real *pzz,
                                           BOXARGS(g2), const real *b, BOXARGS(fb)) {
  box_t br = MKBOX(r), b1 = MKBOX(g1), b2 = MKBOX(g2);
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real x = A(px, b1, i, j, k), y = A(py, b1, i, j, k), z = A(pz, b1, i, j, k);
        real xx = A(pxx, b2, i, j, k), xy = A(pxy, b2, i, j, k), xz = A(pxz, b2, i, j, k);
        real yy = A(pyy, b2, i, j, k), yz = A(pyz, b2, i, j, k), zz = A(pzz, b2, i, j, k);
        real g2 = x * x + y * y + z * z;                     /* :433-435 */
        real curv;
        if (g2 < ZERO_TOL) {
          curv = (real)0.0;
        } else {
          curv = xx * y * y + yy * x * x - (real)2 * xy * x * y + xx * z * z + zz * x * x -
                 (real)2 * xz * x * z + yy * z * z + zz * y * y - (real)2 * yz * y * z;
          curv = curv / g2;
        }
        A(rhs, br, i, j, k) = A(rhs, br, i, j, k) + *b * curv;   /* :451 */
      }
}

/* Godunov selection for the combined external + normal velocity
 * Hamiltonian, lsm_level_set_evolution3d.f.in:630-650. */
static real ext_normal_select(real vcomp, real vn, real dp, real dm) {
  real Hp = vcomp + vn * dp, Hm = vcomp + vn * dm;
  if (Hp > (real)0.0) {
    if (Hm > (real)0.0) return dm;
    return -(vcomp / vn);
  }
  if (Hm < (real)0.0) return dp;
  if (fabs(Hp) > fabs(Hm)) return dp;
  return dm;
}

void oracle_lsm3daddexternalandnormalveltermtolserhs_(
    real *rhs, BOXARGS(r), const real *xp, const real *yp, const real *zp, BOXARGS(gp),
    const real *xm, const real *ym, const real *zm, BOXARGS(gm), const real *vn,
    const real *vx, const real *vy, const real *vz, BOXARGS(vel), BOXARGS(rfb)) {
  /* :532-719.  The reference loops k = klo_rhs_fb..khi_rhs_gb,
   * j = jlo_rhs_fb..jhi_rhs_gb, i = ilo_rhs_gb..ihi_rhs_gb (:620-622);
   * restated as written. */
  box_t br = MKBOX(r), bp = MKBOX(gp), bm = MKBOX(gm), bv = MKBOX(vel);
  for (int k = *klo_rfb; k <= *khi_r; k++)
    for (int j = *jlo_rfb; j <= *jhi_r; j++)
      for (int i = *ilo_r; i <= *ihi_r; i++) {
        real vnc = A(vn, bv, i, j, k), vxc = A(vx, bv, i, j, k);
        real vyc = A(vy, bv, i, j, k), vzc = A(vz, bv, i, j, k);
        real px = ext_normal_select(vxc, vnc, A(xp, bp, i, j, k), A(xm, bm, i, j, k));
        real py = ext_normal_select(vyc, vnc, A(yp, bp, i, j, k), A(ym, bm, i, j, k));
        real pz = ext_normal_select(vzc, vnc, A(zp, bp, i, j, k), A(zm, bm, i, j, k));
        real n2 = px * px + py * py + pz * pz;                     /* :695-696 */
        if (fabs(vnc) >= ZERO_TOL && n2 >= ZERO_TOL)                 /* :700-708 */
          A(rhs, br, i, j, k) = A(rhs, br, i, j, k) - vxc * px - vyc * py - vzc * pz -
                                vnc * (real)sqrt(n2);
        else
          A(rhs, br, i, j, k) = A(rhs, br, i, j, k) - vxc * px - vyc * py - vzc * pz;
      }
}

/* ---------------------------------------------------------------------- */
/* TVD Runge-Kutta stages  lsm_tvd_runge_kutta3d.f                          */
/* ---------------------------------------------------------------------- */
void oracle_lsm3drk1step_(real *un, BOXARGS(un), const real *uc, BOXARGS(uc),
                          const real *rhs, BOXARGS(r), BOXARGS(fb), const real *dt) {
  /* :32-87, update at :78 */
  box_t bn = MKBOX(un), bc = MKBOX(uc), br = MKBOX(r);
  real d = *dt;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++)
        A(un, bn, i, j, k) = A(uc, bc, i, j, k) + d * A(rhs, br, i, j, k);
}

void oracle_lsm3dtvdrk2stage1_(real *u1, BOXARGS(u1), const real *uc, BOXARGS(uc),
                               const real *rhs, BOXARGS(r), BOXARGS(fb), const real *dt) {
  /* :107-155 forwards to lsm3dRK1Step */
  oracle_lsm3drk1step_(u1, ilo_u1, ihi_u1, jlo_u1, jhi_u1, klo_u1, khi_u1, uc, ilo_uc, ihi_uc,
                       jlo_uc, jhi_uc, klo_uc, khi_uc, rhs, ilo_r, ihi_r, jlo_r, jhi_r, klo_r,
                       khi_r, ilo_fb, ihi_fb, jlo_fb, jhi_fb, klo_fb, khi_fb, dt);
}

void oracle_lsm3dtvdrk2stage2_(real *un, BOXARGS(un), const real *u1, BOXARGS(u1),
                               const real *uc, BOXARGS(uc), const real *rhs, BOXARGS(r),
                               BOXARGS(fb), const real *dt) {
  /* :175-250, update at :240-241 */
  box_t bn = MKBOX(un), b1 = MKBOX(u1), bc = MKBOX(uc), br = MKBOX(r);
  real d = *dt;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++)
        A(un, bn, i, j, k) =
            0.5 * (A(uc, bc, i, j, k) + A(u1, b1, i, j, k) + d * A(rhs, br, i, j, k));
}

void oracle_lsm3dtvdrk3stage1_(real *u1, BOXARGS(u1), const real *uc, BOXARGS(uc),
                               const real *rhs, BOXARGS(r), BOXARGS(fb), const real *dt) {
  /* :269-329 forwards to lsm3dRK1Step */
  oracle_lsm3drk1step_(u1, ilo_u1, ihi_u1, jlo_u1, jhi_u1, klo_u1, khi_u1, uc, ilo_uc, ihi_uc,
                       jlo_uc, jhi_uc, klo_uc, khi_uc, rhs, ilo_r, ihi_r, jlo_r, jhi_r, klo_r,
                       khi_r, ilo_fb, ihi_fb, jlo_fb, jhi_fb, klo_fb, khi_fb, dt);
}

void oracle_lsm3dtvdrk3stage2_(real *u2, BOXARGS(u2), const real *u1, BOXARGS(u1),
                               const real *uc, BOXARGS(uc), const real *rhs, BOXARGS(r),
                               BOXARGS(fb), const real *dt) {
  /* :347-413, update at :403-404 */
  box_t b2 = MKBOX(u2), b1 = MKBOX(u1), bc = MKBOX(uc), br = MKBOX(r);
  real d = *dt;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++)
        A(u2, b2, i, j, k) =
            0.75 * A(uc, bc, i, j, k) + 0.25 * (A(u1, b1, i, j, k) + d * A(rhs, br, i, j, k));
}

void oracle_lsm3dtvdrk3stage3_(real *un, BOXARGS(un), const real *u2, BOXARGS(u2),
                               const real *uc, BOXARGS(uc), const real *rhs, BOXARGS(r),
                               BOXARGS(fb), const real *dt) {
  /* :431-501, constants :477-478, update at :486-488 */
  const real one_third = (real)(1.0 / 3.0), two_thirds = (real)(2.0 / 3.0);
  box_t bn = MKBOX(un), b2 = MKBOX(u2), bc = MKBOX(uc), br = MKBOX(r);
  real d = *dt;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++)
        A(un, bn, i, j, k) = one_third * A(uc, bc, i, j, k) +
                             two_thirds * (A(u2, b2, i, j, k) + d * A(rhs, br, i, j, k));
}

/* ---------------------------------------------------------------------- */
/* Reinitialization / orthogonalization  lsm_reinitialization3d.f.in        */
/* ---------------------------------------------------------------------- */
/* Godunov selection of |grad phi|* for the sign of s (:168-185). */
static void godunov_star(real s, const real gp[3], const real gm[3], real gs[3]) {
  for (int d = 0; d < 3; d++) {
    real p, m;
    if (s > 0.0) {
      p = (real)fmax2(-gp[d], 0.0);
      m = (real)fmax2(gm[d], 0.0);
    } else {
      p = (real)fmax2(gp[d], 0.0);
      m = (real)fmax2(-gm[d], 0.0);
    }
    gs[d] = (real)fmax2(p, m);
  }
}

void oracle_lsm3dcomputereinitializationeqnrhs_(
    real *rhs, BOXARGS(r), const real *phi, BOXARGS(phi), const real *phi0, BOXARGS(phi0),
    const real *xp, const real *yp, const real *zp, BOXARGS(gp), const real *xm,
    const real *ym, const real *zm, BOXARGS(gm), BOXARGS(fb), const real *dx, const real *dy,
    const real *dz, const int *use_phi0_for_sgn) {
  /* :48-268 */
  box_t br = MKBOX(r), bphi = MKBOX(phi), bphi0 = MKBOX(phi0), bp = MKBOX(gp), bm = MKBOX(gm);
  const real one = (real)1.0;
  real dx_sq = (real)fmax2(fmax2(*dx, *dy), *dz);                   /* :140-141 */
  dx_sq = dx_sq * dx_sq;
  int use0 = (*use_phi0_for_sgn == 1);
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real cur = use0 ? A(phi0, bphi0, i, j, k) : A(phi, bphi, i, j, k);   /* :157 / :215 */
        real gpv[3] = {A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k)};
        real gmv[3] = {A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k)};
        real gs[3];
        godunov_star(cur, gpv, gmv, gs);
        if (fabs(cur) >= ZERO_TOL) {                               /* :189 / :247 */
          real n2 = gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2];
          real sgn;
          if (!use0) sgn = cur / (real)sqrt(cur * cur + n2 * dx_sq);      /* :193-194 */
          else sgn = cur / (real)sqrt(cur * cur + dx_sq);                 /* :251 */
          A(rhs, br, i, j, k) = sgn * (one - (real)sqrt(n2));
        } else {
          A(rhs, br, i, j, k) = (real)0.0;
        }
      }
}

void oracle_lsm3dcomputeorthogonalizationeqnrhs_(
    real *rhs, BOXARGS(r), const real *xp, const real *yp, const real *zp, BOXARGS(gp),
    const real *xm, const real *ym, const real *zm, BOXARGS(gm), const real *psi,
    BOXARGS(psi), const real *sxp, const real *syp, const real *szp, BOXARGS(sp),
    const real *sxm, const real *sym, const real *szm, BOXARGS(sm), BOXARGS(fb),
    const real *dx, const real *dy, const real *dz) {
  /* :296-471 */
  box_t br = MKBOX(r), bp = MKBOX(gp), bm = MKBOX(gm), bpsi = MKBOX(psi);
  box_t bsp = MKBOX(sp), bsm = MKBOX(sm);
  const real half = (real)0.5;
  real dx_sq = (real)fmax2(fmax2(*dx, *dy), *dz);                   /* :396-397 */
  dx_sq = dx_sq * dx_sq;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real g0 = half * (A(sxp, bsp, i, j, k) + A(sxm, bsm, i, j, k));   /* :411-416 */
        real g1 = half * (A(syp, bsp, i, j, k) + A(sym, bsm, i, j, k));
        real g2 = half * (A(szp, bsp, i, j, k) + A(szm, bsm, i, j, k));
        real ng = g0 * g0 + g1 * g1 + g2 * g2;                     /* :419-422 */
        ng = (real)sqrt(ng);
        real ps = A(psi, bpsi, i, j, k);
        if (fabs(ps) > ZERO_TOL && ng > ZERO_TOL) {                /* :425-426 */
          real s = ps / (real)sqrt(ps * ps + dx_sq * (ng * ng));   /* :430-431 */
          double acc = fmax2(s * g0, 0.0) * A(xm, bm, i, j, k);    /* :433-445 */
          acc = acc + fmin2(s * g0, 0.0) * A(xp, bp, i, j, k);
          acc = acc + fmax2(s * g1, 0.0) * A(ym, bm, i, j, k);
          acc = acc + fmin2(s * g1, 0.0) * A(yp, bp, i, j, k);
          acc = acc + fmax2(s * g2, 0.0) * A(zm, bm, i, j, k);
          acc = acc + fmin2(s * g2, 0.0) * A(zp, bp, i, j, k);
          A(rhs, br, i, j, k) = (real)(-1.0 / ng * acc);
        } else {
          A(rhs, br, i, j, k) = (real)0.0;
        }
      }
}

void oracle_lsm3dcomputedistanceforsubcellfix_(real *dist0, const real *phi0, BOXARGS(p0),
                                               BOXARGS(fb), const real *dx, const real *dy,
                                               const real *dz) {
  /* :477-581 (Russo-Smereka "version 2", :548-562) */
  box_t b = MKBOX(p0);
  real max_dx = (real)fmax2(fmax2(*dx, *dy), *dz);
  real flag = -1000.0 * max_dx;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real c = A(phi0, b, i, j, k);
        int near = 0;
        const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int n = 0; n < 6; n++) {                              /* :518-540 */
          real o = A(phi0, b, i + nb[n][0], j + nb[n][1], k + nb[n][2]);
          if ((o > 0 && c <= 0) || (o <= 0 && c > 0)) near = 1;
        }
        if (!near) {
          A(dist0, b, i, j, k) = flag;
        } else {
          real delta = ZERO_TOL;
          for (int d = 0; d < 3; d++) {
            int a = DI(d), bb = DJ(d), cc = DK(d);
            real pp = A(phi0, b, i + a, j + bb, k + cc), pm = A(phi0, b, i - a, j - bb, k - cc);
            real d1 = (real)(fabs(pp - pm) / 2.0);
            real d2 = (real)fabs(pp - c);
            real d3 = (real)fabs(c - pm);
            /* MAX(d1,d2,d3,prev) evaluated left to right */
            real m = (real)fmax2(d1, d2);
            m = (real)fmax2(m, d3);
            delta = (real)fmax2(m, delta);
          }
          A(dist0, b, i, j, k) = c / delta;
        }
      }
}

void oracle_lsm3dcomputereinitializationeqnrhssubcellfixorder1_(
    real *rhs, BOXARGS(r), const real *phi, BOXARGS(phi), const real *phi0,
    const real *dist0, BOXARGS(p0), const real *xp, const real *yp, const real *zp,
    BOXARGS(gp), const real *xm, const real *ym, const real *zm, BOXARGS(gm), BOXARGS(fb),
    const real *dx, const real *dy, const real *dz) {
  /* :586-754 */
  box_t br = MKBOX(r), bphi = MKBOX(phi), b0 = MKBOX(p0), bp = MKBOX(gp), bm = MKBOX(gm);
  const real zero = (real)0.0;
  real max_dx = (real)fmax2(fmax2(*dx, *dy), *dz);
  real flag = -1000.0 * max_dx;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real p0v = A(phi0, b0, i, j, k), s0;
        if (fabs(p0v) > ZERO_TOL) s0 = (real)copysign(1.0, p0v);  /* :676-680 SIGN(one,phi0) */
        else s0 = zero;
        if (A(dist0, b0, i, j, k) == flag) {
          real cur = A(phi, bphi, i, j, k);
          real gpv[3] = {A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k)};
          real gmv[3] = {A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k)};
          real gs[3];
          godunov_star(cur, gpv, gmv, gs);
          if (fabs(cur) > ZERO_TOL) {
            real n2 = gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2];
            A(rhs, br, i, j, k) = s0 - s0 * (real)sqrt(n2);        /* :722-723 */
          } else {
            A(rhs, br, i, j, k) = (real)0.0;
          }
        } else {
          A(rhs, br, i, j, k) =
              A(dist0, b0, i, j, k) - s0 * (real)fabs(A(phi, bphi, i, j, k)) / max_dx; /* :731-732 */
        }
      }
}

/* ---------------------------------------------------------------------- */
/* Reductions  lsm_utilities3d.f.in, geometry/lsm_geometry3d_fort.f.in      */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dmaxnormdiff_(real *out, const real *f1, BOXARGS(f1), const real *f2,
                              BOXARGS(f2), BOXARGS(ib)) {
  /* lsm_utilities3d.f.in:16-76 */
  box_t b1 = MKBOX(f1), b2 = MKBOX(f2);
  real m = (real)fabs(A(f1, b1, *ilo_ib, *jlo_ib, *klo_ib) - A(f2, b2, *ilo_ib, *jlo_ib, *klo_ib));
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real d = (real)fabs(A(f1, b1, i, j, k) - A(f2, b2, i, j, k));
        if (d > m) m = d;
      }
  *out = m;
}

void oracle_lsm3dmaxnormdiffcontrolvolume_(real *out, const real *f1, BOXARGS(f1),
                                           const real *f2, BOXARGS(f2), const real *cv,
                                           BOXARGS(cv), const int *cv_sgn, BOXARGS(ib)) {
  /* lsm_utilities3d.f.in:852-955; initial value ignores the control volume */
  box_t b1 = MKBOX(f1), b2 = MKBOX(f2), bc = MKBOX(cv);
  real m = (real)fabs(A(f1, b1, *ilo_ib, *jlo_ib, *klo_ib) - A(f2, b2, *ilo_ib, *jlo_ib, *klo_ib));
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real c = A(cv, bc, i, j, k);
        if ((*cv_sgn > 0) ? (c > 0.0) : (c < 0.0)) {
          real d = (real)fabs(A(f1, b1, i, j, k) - A(f2, b2, i, j, k));
          if (d > m) m = d;
        }
      }
  *out = m;
}

static real adv_u_over_dx(real ux, real uy, real uz, real ix, real iy, real iz) {
  /* lsm_utilities3d.f.in:306-308 */
  return (real)fabs(ux) * ix + (real)fabs(uy) * iy + (real)fabs(uz) * iz;
}

void oracle_lsm3dcomputestableadvectiondt_(real *dt, const real *vx, const real *vy,
                                           const real *vz, BOXARGS(vel), BOXARGS(ib),
                                           const real *dx, const real *dy, const real *dz,
                                           const real *cfl) {
  /* lsm_utilities3d.f.in:253-327 */
  box_t bv = MKBOX(vel);
  real mx = (real)-1.0;                                             /* :297 */
  real ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real u = adv_u_over_dx(A(vx, bv, i, j, k), A(vy, bv, i, j, k), A(vz, bv, i, j, k), ix, iy, iz);
        if (u > mx) mx = u;
      }
  *dt = *cfl / (mx + TINY);                                         /* :323 */
}

void oracle_lsm3dcomputestableadvectiondtcontrolvolume_(
    real *dt, const real *vx, const real *vy, const real *vz, BOXARGS(vel), const real *cv,
    BOXARGS(cv), const int *cv_sgn, BOXARGS(ib), const real *dx, const real *dy,
    const real *dz, const real *cfl) {
  /* lsm_utilities3d.f.in:960-1079 */
  box_t bv = MKBOX(vel), bc = MKBOX(cv);
  real mx = (real)-1.0;
  real ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real c = A(cv, bc, i, j, k);
        if ((*cv_sgn > 0) ? (c > 0.0) : (c < 0.0)) {
          real u = adv_u_over_dx(A(vx, bv, i, j, k), A(vy, bv, i, j, k), A(vz, bv, i, j, k), ix, iy, iz);
          if (u > mx) mx = u;
        }
      }
  *dt = *cfl / (mx + TINY);
}

/* H/dX for the normal-velocity CFL, lsm_utilities3d.f.in:404-429. */
static real normal_h_over_dx(real avn, real xp, real yp, real zp, real xm, real ym, real zm,
                             real max_dx_sq, real ix, real iy, real iz) {
  real px = (real)fmax2(fabs(xp), fabs(xm));
  real py = (real)fmax2(fabs(yp), fabs(ym));
  real pz = (real)fmax2(fabs(zp), fabs(zm));
  real ng = (real)sqrt(px * px + py * py + pz * pz + max_dx_sq);
  return avn / ng * (px * ix + py * iy + pz * iz);
}

void oracle_lsm3dcomputestablenormalveldt_(real *dt, const real *vn, BOXARGS(vel),
                                           const real *xp, const real *yp, const real *zp,
                                           BOXARGS(gp), const real *xm, const real *ym,
                                           const real *zm, BOXARGS(gm), BOXARGS(ib),
                                           const real *dx, const real *dy, const real *dz,
                                           const real *cfl) {
  /* lsm_utilities3d.f.in:332-450 */
  box_t bv = MKBOX(vel), bp = MKBOX(gp), bm = MKBOX(gm);
  real mdx = (real)fmax2(fmax2(*dx, *dy), *dz);
  real max_dx_sq = mdx * mdx;
  real mx = (real)-1.0;
  real ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real h = normal_h_over_dx((real)fabs(A(vn, bv, i, j, k)), A(xp, bp, i, j, k),
                                  A(yp, bp, i, j, k), A(zp, bp, i, j, k), A(xm, bm, i, j, k),
                                  A(ym, bm, i, j, k), A(zm, bm, i, j, k), max_dx_sq, ix, iy, iz);
        if (h > mx) mx = h;
      }
  *dt = *cfl / (mx + TINY);
}

void oracle_lsm3dcomputestableconstnormalveldt_(real *dt, const real *vn, const real *xp,
                                                const real *yp, const real *zp, BOXARGS(gp),
                                                const real *xm, const real *ym, const real *zm,
                                                BOXARGS(gm), BOXARGS(ib), const real *dx,
                                                const real *dy, const real *dz,
                                                const real *cfl) {
  /* lsm_utilities3d.f.in:455-568 */
  box_t bp = MKBOX(gp), bm = MKBOX(gm);
  real mdx = (real)fmax2(fmax2(*dx, *dy), *dz);
  real max_dx_sq = mdx * mdx;
  real mx = (real)-1;
  real ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  real avn = (real)fabs(*vn);
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real h = normal_h_over_dx(avn, A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k),
                                  A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k),
                                  max_dx_sq, ix, iy, iz);
        if (h > mx) mx = h;
      }
  *dt = *cfl / (mx + TINY);
}

/* pi and the 14-digit 1/pi the reference uses for smoothed Heaviside
 * integrals (lsm_utilities3d.f.in:619-621, lsm_geometry3d_fort.f.in:299-301) */
#define PI_REF ((real)3.14159265358979323846)
#define ONE_OVER_PI_REF ((real)0.31830988618379)

void oracle_lsm3dvolumeintegralphilessthanzero_(real *out, const real *F, BOXARGS(F),
                                                const real *phi, BOXARGS(phi), BOXARGS(ib),
                                                const real *dx, const real *dy, const real *dz,
                                                const real *eps) {
  /* lsm_utilities3d.f.in:573-654 */
  box_t bF = MKBOX(F), bphi = MKBOX(phi);
  real dV = *dx * *dy * *dz, acc = (real)0.0, e = *eps;
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real p = A(phi, bphi, i, j, k), x = p / e;
        if (p < -e) {
          acc = acc + A(F, bF, i, j, k) * dV;
        } else if (p < e) {
          real omh = 0.5 * (1.0 - x - ONE_OVER_PI_REF * SIN(PI_REF * x));
          acc = acc + omh * A(F, bF, i, j, k) * dV;
        }
      }
  *out = acc;
}

void oracle_lsm3dvolumeintegralphigreaterthanzero_(real *out, const real *F, BOXARGS(F),
                                                   const real *phi, BOXARGS(phi), BOXARGS(ib),
                                                   const real *dx, const real *dy,
                                                   const real *dz, const real *eps) {
  /* lsm_utilities3d.f.in:659-740 */
  box_t bF = MKBOX(F), bphi = MKBOX(phi);
  real dV = *dx * *dy * *dz, acc = (real)0.0, e = *eps;
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real p = A(phi, bphi, i, j, k), x = p / e;
        if (p > e) {
          acc = acc + A(F, bF, i, j, k) * dV;
        } else if (p > -e) {
          real H = 0.5 * (1.0 + x + ONE_OVER_PI_REF * SIN(PI_REF * x));
          acc = acc + H * A(F, bF, i, j, k) * dV;
        }
      }
  *out = acc;
}

void oracle_lsm3dvolumeregionphilessthanzero_(real *out, const real *phi, BOXARGS(phi),
                                              BOXARGS(ib), const real *dx, const real *dy,
                                              const real *dz, const real *eps) {
  /* lsm_geometry3d_fort.f.in:263-345 */
  box_t bphi = MKBOX(phi);
  real dV = *dx * *dy * *dz, vol = (real)0.0, e = *eps;
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real p = A(phi, bphi, i, j, k), x = p / e;
        if (p < -e) {
          vol = vol + dV;
        } else if (p < e) {
          real omh = 0.5 * (1 - x - ONE_OVER_PI_REF * SIN(PI_REF * x));   /* :321-323 */
          vol = vol + omh * dV;
        }
      }
  *out = vol;
}

void oracle_lsm3dvolumeregionphigreaterthanzero_(real *out, const real *phi, BOXARGS(phi),
                                                 BOXARGS(ib), const real *dx, const real *dy,
                                                 const real *dz, const real *eps) {
  /* lsm_geometry3d_fort.f.in:352-422 */
  box_t bphi = MKBOX(phi);
  real dV = *dx * *dy * *dz, vol = (real)0.0, e = *eps;
  for (int k = *klo_ib; k <= *khi_ib; k++)
    for (int j = *jlo_ib; j <= *jhi_ib; j++)
      for (int i = *ilo_ib; i <= *ihi_ib; i++) {
        real p = A(phi, bphi, i, j, k), x = p / e;
        if (p > e) {
          vol = vol + dV;
        } else if (p > -e) {
          real H = (real)0.5 * (1 + x + ONE_OVER_PI_REF * SIN(PI_REF * x)); /* :409-410 */
          vol = vol + H * dV;
        }
      }
  *out = vol;
}

void oracle_lsmheaviside_(real *H, const real *x, const real *eps) {
  /* lsm_calculus_toolbox.f:25-54 */
  real e = *eps, xv = *x;
  if (xv < -e) *H = (real)0.0;
  else if (xv > e) *H = (real)1.0;
  else {
    real ioe = 1.0 / e;
    *H = 0.5 * (1.0 + xv * ioe + ONE_OVER_PI_REF * SIN(PI_REF * xv * ioe));
  }
}

/* ---------------------------------------------------------------------- */
/* Central differences  lsm_spatial_derivatives3d.f.in:3016-3264, 3586-3650 */
/* ---------------------------------------------------------------------- */
void oracle_lsm3dcentralgradorder2_(real *gx, real *gy, real *gz, BOXARGS(g), const real *phi,
                                    BOXARGS(phi), BOXARGS(fb), const real *dx, const real *dy,
                                    const real *dz) {
  /* :3016-3088; the factors are 0.5d0/dx rounded to real (:3058-3060) */
  box_t bg = MKBOX(g), bp = MKBOX(phi);
  real fx = 0.5 / *dx, fy = 0.5 / *dy, fz = 0.5 / *dz;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        A(gx, bg, i, j, k) = (A(phi, bp, i + 1, j, k) - A(phi, bp, i - 1, j, k)) * fx;
        A(gy, bg, i, j, k) = (A(phi, bp, i, j + 1, k) - A(phi, bp, i, j - 1, k)) * fy;
        A(gz, bg, i, j, k) = (A(phi, bp, i, j, k + 1) - A(phi, bp, i, j, k - 1)) * fz;
      }
}

void oracle_lsm3dcentralgradorder4_(real *gx, real *gy, real *gz, BOXARGS(g), const real *phi,
                                    BOXARGS(phi), BOXARGS(fb), const real *dx, const real *dy,
                                    const real *dz) {
  /* :3094-3174; eight is a real parameter, the factor 1/12 a d0 literal */
  box_t bg = MKBOX(g), bp = MKBOX(phi);
  const real eight = (real)8.0;
  real fx = 0.0833333333333333333333 / *dx, fy = 0.0833333333333333333333 / *dy,
       fz = 0.0833333333333333333333 / *dz;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        A(gx, bg, i, j, k) = (-A(phi, bp, i + 2, j, k) + eight * A(phi, bp, i + 1, j, k) +
                              A(phi, bp, i - 2, j, k) - eight * A(phi, bp, i - 1, j, k)) * fx;
        A(gy, bg, i, j, k) = (-A(phi, bp, i, j + 2, k) + eight * A(phi, bp, i, j + 1, k) +
                              A(phi, bp, i, j - 2, k) - eight * A(phi, bp, i, j - 1, k)) * fy;
        A(gz, bg, i, j, k) = (-A(phi, bp, i, j, k + 2) + eight * A(phi, bp, i, j, k + 1) +
                              A(phi, bp, i, j, k - 2) - eight * A(phi, bp, i, j, k - 1)) * fz;
      }
}

void oracle_lsm3dlaplacianorder2_(real *lap, BOXARGS(l), const real *phi, BOXARGS(phi),
                                  BOXARGS(fb), const real *dx, const real *dy, const real *dz) {
  /* :3180-3264.  1.0d0/dx/dx is rounded to real; 2.0d0*phi promotes each
   * bracket, the products and the sum to double before the store. */
  box_t bl = MKBOX(l), bp = MKBOX(phi);
  real ix2 = 1.0 / *dx / *dx, iy2 = 1.0 / *dy / *dy, iz2 = 1.0 / *dz / *dz;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        double c = 2.0 * A(phi, bp, i, j, k);
        A(lap, bl, i, j, k) =
            (real)(ix2 * (A(phi, bp, i + 1, j, k) - c + A(phi, bp, i - 1, j, k)) +
                   iy2 * (A(phi, bp, i, j + 1, k) - c + A(phi, bp, i, j - 1, k)) +
                   iz2 * (A(phi, bp, i, j, k + 1) - c + A(phi, bp, i, j, k - 1)));
      }
}

void oracle_lsm3dgradientmagnitude_(const real *gx, const real *gy, const real *gz, real *mag,
                                    BOXARGS(g), BOXARGS(fb)) {
  /* :3586-3650: squares below tol = 1e-8 (a real parameter) count as 0,
   * and a magnitude below tol is stored as 0 */
  box_t bg = MKBOX(g);
  const real tol = (real)1.e-8;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real x2 = A(gx, bg, i, j, k) * A(gx, bg, i, j, k);
        real y2 = A(gy, bg, i, j, k) * A(gy, bg, i, j, k);
        real z2 = A(gz, bg, i, j, k) * A(gz, bg, i, j, k);
        if (x2 < tol) x2 = (real)0.0;
        if (y2 < tol) y2 = (real)0.0;
        if (z2 < tol) z2 = (real)0.0;
        real t = (real)sqrt(x2 + y2 + z2);
        A(mag, bg, i, j, k) = (t < tol) ? (real)0.0 : t;
      }
}

/* ---------------------------------------------------------------------- */
/* Normals and the field-extension right-hand side                         */
/* ---------------------------------------------------------------------- */
/* phi-upwind selection of one component (lsm_spatial_derivatives3d.f.in:3270-3480):
 * returns 1 and sets *out, or 0 when phi == 0 (the reference leaves F alone). */
static int phi_upwind_pick(real pc, real pm, real pp, real fp, real fm, real *out) {
  if (pc > 0) {
    if (pm <= pc) *out = (pp < pm) ? fp : fm;
    else if (pp <= pc) *out = fp;
    else *out = (real)0.0;
    return 1;
  }
  if (pc < 0) {
    if (pm >= pc) *out = (pp > pm) ? fp : fm;
    else if (pp >= pc) *out = fp;
    else *out = (real)0.0;
    return 1;
  }
  return 0;
}

void oracle_lsm3dphiupwindgradf_(real *fx, real *fy, real *fz, BOXARGS(g), const real *fxp,
                                 const real *fyp, const real *fzp, BOXARGS(gp), const real *fxm,
                                 const real *fym, const real *fzm, BOXARGS(gm), const real *phi,
                                 BOXARGS(phi), BOXARGS(fb)) {
  box_t bg = MKBOX(g), bp = MKBOX(gp), bm = MKBOX(gm), bphi = MKBOX(phi);
  real *out[3] = {fx, fy, fz};
  const real *plus[3] = {fxp, fyp, fzp}, *minus[3] = {fxm, fym, fzm};
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real pc = A(phi, bphi, i, j, k);
        for (int d = 0; d < 3; d++) {
          int a = DI(d), b = DJ(d), c = DK(d);
          real v;
          if (phi_upwind_pick(pc, A(phi, bphi, i - a, j - b, k - c), A(phi, bphi, i + a, j + b, k + c),
                              A(plus[d], bp, i, j, k), A(minus[d], bm, i, j, k), &v))
            A(out[d], bg, i, j, k) = v;
        }
      }
}

void oracle_lsm3daveragegradphi_(real *gx, real *gy, real *gz, BOXARGS(g), const real *xp,
                                 const real *yp, const real *zp, BOXARGS(gp), const real *xm,
                                 const real *ym, const real *zm, BOXARGS(gm), BOXARGS(fb)) {
  /* lsm_spatial_derivatives3d.f.in:3486-3580; half is a real parameter */
  box_t bg = MKBOX(g), bp = MKBOX(gp), bm = MKBOX(gm);
  const real half = (real)0.5;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        A(gx, bg, i, j, k) = half * (A(xp, bp, i, j, k) + A(xm, bm, i, j, k));
        A(gy, bg, i, j, k) = half * (A(yp, bp, i, j, k) + A(ym, bm, i, j, k));
        A(gz, bg, i, j, k) = half * (A(zp, bp, i, j, k) + A(zm, bm, i, j, k));
      }
}

void oracle_lsm3dcomputeunitnormal_(real *nx, real *ny, real *nz, BOXARGS(n), const real *px,
                                    const real *py, const real *pz, BOXARGS(g), BOXARGS(fb)) {
  /* lsm_geometry3d_fort.f.in:27-110; 1.0d0/norm rounded to real */
  box_t bn = MKBOX(n), bg = MKBOX(g);
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real x = A(px, bg, i, j, k), y = A(py, bg, i, j, k), z = A(pz, bg, i, j, k);
        real norm = (real)sqrt(x * x + y * y + z * z);
        if (norm >= ZERO_TOL) {
          real inv = 1.0 / norm;
          A(nx, bn, i, j, k) = x * inv;
          A(ny, bn, i, j, k) = y * inv;
          A(nz, bn, i, j, k) = z * inv;
        } else {
          A(nx, bn, i, j, k) = (real)1.0;
          A(ny, bn, i, j, k) = (real)0.0;
          A(nz, bn, i, j, k) = (real)0.0;
        }
      }
}

void oracle_lsm3dcomputesignedunitnormal_(real *nx, real *ny, real *nz, BOXARGS(n), const real *px,
                                          const real *py, const real *pz, BOXARGS(g), const real *phi,
                                          BOXARGS(phi), BOXARGS(fb), const real *dx, const real *dy,
                                          const real *dz) {
  /* lsm_geometry3d_fort.f.in:129-240 */
  box_t bn = MKBOX(n), bg = MKBOX(g), bphi = MKBOX(phi);
  real m = *dx;
  if (*dy > m) m = *dy;
  if (*dz > m) m = *dz;
  real dx_sq = m * m;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real pc = A(phi, bphi, i, j, k);
        if (fabs(pc) > ZERO_TOL) {
          real x = A(px, bg, i, j, k), y = A(py, bg, i, j, k), z = A(pz, bg, i, j, k);
          real n2 = x * x + y * y + z * z;
          if (n2 >= ZERO_TOL) {
            real sgn = pc / (real)sqrt(pc * pc + n2 * dx_sq);     /* :215-216 */
            real inv = 1.0 / (real)sqrt(n2);                      /* :218 */
            A(nx, bn, i, j, k) = sgn * x * inv;
            A(ny, bn, i, j, k) = sgn * y * inv;
            A(nz, bn, i, j, k) = sgn * z * inv;
          } else {
            A(nx, bn, i, j, k) = (real)1.0;
            A(ny, bn, i, j, k) = (real)0.0;
            A(nz, bn, i, j, k) = (real)0.0;
          }
        } else {
          A(nx, bn, i, j, k) = (real)0.0;
          A(ny, bn, i, j, k) = (real)0.0;
          A(nz, bn, i, j, k) = (real)0.0;
        }
      }
}

void oracle_lsm3dcomputefieldextensioneqnrhs_(real *rhs, BOXARGS(r), const real *S, BOXARGS(S),
                                              const real *phi, BOXARGS(phi), const real *sx,
                                              const real *sy, const real *sz, BOXARGS(gs),
                                              const real *nx, const real *ny, const real *nz,
                                              BOXARGS(n), BOXARGS(fb), const real *dx,
                                              const real *dy, const real *dz) {
  /* lsm_field_extension3d.f:31-158: -(n . grad S) away from the zero level set
   * (|phi| > 3 max(dx) or phi of one sign on the 6 neighbours), else 0 */
  box_t br = MKBOX(r), bphi = MKBOX(phi), bs = MKBOX(gs), bn = MKBOX(n);
  (void)S; (void)ilo_S; (void)ihi_S; (void)jlo_S; (void)jhi_S; (void)klo_S; (void)khi_S;
  real m = *dx;
  if (*dy > m) m = *dy;
  if (*dz > m) m = *dz;
  real cutoff = 3.0 * m;                                           /* :115 */
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real pc = A(phi, bphi, i, j, k);
        int active = fabs(pc) > cutoff;
        if (!active) {
          real z = (real)0.0;
          active = pc * A(phi, bphi, i - 1, j, k) > z && pc * A(phi, bphi, i + 1, j, k) > z &&
                   pc * A(phi, bphi, i, j - 1, k) > z && pc * A(phi, bphi, i, j + 1, k) > z &&
                   pc * A(phi, bphi, i, j, k - 1) > z && pc * A(phi, bphi, i, j, k + 1) > z;
        }
        if (active)
          A(rhs, br, i, j, k) = -(A(nx, bn, i, j, k) * A(sx, bs, i, j, k) +
                                  A(ny, bn, i, j, k) * A(sy, bs, i, j, k) +
                                  A(nz, bn, i, j, k) * A(sz, bs, i, j, k));
        else
          A(rhs, br, i, j, k) = (real)0.0;
      }
}

/* ---------------------------------------------------------------------- */
/* The remaining 3D reductions: lsm_utilities3d.f.in, lsm_geometry3d_fort.f.in */
/* and lsm_calculus_toolbox3d.f.in                                           */
/* ---------------------------------------------------------------------- */
#define CV_IN(c, sgn) (((sgn) > 0) ? ((c) > 0.0) : ((c) < 0.0))
#define LOOP_IB                                       \
  for (int k = *klo_ib; k <= *khi_ib; k++)            \
    for (int j = *jlo_ib; j <= *jhi_ib; j++)          \
      for (int i = *ilo_ib; i <= *ihi_ib; i++)

void oracle_lsm3daveabsdiff_(real *out, const real *f1, BOXARGS(f1), const real *f2, BOXARGS(f2), BOXARGS(ib)) {
  /* lsm_utilities3d.f.in:81-149: the sum and the point count are both real */
  box_t b1 = MKBOX(f1), b2 = MKBOX(f2);
  real sum = (real)0.0, num = (real)0.0;
  LOOP_IB {
    real d = (real)fabs(A(f1, b1, i, j, k) - A(f2, b2, i, j, k));
    sum = sum + d;
    num = num + (real)1.0;
  }
  *out = (num > (real)0.0) ? sum / num : (real)0.0;
}

static void voxel_count(int *count, const real *phi, box_t bp, const real *cv, box_t bc, int cv_sgn, int greater,
                        const int *ilo_ib, const int *ihi_ib, const int *jlo_ib, const int *jhi_ib,
                        const int *klo_ib, const int *khi_ib) {
  /* lsm_utilities3d.f.in:154-247; with a control volume :1849-2007, where
   * the control volume test is >= 0 / <= 0 (unlike the other routines' > / <) */
  int n = 0;
  LOOP_IB {
    real p = A(phi, bp, i, j, k);
    int in = greater ? (p > 0.0) : (p < 0.0);
    if (cv) {
      real c = A(cv, bc, i, j, k);
      in = in && ((cv_sgn > 0) ? (c >= 0.0) : (c <= 0.0));
    }
    if (in) n++;
  }
  *count = n;
}

void oracle_lsm3dvoxelcountgreaterthanzero_(int *count, const real *phi, BOXARGS(g), BOXARGS(ib)) {
  voxel_count(count, phi, MKBOX(g), NULL, MKBOX(g), 0, 1, ilo_ib, ihi_ib, jlo_ib, jhi_ib, klo_ib, khi_ib);
}
void oracle_lsm3dvoxelcountlessthanzero_(int *count, const real *phi, BOXARGS(g), BOXARGS(ib)) {
  voxel_count(count, phi, MKBOX(g), NULL, MKBOX(g), 0, 0, ilo_ib, ihi_ib, jlo_ib, jhi_ib, klo_ib, khi_ib);
}
void oracle_lsm3dvoxelcountgreaterthanzerocontrolvolume_(int *count, const real *phi, BOXARGS(g), const real *cv,
                                                         BOXARGS(cv), const int *cv_sgn, BOXARGS(ib)) {
  voxel_count(count, phi, MKBOX(g), cv, MKBOX(cv), *cv_sgn, 1, ilo_ib, ihi_ib, jlo_ib, jhi_ib, klo_ib, khi_ib);
}
void oracle_lsm3dvoxelcountlessthanzerocontrolvolume_(int *count, const real *phi, BOXARGS(g), const real *cv,
                                                      BOXARGS(cv), const int *cv_sgn, BOXARGS(ib)) {
  voxel_count(count, phi, MKBOX(g), cv, MKBOX(cv), *cv_sgn, 0, ilo_ib, ihi_ib, jlo_ib, jhi_ib, klo_ib, khi_ib);
}

/* delta_eps(phi) |grad phi| summands.  one_plus: 1 for lsm_utilities3d's
 * (1.d0 + cos, promoted to double), 0 for lsm_geometry3d's (1 + cos, real). */
static real smoothed_delta(real p, real ooe, int double_one) {
  real c = COS(PI_REF * p * ooe);
  if (double_one) return (real)(0.5 * (double)ooe * (1.0 + (double)c));
  return (real)(0.5 * (double)ooe * (double)((real)1 + c));
}

static real surface_sum(const real *F, box_t bF, const real *phi, box_t bp, const real *px, const real *py,
                        const real *pz, box_t bg, const real *cv, box_t bc, int cv_sgn, const int *ilo_ib,
                        const int *ihi_ib, const int *jlo_ib, const int *jhi_ib, const int *klo_ib,
                        const int *khi_ib, real dV, real eps, int double_one) {
  real acc = (real)0.0, ooe = 1.0 / eps;
  LOOP_IB {
    if (cv && !CV_IN(A(cv, bc, i, j, k), cv_sgn)) continue;
    real p = A(phi, bp, i, j, k);
    if (fabs(p) < eps) {
      real delta = smoothed_delta(p, ooe, double_one);
      real x = A(px, bg, i, j, k), y = A(py, bg, i, j, k), z = A(pz, bg, i, j, k);
      real ng = (real)sqrt(x * x + y * y + z * z);
      if (F) acc = acc + delta * ng * A(F, bF, i, j, k) * dV;
      else acc = acc + delta * ng * dV;
    }
  }
  return acc;
}

void oracle_lsm3dsurfaceintegral_(real *out, const real *F, BOXARGS(F), const real *phi, BOXARGS(phi),
                                  const real *px, const real *py, const real *pz, BOXARGS(g), BOXARGS(ib),
                                  const real *dx, const real *dy, const real *dz, const real *eps) {
  /* lsm_utilities3d.f.in:745-846 */
  *out = surface_sum(F, MKBOX(F), phi, MKBOX(phi), px, py, pz, MKBOX(g), NULL, MKBOX(g), 0, ilo_ib, ihi_ib, jlo_ib,
                     jhi_ib, klo_ib, khi_ib, *dx * *dy * *dz, *eps, 1);
}
void oracle_lsm3dsurfaceintegralcontrolvolume_(real *out, const real *F, BOXARGS(F), const real *phi, BOXARGS(phi),
                                               const real *px, const real *py, const real *pz, BOXARGS(g),
                                               const real *cv, BOXARGS(cv), const int *cv_sgn, BOXARGS(ib),
                                               const real *dx, const real *dy, const real *dz, const real *eps) {
  /* lsm_utilities3d.f.in:1691-1844 */
  *out = surface_sum(F, MKBOX(F), phi, MKBOX(phi), px, py, pz, MKBOX(g), cv, MKBOX(cv), *cv_sgn, ilo_ib, ihi_ib,
                     jlo_ib, jhi_ib, klo_ib, khi_ib, *dx * *dy * *dz, *eps, 1);
}
void oracle_lsm3dsurfaceareazerolevelset_(real *out, const real *phi, BOXARGS(phi), const real *px, const real *py,
                                          const real *pz, BOXARGS(g), BOXARGS(ib), const real *dx, const real *dy,
                                          const real *dz, const real *eps) {
  /* lsm_geometry3d_fort.f.in:442-533 */
  *out = surface_sum(NULL, MKBOX(phi), phi, MKBOX(phi), px, py, pz, MKBOX(g), NULL, MKBOX(g), 0, ilo_ib, ihi_ib,
                     jlo_ib, jhi_ib, klo_ib, khi_ib, *dx * *dy * *dz, *eps, 0);
}
void oracle_lsm3dsurfaceareazerolevelsetcontrolvolume_(real *out, const real *phi, BOXARGS(phi), const real *px,
                                                       const real *py, const real *pz, BOXARGS(g), const real *cv,
                                                       BOXARGS(cv), const int *cv_sgn, BOXARGS(ib), const real *dx,
                                                       const real *dy, const real *dz, const real *eps) {
  /* lsm_geometry3d_fort.f.in:929-1072 */
  *out = surface_sum(NULL, MKBOX(phi), phi, MKBOX(phi), px, py, pz, MKBOX(g), cv, MKBOX(cv), *cv_sgn, ilo_ib,
                     ihi_ib, jlo_ib, jhi_ib, klo_ib, khi_ib, *dx * *dy * *dz, *eps, 0);
}

static real delta_area_sum(const real *dphi, box_t bd, const real *px, const real *py, const real *pz, box_t bg,
                           const real *cv, box_t bc, int cv_sgn, const int *ilo_ib, const int *ihi_ib,
                           const int *jlo_ib, const int *jhi_ib, const int *klo_ib, const int *khi_ib, real dV) {
  /* lsm_geometry3d_fort.f.in:537-613 and :1077-1197 */
  real area = (real)0.0;
  LOOP_IB {
    if (cv && !CV_IN(A(cv, bc, i, j, k), cv_sgn)) continue;
    real d = A(dphi, bd, i, j, k);
    if (d > 0.0) {
      real x = A(px, bg, i, j, k), y = A(py, bg, i, j, k), z = A(pz, bg, i, j, k);
      real ng = (real)sqrt(x * x + y * y + z * z);
      area = area + d * ng * dV;
    }
  }
  return area;
}

void oracle_lsm3dsurfaceareazerolevelsetdelta_(real *out, const real *dphi, BOXARGS(d), const real *px,
                                               const real *py, const real *pz, BOXARGS(g), BOXARGS(ib),
                                               const real *dx, const real *dy, const real *dz) {
  *out = delta_area_sum(dphi, MKBOX(d), px, py, pz, MKBOX(g), NULL, MKBOX(g), 0, ilo_ib, ihi_ib, jlo_ib, jhi_ib,
                        klo_ib, khi_ib, *dx * *dy * *dz);
}
void oracle_lsm3dsurfaceareazerolevelsetdeltacontrolvolume_(real *out, const real *dphi, BOXARGS(d), const real *px,
                                                            const real *py, const real *pz, BOXARGS(g),
                                                            const real *cv, BOXARGS(cv), const int *cv_sgn,
                                                            BOXARGS(ib), const real *dx, const real *dy,
                                                            const real *dz) {
  *out = delta_area_sum(dphi, MKBOX(d), px, py, pz, MKBOX(g), cv, MKBOX(cv), *cv_sgn, ilo_ib, ihi_ib, jlo_ib,
                        jhi_ib, klo_ib, khi_ib, *dx * *dy * *dz);
}

void oracle_lsm3dcomputestablenormalveldtcontrolvolume_(real *dt, const real *vn, BOXARGS(vel), const real *xp,
                                                        const real *yp, const real *zp, BOXARGS(gp), const real *xm,
                                                        const real *ym, const real *zm, BOXARGS(gm), const real *cv,
                                                        BOXARGS(cv), const int *cv_sgn, BOXARGS(ib),
                                                        const real *dx, const real *dy, const real *dz,
                                                        const real *cfl) {
  /* lsm_utilities3d.f.in:1080-1250 */
  box_t bv = MKBOX(vel), bp = MKBOX(gp), bm = MKBOX(gm), bc = MKBOX(cv);
  real mdx = (real)fmax2(fmax2(*dx, *dy), *dz);
  real max_dx_sq = mdx * mdx, mx = (real)-1.0;
  real ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  LOOP_IB {
    if (!CV_IN(A(cv, bc, i, j, k), *cv_sgn)) continue;
    real h = normal_h_over_dx((real)fabs(A(vn, bv, i, j, k)), A(xp, bp, i, j, k), A(yp, bp, i, j, k),
                              A(zp, bp, i, j, k), A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k),
                              max_dx_sq, ix, iy, iz);
    if (h > mx) mx = h;
  }
  *dt = *cfl / (mx + TINY);
}

void oracle_lsm3dcomputestableconstnormalveldtcontrolvolume_(real *dt, const real *vn, const real *xp,
                                                             const real *yp, const real *zp, BOXARGS(gp),
                                                             const real *xm, const real *ym, const real *zm,
                                                             BOXARGS(gm), const real *cv, BOXARGS(cv),
                                                             const int *cv_sgn, BOXARGS(ib), const real *dx,
                                                             const real *dy, const real *dz, const real *cfl) {
  /* lsm_utilities3d.f.in:1255-1415 */
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bc = MKBOX(cv);
  real mdx = (real)fmax2(fmax2(*dx, *dy), *dz);
  real max_dx_sq = mdx * mdx, mx = (real)-1.0;
  real ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz, avn = (real)fabs(*vn);
  LOOP_IB {
    if (!CV_IN(A(cv, bc, i, j, k), *cv_sgn)) continue;
    real h = normal_h_over_dx(avn, A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k), A(xm, bm, i, j, k),
                              A(ym, bm, i, j, k), A(zm, bm, i, j, k), max_dx_sq, ix, iy, iz);
    if (h > mx) mx = h;
  }
  *dt = *cfl / (mx + TINY);
}

void oracle_lsm3dvolumeintegralphilessthanzerocontrolvolume_(real *out, const real *F, BOXARGS(F), const real *phi,
                                                             BOXARGS(phi), const real *cv, BOXARGS(cv),
                                                             const int *cv_sgn, BOXARGS(ib), const real *dx,
                                                             const real *dy, const real *dz, const real *eps) {
  /* lsm_utilities3d.f.in:1420-1551 */
  box_t bF = MKBOX(F), bphi = MKBOX(phi), bc = MKBOX(cv);
  real dV = *dx * *dy * *dz, acc = (real)0.0, e = *eps;
  LOOP_IB {
    if (!CV_IN(A(cv, bc, i, j, k), *cv_sgn)) continue;
    real p = A(phi, bphi, i, j, k), x = p / e;
    if (p < -e) {
      acc = acc + A(F, bF, i, j, k) * dV;
    } else if (p < e) {
      real omh = 0.5 * (1.0 - x - ONE_OVER_PI_REF * SIN(PI_REF * x));
      acc = acc + omh * A(F, bF, i, j, k) * dV;
    }
  }
  *out = acc;
}

void oracle_lsm3dvolumeintegralphigreaterthanzerocontrolvolume_(real *out, const real *F, BOXARGS(F),
                                                                const real *phi, BOXARGS(phi), const real *cv,
                                                                BOXARGS(cv), const int *cv_sgn, BOXARGS(ib),
                                                                const real *dx, const real *dy, const real *dz,
                                                                const real *eps) {
  /* lsm_utilities3d.f.in:1556-1686 */
  box_t bF = MKBOX(F), bphi = MKBOX(phi), bc = MKBOX(cv);
  real dV = *dx * *dy * *dz, acc = (real)0.0, e = *eps;
  LOOP_IB {
    if (!CV_IN(A(cv, bc, i, j, k), *cv_sgn)) continue;
    real p = A(phi, bphi, i, j, k), x = p / e;
    if (p > e) {
      acc = acc + A(F, bF, i, j, k) * dV;
    } else if (p > -e) {
      real H = 0.5 * (1.0 + x + ONE_OVER_PI_REF * SIN(PI_REF * x));
      acc = acc + H * A(F, bF, i, j, k) * dV;
    }
  }
  *out = acc;
}

void oracle_lsm3dvolumeregionphilessthanzerocontrolvolume_(real *out, const real *phi, BOXARGS(phi), const real *cv,
                                                           BOXARGS(cv), const int *cv_sgn, BOXARGS(ib),
                                                           const real *dx, const real *dy, const real *dz,
                                                           const real *eps) {
  /* lsm_geometry3d_fort.f.in:638-760 */
  box_t bphi = MKBOX(phi), bc = MKBOX(cv);
  real dV = *dx * *dy * *dz, vol = (real)0.0, e = *eps;
  LOOP_IB {
    if (!CV_IN(A(cv, bc, i, j, k), *cv_sgn)) continue;
    real p = A(phi, bphi, i, j, k), x = p / e;
    if (p < -e) {
      vol = vol + dV;
    } else if (p < e) {
      real omh = 0.5 * (1 - x - ONE_OVER_PI_REF * SIN(PI_REF * x));
      vol = vol + omh * dV;
    }
  }
  *out = vol;
}

void oracle_lsm3dvolumeregionphigreaterthanzerocontrolvolume_(real *out, const real *phi, BOXARGS(phi),
                                                              const real *cv, BOXARGS(cv), const int *cv_sgn,
                                                              BOXARGS(ib), const real *dx, const real *dy,
                                                              const real *dz, const real *eps) {
  /* lsm_geometry3d_fort.f.in:784-903 */
  box_t bphi = MKBOX(phi), bc = MKBOX(cv);
  real dV = *dx * *dy * *dz, vol = (real)0.0, e = *eps;
  LOOP_IB {
    if (!CV_IN(A(cv, bc, i, j, k), *cv_sgn)) continue;
    real p = A(phi, bphi, i, j, k), x = p / e;
    if (p > e) {
      vol = vol + dV;
    } else if (p > -e) {
      real H = (real)0.5 * (1 + x + ONE_OVER_PI_REF * SIN(PI_REF * x));
      vol = vol + H * dV;
    }
  }
  *out = vol;
}

void oracle_lsm3ddeltafunctionorder1_(const real *phi, real *delta, BOXARGS(g), const real *nx, const real *ny,
                                      const real *nz, BOXARGS(n), BOXARGS(fb), const real *dx, const real *dy,
                                      const real *dz) {
  /* lsm_calculus_toolbox3d.f.in:15-120.  The z terms use norm_phi_x and
   * 1/dx, as the reference writes them (:98-110). */
  box_t bg = MKBOX(g), bn = MKBOX(n);
  real odx = 1.0 / *dx, ody = 1.0 / *dy;
  (void)dz;
  for (int k = *klo_fb; k <= *khi_fb; k++)
    for (int j = *jlo_fb; j <= *jhi_fb; j++)
      for (int i = *ilo_fb; i <= *ihi_fb; i++) {
        real c = A(phi, bg, i, j, k), t[6];
        const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
        for (int q = 0; q < 6; q++) {
          real nb = A(phi, bg, i + di[q], j + dj[q], k + dk[q]);
          const real *nf = (q < 2) ? nx : (q < 4) ? ny : nx;
          real f = (q >= 2 && q < 4) ? ody : odx;
          if (nb * c <= (real)0.0) {
            t[q] = nb * A(nf, bn, i, j, k) * f;
            t[q] = t[q] / (nb - c);
          } else {
            t[q] = (real)0.0;
          }
        }
        A(delta, bg, i, j, k) = (real)fabs(t[0]) + (real)fabs(t[1]) + (real)fabs(t[2]) + (real)fabs(t[3]) +
                                (real)fabs(t[4]) + (real)fabs(t[5]);
      }
}

/* ---------------------------------------------------------------------- */
/* Mean and Gaussian curvature: lsm_curvature3d.f.in                        */
/* ---------------------------------------------------------------------- */
/* Second derivatives of one cell.  Order 2 (:78-101): the pure ones are
 * computed in double (the 2.0d0 literal promotes the sum) and rounded, the
 * mixed ones in LSMLIB_REAL; the factors 1.d0/dx/dx and 0.25d0/dx/dy are
 * double quotients rounded to LSMLIB_REAL.  Order 4 (:318-349): sixteen,
 * thirty and eight are LSMLIB_REAL parameters, the mixed derivatives are
 * fourth-order differences of the given first derivatives (phi_x along y,
 * phi_y along z, phi_z along x). */
typedef struct {
  real xx, yy, zz, xy, xz, yz;
} hess_t;

static hess_t hessian(int order, const real *phi, box_t bp, const real *px, const real *py, const real *pz,
                      box_t bg, int i, int j, int k, double dx, double dy, double dz) {
  hess_t h;
#define P(a, b, c) A(phi, bp, a, b, c)
  if (order == 2) {
    real fxx = (real)(1.0 / dx / dx), fyy = (real)(1.0 / dy / dy), fzz = (real)(1.0 / dz / dz);
    real fxy = (real)(0.25 / dx / dy), fxz = (real)(0.25 / dx / dz), fyz = (real)(0.25 / dy / dz);
    h.xx = (real)(((double)P(i + 1, j, k) - 2.0 * (double)P(i, j, k) + (double)P(i - 1, j, k)) * (double)fxx);
    h.yy = (real)(((double)P(i, j + 1, k) - 2.0 * (double)P(i, j, k) + (double)P(i, j - 1, k)) * (double)fyy);
    h.zz = (real)(((double)P(i, j, k + 1) - 2.0 * (double)P(i, j, k) + (double)P(i, j, k - 1)) * (double)fzz);
    h.xy = (P(i + 1, j + 1, k) - P(i + 1, j - 1, k) - P(i - 1, j + 1, k) + P(i - 1, j - 1, k)) * fxy;
    h.xz = (P(i + 1, j, k + 1) - P(i + 1, j, k - 1) - P(i - 1, j, k + 1) + P(i - 1, j, k - 1)) * fxz;
    h.yz = (P(i, j + 1, k + 1) - P(i, j + 1, k - 1) - P(i, j - 1, k + 1) + P(i, j - 1, k - 1)) * fyz;
  } else {
    const double c12 = 0.0833333333333333333333;
    real f2x = (real)(c12 / dx / dx), f2y = (real)(c12 / dy / dy), f2z = (real)(c12 / dz / dz);
    real f1x = (real)(c12 / dx), f1y = (real)(c12 / dy), f1z = (real)(c12 / dz);
    const real s16 = (real)16.0, s30 = (real)30.0, e8 = (real)8.0;
    h.xx = (-P(i + 2, j, k) + s16 * P(i + 1, j, k) - s30 * P(i, j, k) - P(i - 2, j, k) + s16 * P(i - 1, j, k)) * f2x;
    h.yy = (-P(i, j + 2, k) + s16 * P(i, j + 1, k) - s30 * P(i, j, k) - P(i, j - 2, k) + s16 * P(i, j - 1, k)) * f2y;
    h.zz = (-P(i, j, k + 2) + s16 * P(i, j, k + 1) - s30 * P(i, j, k) - P(i, j, k - 2) + s16 * P(i, j, k - 1)) * f2z;
    h.xy = (-A(px, bg, i, j + 2, k) + e8 * A(px, bg, i, j + 1, k) + A(px, bg, i, j - 2, k) -
            e8 * A(px, bg, i, j - 1, k)) * f1y;
    h.yz = (-A(py, bg, i, j, k + 2) + e8 * A(py, bg, i, j, k + 1) + A(py, bg, i, j, k - 2) -
            e8 * A(py, bg, i, j, k - 1)) * f1z;
    h.xz = (-A(pz, bg, i + 2, j, k) + e8 * A(pz, bg, i + 1, j, k) + A(pz, bg, i - 2, j, k) -
            e8 * A(pz, bg, i - 1, j, k)) * f1x;
  }
#undef P
  return h;
}

/* kappa = (sum of the three 2x2 terms) / |grad phi|^3 (mean, :103-120) or
 * (the cofactor expansion) / |grad phi|^4 (Gaussian, :224-239); 0 where
 * that power of |grad phi| is below zero_tol. */
static void curvature(int order, int gauss, real *kappa, box_t bk, const real *phi, box_t bp, const real *px,
                      const real *py, const real *pz, const real *gm, box_t bg, int ilo, int ihi, int jlo, int jhi,
                      int klo, int khi, double dx, double dy, double dz) {
  for (int k = klo; k <= khi; k++)
    for (int j = jlo; j <= jhi; j++)
      for (int i = ilo; i <= ihi; i++) {
        hess_t h = hessian(order, phi, bp, px, py, pz, bg, i, j, k, dx, dy, dz);
        real x = A(px, bg, i, j, k), y = A(py, bg, i, j, k), z = A(pz, bg, i, j, k), g = A(gm, bg, i, j, k);
        real den = gauss ? g * g * g * g : g * g * g;
        if (den < ZERO_TOL) {
          A(kappa, bk, i, j, k) = (real)0.0;
          continue;
        }
        real t;
        if (gauss)
          t = x * x * (h.yy * h.zz - h.yz * h.yz) + y * y * (h.xx * h.zz - h.xz * h.xz) +
              z * z * (h.xx * h.yy - h.xy * h.xy) +
              2 * (x * y * (h.xz * h.yz - h.xy * h.zz) + y * z * (h.xy * h.xz - h.yz * h.xx) +
                   x * z * (h.xy * h.yz - h.xz * h.yy));
        else
          t = h.xx * y * y + h.yy * x * x - 2 * h.xy * x * y + h.xx * z * z + h.zz * x * x - 2 * h.xz * x * z +
              h.yy * z * z + h.zz * y * y - 2 * h.yz * y * z;
        A(kappa, bk, i, j, k) = t / den;
      }
}

#define CURV_ARGS                                                                                        \
  real *kappa, BOXARGS(k), const real *phi, BOXARGS(phi), const real *px, const real *py, const real *pz, \
      const real *gm, BOXARGS(g)
#define CURV_CALL(order, gauss)                                                                             \
  curvature(order, gauss, kappa, MKBOX(k), phi, MKBOX(phi), px, py, pz, gm, MKBOX(g), *ilo_fb, *ihi_fb, *jlo_fb, \
            *jhi_fb, *klo_fb, *khi_fb, *dx, *dy, *dz)

void oracle_lsm3dcomputemeancurvatureorder2_(CURV_ARGS, BOXARGS(fb), const real *dx, const real *dy,
                                             const real *dz) {
  /* lsm_curvature3d.f.in:15-129 */
  CURV_CALL(2, 0);
}
void oracle_lsm3dcomputegaussiancurvatureorder2_(CURV_ARGS, BOXARGS(fb), const real *dx, const real *dy,
                                                 const real *dz) {
  /* lsm_curvature3d.f.in:136-248 */
  CURV_CALL(2, 1);
}
void oracle_lsm3dcomputemeancurvatureorder4_(CURV_ARGS, BOXARGS(fb), const real *dx, const real *dy,
                                             const real *dz) {
  /* lsm_curvature3d.f.in:253-376 */
  CURV_CALL(4, 0);
}
void oracle_lsm3dcomputegaussiancurvatureorder4_(CURV_ARGS, BOXARGS(fb), const real *dx, const real *dy,
                                                 const real *dz) {
  /* lsm_curvature3d.f.in:383-507 */
  CURV_CALL(4, 1);
}

void oracle_lsm3daddconstprecomputedcurvtermtolserhs_(real *rhs, BOXARGS(r), const real *kappa, BOXARGS(k),
                                                      const real *gm, BOXARGS(g), const real *b, BOXARGS(fb)) {
  /* lsm_level_set_evolution3d.f.in:465-526.  The loops run k from the fill
   * box's klo to the ghost box's khi, j likewise, and i over the whole ghost
   * box (:514-516), as the reference writes them. */
  box_t br = MKBOX(r), bk = MKBOX(k), bg = MKBOX(g);
  for (int k = *klo_fb; k <= *khi_r; k++)
    for (int j = *jlo_fb; j <= *jhi_r; j++)
      for (int i = *ilo_r; i <= *ihi_r; i++)
        A(rhs, br, i, j, k) = A(rhs, br, i, j, k) + *b * A(kappa, bk, i, j, k) * A(gm, bg, i, j, k);
}

/* ---------------------------------------------------------------------- */
/* Narrow band (localization): lsm_localization3d.f and the *LOCAL kernels  */
/* Index lists index_x/y/z(nlo:nhi) and the integer*1 narrow_band marks     */
/* (signed in the Fortran; compared as signed char here).                   */
/* ---------------------------------------------------------------------- */
typedef signed char nb_t;
#define LIDX(p, l, nlo) (p)[(l) - (nlo)]

/* lsm_localization3d.f:48-168; an empty previous level (n_lo = n_hi = -1)
 * has no parents here (the reference would read index -1) */
static void mark_nb_neighbors(nb_t *nb, box_t bn, int *ix, int *iy, int *iz, int nlo, int *n_lo, int *n_hi,
                              int level) {
  int count = n_hi[0] + 1;
  for (int l = 1; l <= level; l++) {
    nb_t mark = (nb_t)(l + 1);
    if (n_lo[l - 1] >= nlo)
      for (int m = n_lo[l - 1]; m <= n_hi[l - 1]; m++) {
        int i = LIDX(ix, m, nlo), j = LIDX(iy, m, nlo), k = LIDX(iz, m, nlo);
        const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
        for (int q = 0; q < 6; q++) {
          int a = i + di[q], b = j + dj[q], c = k + dk[q];
          int inside = (q == 0) ? (i < bn.ihi) : (q == 1) ? (i > bn.ilo) : (q == 2) ? (j < bn.jhi)
                     : (q == 3) ? (j > bn.jlo) : (q == 4) ? (k < bn.khi) : (k > bn.klo);
          if (inside && A(nb, bn, a, b, c) == 0) {
            LIDX(ix, count, nlo) = a; LIDX(iy, count, nlo) = b; LIDX(iz, count, nlo) = c;
            A(nb, bn, a, b, c) = mark;
            count++;
          }
        }
      }
    if (count > n_hi[l - 1] + 1) { n_lo[l] = n_hi[l - 1] + 1; n_hi[l] = count - 1; }
    else { n_lo[l] = -1; n_hi[l] = -1; }
  }
}

void oracle_lsm3ddeterminenarrowband_(const real *phi, BOXARGS(g), nb_t *nb, BOXARGS(nb), int *ix, int *iy,
                                      int *iz, const int *nlo_index, const int *nhi_index, int *n_lo, int *n_hi,
                                      int *index_outer, const int *nlo_index_outer, const int *nhi_index_outer,
                                      int *nlo_outer_plus, int *nhi_outer_plus, int *nlo_outer_minus,
                                      int *nhi_outer_minus, const real *width, const real *width_inner,
                                      const int *level) {
  /* lsm_localization3d.f:174-298 */
  box_t bg = MKBOX(g), bn = MKBOX(nb);
  (void)nhi_index;
  int nlo = *nlo_index, count = nlo, no = *nlo_index_outer;
  n_lo[0] = nlo;
  int cm = *nlo_index_outer, cp = *nhi_index_outer;
  *nlo_outer_minus = *nlo_index_outer;
  *nhi_outer_plus = *nhi_index_outer;
  for (int k = bg.klo; k <= bg.khi; k++)
    for (int j = bg.jlo; j <= bg.jhi; j++)
      for (int i = bg.ilo; i <= bg.ihi; i++) {
        real p = A(phi, bg, i, j, k), a = (real)fabs(p);
        if (a < *width) {
          LIDX(ix, count, nlo) = i; LIDX(iy, count, nlo) = j; LIDX(iz, count, nlo) = k;
          A(nb, bn, i, j, k) = 1;
          if (a >= *width_inner) {
            if (p <= 0.0) { index_outer[cm - no] = count; cm++; }
            else { index_outer[cp - no] = count; cp--; }
          }
          count++;
        } else {
          A(nb, bn, i, j, k) = 0;
        }
      }
  if (count > nlo) {
    n_hi[0] = count - 1;
    *nhi_outer_minus = cm - 1;
    *nlo_outer_plus = cp + 1;
    mark_nb_neighbors(nb, bn, ix, iy, iz, nlo, n_lo, n_hi, *level);
  } else {
    n_hi[0] = count;
    *nhi_outer_minus = cm;
    *nlo_outer_plus = cp;
  }
}

void oracle_lsm3ddeterminenarrowbandfrommask_(const real *mask, BOXARGS(g), nb_t *nb, BOXARGS(nb), int *ix,
                                              int *iy, int *iz, const int *nlo_index, const int *nhi_index,
                                              int *n_lo, int *n_hi, const int *level, const int *use_mask_sign) {
  /* lsm_localization3d.f:385-476 */
  box_t bg = MKBOX(g), bn = MKBOX(nb);
  (void)nhi_index;
  int nlo = *nlo_index, count = nlo;
  n_lo[0] = nlo;
  for (int k = bg.klo; k <= bg.khi; k++)
    for (int j = bg.jlo; j <= bg.jhi; j++)
      for (int i = bg.ilo; i <= bg.ihi; i++) {
        real m = A(mask, bg, i, j, k);
        int in = (*use_mask_sign > 0) ? (m >= 0) : (m <= 0);
        if (in) {
          LIDX(ix, count, nlo) = i; LIDX(iy, count, nlo) = j; LIDX(iz, count, nlo) = k;
          A(nb, bn, i, j, k) = 1;
          count++;
        } else {
          A(nb, bn, i, j, k) = 0;
        }
      }
  n_hi[0] = count - 1;
  mark_nb_neighbors(nb, bn, ix, iy, iz, nlo, n_lo, n_hi, *level);
}

void oracle_lsm3dmarknarrowbandboundarylayer_(nb_t *nb, BOXARGS(g), BOXARGS(fb), const nb_t *mark) {
  /* lsm_localization3d.f:303-378: the six fill-box face layers, each across the ghost box */
  box_t bg = MKBOX(g);
  for (int k = bg.klo; k <= bg.khi; k++)
    for (int j = bg.jlo; j <= bg.jhi; j++)
      for (int i = bg.ilo; i <= bg.ihi; i++)
        if (i == *ilo_fb || i == *ihi_fb || j == *jlo_fb || j == *jhi_fb || k == *klo_fb || k == *khi_fb)
          A(nb, bg, i, j, k) = *mark;
}

/* loop over index_*(nlo..nhi); (i, j, k) of entry l */
#define FOR_LIST(nlo_p, nhi_p)                                                 \
  for (int l = *(nlo_p); l <= *(nhi_p); l++) {                                 \
    int i = LIDX(ix, l, *(nlo_p)), j = LIDX(iy, l, *(nlo_p)), k = LIDX(iz, l, *(nlo_p));
#define END_LIST }
#define XYZ const int *ix, const int *iy, const int *iz, const int *nlo_index, const int *nhi_index
#define NB const nb_t *nb, BOXARGS(nb)

void oracle_lsm3dmultiplycutofflserhslocal_(const real *phi, real *rhs, BOXARGS(g), XYZ, NB, const nb_t *mark_fb,
                                            const real *beta, const real *gamma) {
  /* lsm_localization3d.f:482-562 */
  box_t bg = MKBOX(g), bn = MKBOX(nb);
  const real b = *beta, gm = *gamma;
  real c1 = gm - 3 * b;
  real c2 = (gm - b);
  c2 = c2 * c2 * c2;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    real a = (real)fabs(A(phi, bg, i, j, k)), cut;
    if (a <= b) cut = 1;
    else if (a <= gm) {
      real t = (a - gm);
      cut = (t * t * (2 * a + c1)) / c2;
    } else cut = 0;
    A(rhs, bg, i, j, k) = cut * A(rhs, bg, i, j, k);
  }
  END_LIST
}

void oracle_lsm3dcheckouternarrowbandlayer_(int *change_sign, const real *phi, BOXARGS(g), XYZ,
                                            const int *index_outer, const int *nlo_outer, const int *nhi_outer,
                                            const int *nlo_plus, const int *nhi_plus, const int *nlo_minus,
                                            const int *nhi_minus) {
  /* lsm_localization3d.f:568-654 */
  box_t bg = MKBOX(g);
  (void)nhi_index; (void)nhi_outer;
  int nlo = *nlo_index, no = *nlo_outer;
  *change_sign = 0;
  for (int pass = 0; pass < 2 && !*change_sign; pass++) {
    int lo = pass ? *nlo_minus : *nlo_plus, hi = pass ? *nhi_minus : *nhi_plus;
    int sp = 0, sm = 0;
    for (int m = lo; m <= hi; m++) {
      int l = index_outer[m - no];
      real p = A(phi, bg, LIDX(ix, l, nlo), LIDX(iy, l, nlo), LIDX(iz, l, nlo));
      if (p > 0.0) sp = 1; else sm = 1;
      if (sp && sm) { *change_sign = 1; break; }
    }
  }
}

void oracle_lsm3dimposemasklocal_(real *dest, BOXARGS(d), const real *src, BOXARGS(s), const real *mask,
                                  BOXARGS(m), XYZ) {
  /* lsm_localization3d.f:660-718: max(mask, src) */
  box_t bd = MKBOX(d), bs = MKBOX(s), bm = MKBOX(m);
  FOR_LIST(nlo_index, nhi_index)
  real a = A(mask, bm, i, j, k), b = A(src, bs, i, j, k);
  A(dest, bd, i, j, k) = (a > b) ? a : b;
  END_LIST
}

void oracle_lsm3dcopydatalocal_(real *dest, BOXARGS(d), const real *src, BOXARGS(s), XYZ) {
  /* lsm_localization3d.f:724-776 */
  box_t bd = MKBOX(d), bs = MKBOX(s);
  FOR_LIST(nlo_index, nhi_index)
  A(dest, bd, i, j, k) = A(src, bs, i, j, k);
  END_LIST
}

/* TVD Runge-Kutta on the band (lsm_tvd_runge_kutta3d_local.f) */
#define RK_LOCAL_1(name)                                                                              \
  void oracle_##name(real *un, BOXARGS(un), const real *uc, BOXARGS(uc), const real *rhs, BOXARGS(r), \
                     const real *dt, XYZ, NB, const nb_t *mark_fb) {                                  \
    box_t bu = MKBOX(un), bc = MKBOX(uc), br = MKBOX(r), bn = MKBOX(nb);                              \
    FOR_LIST(nlo_index, nhi_index)                                                                    \
    if (A(nb, bn, i, j, k) <= *mark_fb) A(un, bu, i, j, k) = A(uc, bc, i, j, k) + *dt * A(rhs, br, i, j, k); \
    END_LIST                                                                                          \
  }
RK_LOCAL_1(lsm3drk1steplocal_)          /* :34-131 */
RK_LOCAL_1(lsm3dtvdrk2stage1local_)     /* :137-237 (calls the RK1 step) */
RK_LOCAL_1(lsm3dtvdrk3stage1local_)     /* :357-459 */

void oracle_lsm3dtvdrk2stage2local_(real *un, BOXARGS(un), const real *u1, BOXARGS(u1), const real *uc,
                                    BOXARGS(uc), const real *rhs, BOXARGS(r), const real *dt, XYZ, NB,
                                    const nb_t *mark_fb) {
  /* :243-351 */
  box_t bu = MKBOX(un), b1 = MKBOX(u1), bc = MKBOX(uc), br = MKBOX(r), bn = MKBOX(nb);
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb)
    A(un, bu, i, j, k) = 0.5 * (A(uc, bc, i, j, k) + A(u1, b1, i, j, k) + *dt * A(rhs, br, i, j, k));
  END_LIST
}

void oracle_lsm3dtvdrk3stage2local_(real *u2, BOXARGS(u2), const real *u1, BOXARGS(u1), const real *uc,
                                    BOXARGS(uc), const real *rhs, BOXARGS(r), const real *dt, XYZ, NB,
                                    const nb_t *mark_fb) {
  /* :465-569 */
  box_t bu = MKBOX(u2), b1 = MKBOX(u1), bc = MKBOX(uc), br = MKBOX(r), bn = MKBOX(nb);
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb)
    A(u2, bu, i, j, k) = 0.75 * A(uc, bc, i, j, k) + 0.25 * (A(u1, b1, i, j, k) + *dt * A(rhs, br, i, j, k));
  END_LIST
}

void oracle_lsm3dtvdrk3stage3local_(real *un, BOXARGS(un), const real *u2, BOXARGS(u2), const real *uc,
                                    BOXARGS(uc), const real *rhs, BOXARGS(r), const real *dt, XYZ, NB,
                                    const nb_t *mark_fb) {
  /* :575-667; one_third, two_thirds are real variables set from d0 literals */
  box_t bu = MKBOX(un), b2 = MKBOX(u2), bc = MKBOX(uc), br = MKBOX(r), bn = MKBOX(nb);
  const real one_third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb)
    A(un, bu, i, j, k) = one_third * A(uc, bc, i, j, k) + two_thirds * (A(u2, b2, i, j, k) + *dt * A(rhs, br, i, j, k));
  END_LIST
}

void oracle_lsm3dzerooutlevelseteqnrhslocal_(real *rhs, BOXARGS(r), XYZ) {
  /* lsm_level_set_evolution3d_local.f.in:16-55 */
  box_t br = MKBOX(r);
  FOR_LIST(nlo_index, nhi_index)
  A(rhs, br, i, j, k) = (real)0.0;
  END_LIST
}

void oracle_lsm3daddconstnormalveltermtolserhslocal_(real *rhs, BOXARGS(r), const real *xp, const real *yp,
                                                     const real *zp, BOXARGS(gp), const real *xm, const real *ym,
                                                     const real *zm, BOXARGS(gm), const real *vn, XYZ, NB,
                                                     const nb_t *mark_fb) {
  /* :282-398 (the Godunov norm is formed for every band cell; the term is
   * added when |vn| >= zero_tol) */
  box_t br = MKBOX(r), bp = MKBOX(gp), bm = MKBOX(gm), bn = MKBOX(nb);
  real v = *vn;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    real n2 = godunov_norm_sq(v, A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k), A(xm, bm, i, j, k),
                              A(ym, bm, i, j, k), A(zm, bm, i, j, k));
    if (fabs(v) >= ZERO_TOL) A(rhs, br, i, j, k) = A(rhs, br, i, j, k) - v * (real)sqrt(n2);
  }
  END_LIST
}

void oracle_lsm3daddconstcurvtermtolserhslocal_(real *rhs, BOXARGS(r), const real *px, const real *py,
                                                const real *pz, BOXARGS(g1), const real *pxx, const real *pxy,
                                                const real *pxz, const real *pyy, const real *pyz,
                                                const real *pzz, BOXARGS(g2), const real *b, XYZ, NB,
                                                const nb_t *mark_fb) {
  /* :404-524, the per-cell term of lsm3dAddConstCurvTermToLSERHS */
  box_t br = MKBOX(r), b1 = MKBOX(g1), b2 = MKBOX(g2), bn = MKBOX(nb);
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    real x = A(px, b1, i, j, k), y = A(py, b1, i, j, k), z = A(pz, b1, i, j, k);
    real xx = A(pxx, b2, i, j, k), xy = A(pxy, b2, i, j, k), xz = A(pxz, b2, i, j, k);
    real yy = A(pyy, b2, i, j, k), yz = A(pyz, b2, i, j, k), zz = A(pzz, b2, i, j, k);
    real g2 = x * x + y * y + z * z, curv;
    if (g2 < ZERO_TOL) curv = (real)0.0;
    else {
      curv = xx * y * y + yy * x * x - (real)2 * xy * x * y + xx * z * z + zz * x * x -
             (real)2 * xz * x * z + yy * z * z + zz * y * y - (real)2 * yz * y * z;
      curv = curv / g2;
    }
    A(rhs, br, i, j, k) = A(rhs, br, i, j, k) + *b * curv;
  }
  END_LIST
}

/* lsm3dComputeDnLOCAL, lsm_spatial_derivatives3d_local.f.in:43-164 */
static void compute_dn_local(real *Dn, box_t bd, const real *Dm, box_t bm, int n, int dir, const int *ix,
                             const int *iy, const int *iz, int nlo_list, int lo, int hi, const nb_t *nb,
                             box_t bn, nb_t mark) {
  int off[3] = {0, 0, 0};
  real sgn;
  if (n % 2 == 1) { off[dir] = 1; sgn = (real)1.0; }
  else { off[dir] = -1; sgn = (real)-1.0; }
  for (int l = lo; l <= hi; l++) {
    int i = LIDX(ix, l, nlo_list), j = LIDX(iy, l, nlo_list), k = LIDX(iz, l, nlo_list);
    if (A(nb, bn, i, j, k) <= mark)
      A(Dn, bd, i, j, k) = sgn * (A(Dm, bm, i, j, k) - A(Dm, bm, i - off[0], j - off[1], k - off[2]));
    else
      A(Dn, bd, i, j, k) = (real)1.e10;
  }
}

void oracle_lsm3dhjeno2local_(real *xp, real *yp, real *zp, BOXARGS(gp), real *xm, real *ym, real *zm,
                              BOXARGS(gm), const real *phi, BOXARGS(phi), real *D1, BOXARGS(d1), real *D2,
                              BOXARGS(d2), const real *dx, const real *dy, const real *dz, const int *ix,
                              const int *iy, const int *iz, const int *nlo0, const int *nhi0, const int *nlo1,
                              const int *nhi1, const int *nlo2, const int *nhi2, NB, const nb_t *mark_fb,
                              const nb_t *mark_d1, const nb_t *mark_d2) {
  /* lsm_spatial_derivatives3d_local.f.in:414-757: per direction, D1 on levels
   * 0-2, D2 on levels 0-1, the ENO2 choice on level 0 */
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bphi = MKBOX(phi), b1 = MKBOX(d1), b2 = MKBOX(d2), bn = MKBOX(nb);
  (void)nlo1; (void)nlo2;
  const real half = (real)0.5;
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *outp[3] = {xp, yp, zp}, *outm[3] = {xm, ym, zm};
  int nlo = *nlo0;
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    compute_dn_local(D1, b1, phi, bphi, 1, d, ix, iy, iz, nlo, *nlo0, *nhi2, nb, bn, *mark_d1);
    compute_dn_local(D2, b2, D1, b1, 2, d, ix, iy, iz, nlo, *nlo0, *nhi1, nb, bn, *mark_d2);
    for (int l = *nlo0; l <= *nhi0; l++) {
      int i = LIDX(ix, l, nlo), j = LIDX(iy, l, nlo), k = LIDX(iz, l, nlo);
      if (A(nb, bn, i, j, k) <= *mark_fb) {
        real d2m = A(D2, b2, i - a, j - b, k - c), d20 = A(D2, b2, i, j, k), d2p = A(D2, b2, i + a, j + b, k + c);
        if (fabs(d20) < fabs(d2p)) A(outp[d], bp, i, j, k) = (A(D1, b1, i + a, j + b, k + c) - half * d20) * inv[d];
        else A(outp[d], bp, i, j, k) = (A(D1, b1, i + a, j + b, k + c) - half * d2p) * inv[d];
        if (fabs(d2m) < fabs(d20)) A(outm[d], bm, i, j, k) = (A(D1, b1, i, j, k) + half * d2m) * inv[d];
        else A(outm[d], bm, i, j, k) = (A(D1, b1, i, j, k) + half * d20) * inv[d];
      }
    }
  }
}

void oracle_lsm3dcentralgradorder2local_(real *gx, real *gy, real *gz, BOXARGS(g), const real *phi, BOXARGS(phi),
                                         const real *dx, const real *dy, const real *dz, XYZ, NB,
                                         const nb_t *mark_fb) {
  /* lsm_spatial_derivatives3d_local.f.in:763-862 */
  box_t bg = MKBOX(g), bp = MKBOX(phi), bn = MKBOX(nb);
  real fx = 0.5 / *dx, fy = 0.5 / *dy, fz = 0.5 / *dz;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    A(gx, bg, i, j, k) = (A(phi, bp, i + 1, j, k) - A(phi, bp, i - 1, j, k)) * fx;
    A(gy, bg, i, j, k) = (A(phi, bp, i, j + 1, k) - A(phi, bp, i, j - 1, k)) * fy;
    A(gz, bg, i, j, k) = (A(phi, bp, i, j, k + 1) - A(phi, bp, i, j, k - 1)) * fz;
  }
  END_LIST
}

void oracle_lsm3dcomputeavegradphilocal_(real *ave, const real *phi, BOXARGS(phi), const real *dx, const real *dy,
                                         const real *dz, XYZ, NB, const nb_t *mark_fb) {
  /* lsm_spatial_derivatives3d_local.f.in:1089-1180: the sum runs in list order */
  box_t bp = MKBOX(phi), bn = MKBOX(nb);
  real fx = 0.5 / *dx, fy = 0.5 / *dy, fz = 0.5 / *dz, s = (real)0.0;
  int count = 0;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    real x = (A(phi, bp, i + 1, j, k) - A(phi, bp, i - 1, j, k)) * fx;
    real y = (A(phi, bp, i, j + 1, k) - A(phi, bp, i, j - 1, k)) * fy;
    real z = (A(phi, bp, i, j, k + 1) - A(phi, bp, i, j, k - 1)) * fz;
    s = s + (real)sqrt(x * x + y * y + z * z);
    count++;
  }
  END_LIST
  if (count > 0) s = s / (real)count;
  *ave = s;
}

void oracle_lsm3dmaxnormdifflocal_(real *out, const real *f1, BOXARGS(f1), const real *f2, BOXARGS(f2), XYZ, NB,
                                   const nb_t *mark_fb) {
  /* lsm_utilities3d_local.f.in:16-92 */
  box_t b1 = MKBOX(f1), b2 = MKBOX(f2), bn = MKBOX(nb);
  real m = (real)0.0;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    real d = (real)fabs(A(f1, b1, i, j, k) - A(f2, b2, i, j, k));
    if (d > m) m = d;
  }
  END_LIST
  *out = m;
}

void oracle_lsm3dcomputestableconstnormalveldtlocal_(real *dt, const real *vn, const real *xp, const real *yp,
                                                     const real *zp, BOXARGS(gp), const real *xm,
                                                     const real *ym, const real *zm, BOXARGS(gm),
                                                     const real *dx, const real *dy, const real *dz,
                                                     const real *cfl, XYZ, NB, const nb_t *mark_fb) {
  /* lsm_utilities3d_local.f.in:628-753.  The reference never assigns
   * abs_vel_n and norm_grad_phi here (they are read uninitialised, :108);
   * this takes the values of lsm3dComputeStableConstNormalVelDt (|vel_n| and
   * sqrt(phi_x^2 + phi_y^2 + phi_z^2 + max_dx^2), lsm_utilities3d.f.in:520-550). */
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bn = MKBOX(nb);
  real m = *dx;
  if (*dy > m) m = *dy;
  if (*dz > m) m = *dz;
  real mdx2 = m * m, mx = (real)-1.0;
  real ix_ = 1.0 / *dx, iy_ = 1.0 / *dy, iz_ = 1.0 / *dz, avn = (real)fabs(*vn);
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    real h = normal_h_over_dx(avn, A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k),
                              A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k), mdx2, ix_, iy_, iz_);
    if (h > mx) mx = h;
  }
  END_LIST
  *dt = *cfl / (mx + TINY);
}

void oracle_lsm3dcomputereinitializationeqnrhslocal_(real *rhs, BOXARGS(r), const real *phi, BOXARGS(phi),
                                                     const real *phi0, BOXARGS(p0), const real *xp, const real *yp,
                                                     const real *zp, BOXARGS(gp), const real *xm, const real *ym,
                                                     const real *zm, BOXARGS(gm), const real *dx, const real *dy,
                                                     const real *dz, const int *use_phi0, XYZ, NB,
                                                     const nb_t *mark_fb) {
  /* lsm_reinitialization3d_local.f.in:50-320, the per-cell rule of
   * lsm3dComputeReinitializationEqnRHS on the band */
  box_t br = MKBOX(r), bphi = MKBOX(phi), b0 = MKBOX(p0), bp = MKBOX(gp), bm = MKBOX(gm), bn = MKBOX(nb);
  real m = *dx;
  if (*dy > m) m = *dy;
  if (*dz > m) m = *dz;
  real dx_sq = m * m;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= *mark_fb) {
    real pc = (*use_phi0 == 1) ? A(phi0, b0, i, j, k) : A(phi, bphi, i, j, k);
    real gp3[3] = {A(xp, bp, i, j, k), A(yp, bp, i, j, k), A(zp, bp, i, j, k)};
    real gm3[3] = {A(xm, bm, i, j, k), A(ym, bm, i, j, k), A(zm, bm, i, j, k)}, gs[3];
    godunov_star(pc, gp3, gm3, gs);
    if (fabs(pc) >= ZERO_TOL) {
      real n2 = gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2];
      real sgn = (*use_phi0 == 1) ? pc / (real)sqrt(pc * pc + dx_sq) : pc / (real)sqrt(pc * pc + n2 * dx_sq);
      A(rhs, br, i, j, k) = sgn * ((real)1.0 - (real)sqrt(n2));
    } else {
      A(rhs, br, i, j, k) = (real)0.0;
    }
  }
  END_LIST
}

/* The remaining *LOCAL kernels.  Where a local routine's per-cell rule is
 * the box routine's (checked against the reference build), it is applied as
 * that routine over the one-cell fill box (i,i,j,j,k,k) of each marked entry. */
#define CELL_BOX &cb[0], &cb[1], &cb[2], &cb[3], &cb[4], &cb[5]
#define FWD(p) ilo_##p, ihi_##p, jlo_##p, jhi_##p, klo_##p, khi_##p
#define FOR_MARKED(nlo_p, nhi_p)                                                       \
  box_t bn_ = MKBOX(nb);                                                               \
  FOR_LIST(nlo_p, nhi_p)                                                               \
  if (A(nb, bn_, i, j, k) > *mark_fb) continue;                                        \
  int cb[6] = {i, i, j, j, k, k};                                                      \
  (void)cb;

void oracle_lsm3dhjeno1local_(real *xp, real *yp, real *zp, BOXARGS(gp), real *xm, real *ym, real *zm,
                              BOXARGS(gm), const real *phi, BOXARGS(phi), real *D1, BOXARGS(d1),
                              const real *dx, const real *dy, const real *dz, const int *ix, const int *iy,
                              const int *iz, const int *nlo0, const int *nhi0, const int *nlo1,
                              const int *nhi1, NB, const nb_t *mark_fb, const nb_t *mark_d1) {
  /* lsm_spatial_derivatives3d_local.f.in:170-404: per direction, D1 on
   * levels 0-1, then plus = D1(i+1)/dx, minus = D1(i)/dx on level 0 */
  box_t bp = MKBOX(gp), bm = MKBOX(gm), bphi = MKBOX(phi), b1 = MKBOX(d1), bn = MKBOX(nb);
  (void)nlo1;
  real inv[3];
  inv[0] = 1.0 / *dx; inv[1] = 1.0 / *dy; inv[2] = 1.0 / *dz;
  real *outp[3] = {xp, yp, zp}, *outm[3] = {xm, ym, zm};
  int nlo = *nlo0;
  for (int d = 0; d < 3; d++) {
    int a = DI(d), b = DJ(d), c = DK(d);
    compute_dn_local(D1, b1, phi, bphi, 1, d, ix, iy, iz, nlo, *nlo0, *nhi1, nb, bn, *mark_d1);
    for (int l = *nlo0; l <= *nhi0; l++) {
      int i = LIDX(ix, l, nlo), j = LIDX(iy, l, nlo), k = LIDX(iz, l, nlo);
      if (A(nb, bn, i, j, k) <= *mark_fb) {
        A(outp[d], bp, i, j, k) = A(D1, b1, i + a, j + b, k + c) * inv[d];
        A(outm[d], bm, i, j, k) = A(D1, b1, i, j, k) * inv[d];
      }
    }
  }
}

void oracle_lsm3dcentralgradorder4local_(real *gx, real *gy, real *gz, BOXARGS(g), const real *phi,
                                         BOXARGS(phi), const real *dx, const real *dy, const real *dz, XYZ, NB,
                                         const nb_t *mark_fb) {
  /* lsm_spatial_derivatives3d_local.f.in:868-960 */
  FOR_MARKED(nlo_index, nhi_index)
  oracle_lsm3dcentralgradorder4_(gx, gy, gz, FWD(g), phi, FWD(phi), CELL_BOX, dx, dy, dz);
  END_LIST
}

void oracle_lsm3dlaplacianorder2local_(real *lap, BOXARGS(l), const real *phi, BOXARGS(phi), const real *dx,
                                       const real *dy, const real *dz, XYZ, NB, const nb_t *mark_fb) {
  /* lsm_spatial_derivatives3d_local.f.in:966-1060 */
  FOR_MARKED(nlo_index, nhi_index)
  oracle_lsm3dlaplacianorder2_(lap, FWD(l), phi, FWD(phi), CELL_BOX, dx, dy, dz);
  END_LIST
}

#define CURV_LOCAL(order, gauss)                                                                        \
  FOR_MARKED(nlo_index, nhi_index)                                                                      \
  curvature(order, gauss, kappa, MKBOX(k), phi, MKBOX(phi), px, py, pz, gm, MKBOX(g), i, i, j, j, k, k, *dx, \
            *dy, *dz);                                                                                  \
  END_LIST

void oracle_lsm3dcomputemeancurvatureorder2local_(CURV_ARGS, const real *dx, const real *dy, const real *dz, XYZ,
                                                  NB, const nb_t *mark_fb) {
  /* lsm_curvature3d_local.f.in:15-145 */
  CURV_LOCAL(2, 0)
}
void oracle_lsm3dcomputegaussiancurvatureorder2local_(CURV_ARGS, const real *dx, const real *dy, const real *dz,
                                                      XYZ, NB, const nb_t *mark_fb) {
  /* lsm_curvature3d_local.f.in:151-280 */
  CURV_LOCAL(2, 1)
}
void oracle_lsm3dcomputemeancurvatureorder4local_(CURV_ARGS, const real *dx, const real *dy, const real *dz, XYZ,
                                                  NB, const nb_t *mark_fb) {
  /* lsm_curvature3d_local.f.in:285-426.  Its loop body divides the
   * Gaussian numerator by |grad phi|^4 (:403-417): it computes the
   * fourth-order Gaussian curvature, as the reference does. */
  CURV_LOCAL(4, 1)
}
void oracle_lsm3dcomputegaussiancurvatureorder4local_(CURV_ARGS, const real *dx, const real *dy, const real *dz,
                                                      XYZ, NB, const nb_t *mark_fb) {
  /* lsm_curvature3d_local.f.in:432-573 */
  CURV_LOCAL(4, 1)
}

void oracle_lsm3dgradientmagnitudelocal_(const real *gx, const real *gy, const real *gz, real *mag, BOXARGS(g),
                                         XYZ, NB, const nb_t *mark_fb) {
  /* lsm_spatial_derivatives3d_local.f.in:1186-1262: like the box routine
   * but with LSMLIB_ZERO_TOL in place of its 1e-8 */
  box_t bg = MKBOX(g);
  FOR_MARKED(nlo_index, nhi_index)
  real x2 = A(gx, bg, i, j, k) * A(gx, bg, i, j, k);
  real y2 = A(gy, bg, i, j, k) * A(gy, bg, i, j, k);
  real z2 = A(gz, bg, i, j, k) * A(gz, bg, i, j, k);
  if (x2 < ZERO_TOL) x2 = (real)0.0;
  if (y2 < ZERO_TOL) y2 = (real)0.0;
  if (z2 < ZERO_TOL) z2 = (real)0.0;
  real t = (real)sqrt(x2 + y2 + z2);
  A(mag, bg, i, j, k) = (t < ZERO_TOL) ? (real)0.0 : t;
  END_LIST
}

void oracle_lsm3daddadvectiontermtolserhslocal_(real *rhs, BOXARGS(r), const real *px, const real *py,
                                                const real *pz, BOXARGS(g), const real *vx, const real *vy,
                                                const real *vz, BOXARGS(vel), XYZ, NB, const nb_t *mark_fb) {
  /* lsm_level_set_evolution3d_local.f.in:60-146 */
  FOR_MARKED(nlo_index, nhi_index)
  oracle_lsm3daddadvectiontermtolserhs_(rhs, FWD(r), px, py, pz, FWD(g), vx, vy, vz, FWD(vel), CELL_BOX);
  END_LIST
}

void oracle_lsm3daddnormalveltermtolserhslocal_(real *rhs, BOXARGS(r), const real *xp, const real *yp,
                                                const real *zp, BOXARGS(gp), const real *xm, const real *ym,
                                                const real *zm, BOXARGS(gm), const real *vn, BOXARGS(vel), XYZ,
                                                NB, const nb_t *mark_fb) {
  /* lsm_level_set_evolution3d_local.f.in:152-276 */
  FOR_MARKED(nlo_index, nhi_index)
  oracle_lsm3daddnormalveltermtolserhs_(rhs, FWD(r), xp, yp, zp, FWD(gp), xm, ym, zm, FWD(gm), vn, FWD(vel),
                                        CELL_BOX);
  END_LIST
}

void oracle_lsm3daddconstprecomputedcurvtermtolserhslocal_(real *rhs, BOXARGS(r), const real *kappa,
                                                           BOXARGS(kap), const real *gmag, BOXARGS(gm),
                                                           const real *b, XYZ, NB, const nb_t *mark_fb) {
  /* lsm_level_set_evolution3d_local.f.in:530-606 */
  box_t br = MKBOX(r), bk = MKBOX(kap), bg = MKBOX(gm);
  FOR_MARKED(nlo_index, nhi_index)
  A(rhs, br, i, j, k) = A(rhs, br, i, j, k) + *b * A(kappa, bk, i, j, k) * A(gmag, bg, i, j, k);
  END_LIST
}

void oracle_lsm3daddexternalandnormalveltermtolserhslocal_(real *rhs, BOXARGS(r), const real *xp,
                                                           const real *yp, const real *zp, BOXARGS(gp),
                                                           const real *xm, const real *ym, const real *zm,
                                                           BOXARGS(gm), const real *vn, const real *vx,
                                                           const real *vy, const real *vz, BOXARGS(vel), XYZ,
                                                           NB, const nb_t *mark_fb) {
  /* lsm_level_set_evolution3d_local.f.in:612-800 */
  box_t br = MKBOX(r), bp = MKBOX(gp), bm = MKBOX(gm), bv = MKBOX(vel);
  FOR_MARKED(nlo_index, nhi_index)
  real vnc = A(vn, bv, i, j, k), vxc = A(vx, bv, i, j, k);
  real vyc = A(vy, bv, i, j, k), vzc = A(vz, bv, i, j, k);
  real px = ext_normal_select(vxc, vnc, A(xp, bp, i, j, k), A(xm, bm, i, j, k));
  real py = ext_normal_select(vyc, vnc, A(yp, bp, i, j, k), A(ym, bm, i, j, k));
  real pz = ext_normal_select(vzc, vnc, A(zp, bp, i, j, k), A(zm, bm, i, j, k));
  real n2 = px * px + py * py + pz * pz;
  if (fabs(vnc) >= ZERO_TOL && n2 >= ZERO_TOL)
    A(rhs, br, i, j, k) = A(rhs, br, i, j, k) - vxc * px - vyc * py - vzc * pz - vnc * (real)sqrt(n2);
  else
    A(rhs, br, i, j, k) = A(rhs, br, i, j, k) - vxc * px - vyc * py - vzc * pz;
  END_LIST
}

void oracle_lsm3dcomputeorthogonalizationeqnrhslocal_(real *rhs, BOXARGS(r), const real *xp, const real *yp,
                                                      const real *zp, BOXARGS(gp), const real *xm,
                                                      const real *ym, const real *zm, BOXARGS(gm),
                                                      const real *psi, BOXARGS(psi), const real *sxp,
                                                      const real *syp, const real *szp, BOXARGS(sp),
                                                      const real *sxm, const real *sym, const real *szm,
                                                      BOXARGS(sm), const real *dx, const real *dy, const real *dz,
                                                      XYZ, NB, const nb_t *mark_fb) {
  /* lsm_reinitialization3d_local.f.in:326-470 */
  FOR_MARKED(nlo_index, nhi_index)
  oracle_lsm3dcomputeorthogonalizationeqnrhs_(rhs, FWD(r), xp, yp, zp, FWD(gp), xm, ym, zm, FWD(gm), psi,
                                              FWD(psi), sxp, syp, szp, FWD(sp), sxm, sym, szm, FWD(sm), CELL_BOX,
                                              dx, dy, dz);
  END_LIST
}

void oracle_lsm3dcomputestableadvectiondtlocal_(real *dt, const real *vx, const real *vy, const real *vz,
                                                BOXARGS(vel), const real *dx, const real *dy, const real *dz,
                                                const real *cfl, XYZ, NB, const nb_t *mark_fb) {
  /* lsm_utilities3d_local.f.in:194-285 */
  box_t bv = MKBOX(vel);
  real mx = (real)-1.0;
  real ix_ = 1.0 / *dx, iy_ = 1.0 / *dy, iz_ = 1.0 / *dz;
  FOR_MARKED(nlo_index, nhi_index)
  real u = adv_u_over_dx(A(vx, bv, i, j, k), A(vy, bv, i, j, k), A(vz, bv, i, j, k), ix_, iy_, iz_);
  if (u > mx) mx = u;
  END_LIST
  *dt = *cfl / (mx + TINY);
}

static void normal_dt_local(real *dt, const real *vn, box_t bv, const real *cv, box_t bc, int cv_sgn,
                            const real *xp, const real *yp, const real *zp, box_t bp, const real *xm,
                            const real *ym, const real *zm, box_t bm, const real *dx, const real *dy,
                            const real *dz, const real *cfl, const int *ix, const int *iy, const int *iz,
                            const int *nlo_index, const int *nhi_index, const nb_t *nb, box_t bn, nb_t mark_fb) {
  /* lsm_utilities3d_local.f.in:290-426 and :432-622: the H/dX of the box
   * routine, taken only where sqrt(...) > LSMLIB_ZERO_TOL */
  real mdx = (real)fmax2(fmax2(*dx, *dy), *dz);
  real max_dx_sq = mdx * mdx, mx = (real)-1.0;
  real ix_ = 1.0 / *dx, iy_ = 1.0 / *dy, iz_ = 1.0 / *dz;
  FOR_LIST(nlo_index, nhi_index)
  if (A(nb, bn, i, j, k) <= mark_fb) {
    if (cv) {
      real c = A(cv, bc, i, j, k);
      if (!((cv_sgn > 0) ? (c > 0.0) : (c < 0.0))) continue;
    }
    real px = (real)fmax2(fabs(A(xp, bp, i, j, k)), fabs(A(xm, bm, i, j, k)));
    real py = (real)fmax2(fabs(A(yp, bp, i, j, k)), fabs(A(ym, bm, i, j, k)));
    real pz = (real)fmax2(fabs(A(zp, bp, i, j, k)), fabs(A(zm, bm, i, j, k)));
    real ng = (real)sqrt(px * px + py * py + pz * pz + max_dx_sq);
    if (ng > ZERO_TOL) {
      real h = (real)fabs(A(vn, bv, i, j, k)) / ng * (px * ix_ + py * iy_ + pz * iz_);
      if (h > mx) mx = h;
    }
  }
  END_LIST
  *dt = *cfl / (mx + TINY);
}

void oracle_lsm3dcomputestablenormalveldtlocal_(real *dt, const real *vn, BOXARGS(vel), const real *xp,
                                                const real *yp, const real *zp, BOXARGS(gp), const real *xm,
                                                const real *ym, const real *zm, BOXARGS(gm), const real *dx,
                                                const real *dy, const real *dz, const real *cfl, XYZ, NB,
                                                const nb_t *mark_fb) {
  normal_dt_local(dt, vn, MKBOX(vel), NULL, MKBOX(vel), 0, xp, yp, zp, MKBOX(gp), xm, ym, zm, MKBOX(gm), dx, dy,
                  dz, cfl, ix, iy, iz, nlo_index, nhi_index, nb, MKBOX(nb), *mark_fb);
}

void oracle_lsm3dcomputestablenormalveldtcontrolvolumelocal_(
    real *dt, const real *vn, BOXARGS(vel), const real *xp, const real *yp, const real *zp, BOXARGS(gp),
    const real *xm, const real *ym, const real *zm, BOXARGS(gm), const real *cv, BOXARGS(cv), const int *cv_sgn,
    const real *dx, const real *dy, const real *dz, const real *cfl, XYZ, NB, const nb_t *mark_fb) {
  normal_dt_local(dt, vn, MKBOX(vel), cv, MKBOX(cv), *cv_sgn, xp, yp, zp, MKBOX(gp), xm, ym, zm, MKBOX(gm), dx,
                  dy, dz, cfl, ix, iy, iz, nlo_index, nhi_index, nb, MKBOX(nb), *mark_fb);
}

/* ---------------------------------------------------------------------- */
/* Boundary conditions  lsm_boundary_conditions3d.f                         */
/* bdry_location_idx: 0 x_lo, 1 x_hi, 2 y_lo, 3 y_hi, 4 z_lo, 5 z_hi (:13-20) */
/* Each face loops over the full ghost box in the two other directions.     */
/* ---------------------------------------------------------------------- */
/* mode 0 linear (:46-170), 1 signed linear (:193-330), 2 copy (:352-463) */
static void extrapolate(real *phi, box_t g, const int fb[6], int loc, int mode) {
  if (loc < 0 || loc > 5) return;
  int d = loc / 2, hi = loc % 2;
  int a = DI(d), b = DJ(d), c = DK(d);
  int gl[3] = {g.ilo, g.jlo, g.klo}, gh[3] = {g.ihi, g.jhi, g.khi};
  int bnd = hi ? fb[2 * d + 1] : fb[2 * d];   /* boundary fill-box index */
  int in = hi ? -1 : 1;                        /* step into the interior */
  int u = (d + 1) % 3, v = (d + 2) % 3;
  for (int q = gl[v]; q <= gh[v]; q++)
    for (int p = gl[u]; p <= gh[u]; p++) {
      int idx[3];
      idx[u] = p; idx[v] = q; idx[d] = bnd;
      real pb = A(phi, g, idx[0], idx[1], idx[2]);
      real pi1 = A(phi, g, idx[0] + in * a, idx[1] + in * b, idx[2] + in * c);
      real slope = (real)0.0;
      if (mode == 0) {
        slope = pb - pi1;
      } else if (mode == 1) {
        real s = (real)copysign(1.0, pb);   /* SIGN(one, phi) */
        real ad = (real)fabs(pb - pi1);
        slope = s * ad;
      }
      int lo = hi ? bnd + 1 : gl[d], up = hi ? gh[d] : bnd - 1;
      for (int t = lo; t <= up; t++) {
        idx[d] = t;
        if (mode == 2) {
          A(phi, g, idx[0], idx[1], idx[2]) = pb;
        } else {
          real dist = (real)(hi ? (t - bnd) : (bnd - t));
          A(phi, g, idx[0], idx[1], idx[2]) = pb + slope * dist;
        }
      }
    }
}

#define BC_ENTRY(name, mode)                                                    \
  void oracle_##name(real *phi, BOXARGS(g), BOXARGS(fb), const int *loc) {      \
    int fb[6] = {*ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb};         \
    extrapolate(phi, MKBOX(g), fb, *loc, mode);                                  \
  }
BC_ENTRY(lsm3dlinearextrapolation_, 0)
BC_ENTRY(lsm3dsignedlinearextrapolation_, 1)
BC_ENTRY(lsm3dcopyextrapolation_, 2)
/* Homogeneous Neumann for ENO1/2/3 and WENO5 reduces to copy extrapolation
 * (lsm_boundary_conditions3d.f:486-640). */
BC_ENTRY(lsm3dhomogeneousneumanneno1_, 2)
BC_ENTRY(lsm3dhomogeneousneumanneno2_, 2)
BC_ENTRY(lsm3dhomogeneousneumanneno3_, 2)
BC_ENTRY(lsm3dhomogeneousneumannweno5_, 2)
