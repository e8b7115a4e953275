/*
 * CPU oracle for the CRUNA time-stepping hot path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference Fortran, used only as the checker
 * for the HIP library (tests/, __graft_entry__.smoke(), bench.py's
 * cpu_baseline leg).  Expression order follows the Fortran array
 * expressions term by term so that FP64 results agree to rounding.
 * Each function cites the reference file:line it restates (paths are
 * relative to the reference root).  Every function is pinned by a test in
 * tests/test_oracle*.py: against the reference's own compiled leaf modules
 * where they build here (cores, closures, characteristic and sponge rows
 * including sponge_rhss_correct), otherwise by closed forms, invariants and
 * convergence to an independent discretisation.
 */
#include "cruna_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NV 5

static size_t npts(const orc_cfg *c) { return (size_t)c->n[0] * c->n[1] * c->n[2]; }

static double *xalloc(size_t n) {
  double *p = (double *)malloc(n * sizeof(double));
  if (!p) {
    fprintf(stderr, "oracle: out of memory (%zu doubles)\n", n);
    abort();
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* LAPACK dgttrf / dgttrs (dgtts2, itrans = 0).  Third-party routines called by
 * the reference at src/diff/diff_pen_core_o6_i_c_p.f90:213,219,135 and
 * src/diff/diff_pen_core_o6_i_c_o.f90:194,129 (and the filter cores).  This
 * restates LAPACK 3.x's published algorithm: tridiagonal LU with partial
 * pivoting, second superdiagonal du2, row interchanges in ipiv (1-based). */
int orc_dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv) {
  int info = 0;
  for (int i = 0; i < n; ++i) ipiv[i] = i + 1;
  for (int i = 0; i < n - 2; ++i) du2[i] = 0.0;
  for (int i = 0; i < n - 2; ++i) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] != 0.0) {
        double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      du2[i] = du[i + 1];
      du[i + 1] = -fact * du[i + 1];
      ipiv[i] = i + 2;
    }
  }
  if (n > 1) {
    int i = n - 2;
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] != 0.0) {
        double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      ipiv[i] = i + 2;
    }
  }
  for (int i = 0; i < n; ++i)
    if (d[i] == 0.0) {
      info = i + 1;
      break;
    }
  return info;
}

void orc_dgttrs(int n, int nrhs, const double *dl, const double *d, const double *du,
                const double *du2, const int *ipiv, double *b, int ldb) {
  for (int j = 0; j < nrhs; ++j) {
    double *x = b + (size_t)j * ldb;
    /* solve L x = b */
    for (int i = 0; i < n - 1; ++i) {
      if (ipiv[i] == i + 1) {
        x[i + 1] = x[i + 1] - dl[i] * x[i];
      } else {
        double temp = x[i];
        x[i] = x[i + 1];
        x[i + 1] = temp - dl[i] * x[i];
      }
    }
    /* solve U x = b */
    x[n - 1] = x[n - 1] / d[n - 1];
    if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; --i)
      x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Factorisations of the left matrices B.                                   */
typedef struct {
  int n;
  double *dl, *dc, *du, *du2, *fbnorm;
  int *ipiv;
  double fbrat;
  int cyclic;
} orc_lu;

static void lu_free(orc_lu *f) {
  free(f->dl);
  free(f->dc);
  free(f->du);
  free(f->du2);
  free(f->fbnorm);
  free(f->ipiv);
}

static void lu_alloc(orc_lu *f, int n) {
  f->n = n;
  f->dl = xalloc(n);
  f->dc = xalloc(n);
  f->du = xalloc(n);
  f->du2 = xalloc(n);
  f->fbnorm = xalloc(n);
  f->ipiv = (int *)malloc(n * sizeof(int));
  f->fbrat = 0.0;
  f->cyclic = 0;
}

/* Cyclic tridiag(alpha,diag,alpha) with the Sherman-Morrison split used by
 * src/diff/diff_pen_core_o6_i_c_p.f90:182-228 (diag 1),
 * src/filter/filter_pen_core_o6_i_c_p.f90:217-246 (diag 1) and
 * src/diff/diff_pen_core_o4_i_c_p_pade.f90:190-230 (diag 2/3). */
static void lu_cyclic_d(orc_lu *f, int n, double alpha, double diag) {
  lu_alloc(f, n);
  f->cyclic = 1;
  double *fa = xalloc(n), *fb = xalloc(n);
  for (int i = 0; i < n - 1; ++i) {
    f->dl[i] = alpha;
    f->du[i] = alpha;
  }
  for (int i = 0; i < n; ++i) {
    f->dc[i] = diag;
    fa[i] = 0.0;
    fb[i] = 0.0;
  }
  fa[0] = -f->dc[0];
  fa[n - 1] = +alpha;
  fb[0] = +1.0;
  fb[n - 1] = -alpha / f->dc[0];
  f->dc[n - 1] = f->dc[n - 1] + (alpha * alpha / f->dc[0]);
  f->dc[0] = f->dc[0] + f->dc[0];
  int info = orc_dgttrf(n, f->dl, f->dc, f->du, f->du2, f->ipiv);
  if (info) fprintf(stderr, "oracle: dgttrf info %d\n", info);
  orc_dgttrs(n, 1, f->dl, f->dc, f->du, f->du2, f->ipiv, fa, n);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += fb[i] * fa[i];
  for (int i = 0; i < n; ++i) f->fbnorm[i] = fa[i] * fb[0] / (1.0 + s);
  f->fbrat = fb[n - 1] / fb[0];
  free(fa);
  free(fb);
}

static void lu_cyclic(orc_lu *f, int n, double alpha) { lu_cyclic_d(f, n, alpha, 1.0); }

/* Open-boundary B of the 4th-order Pade derivative,
 * src/diff/diff_pen_core_o4_i_c_o_pade.f90:150-190: tridiag(1/6, 2/3, 1/6)
 * with dc(1) = dc(n) = 1/3. */
static void lu_pade_open(orc_lu *f, int n) {
  lu_alloc(f, n);
  for (int i = 0; i < n - 1; ++i) {
    f->dl[i] = 1.0 / 6.0;
    f->du[i] = 1.0 / 6.0;
  }
  for (int i = 0; i < n; ++i) f->dc[i] = 2.0 / 3.0;
  f->dc[0] = 1.0 / 3.00;
  f->dc[n - 1] = 1.0 / 3.00;
  int info = orc_dgttrf(n, f->dl, f->dc, f->du, f->du2, f->ipiv);
  if (info) fprintf(stderr, "oracle: dgttrf info %d\n", info);
}

/* Open-boundary B of the Lele o6 derivative, src/diff/diff_pen_core_o6_i_c_o.f90:159-199. */
static void lu_diff_open(orc_lu *f, int n) {
  lu_alloc(f, n);
  const double alpha = 1.0 / 3.0;
  for (int i = 0; i < n - 1; ++i) {
    f->dl[i] = alpha;
    f->du[i] = alpha;
  }
  for (int i = 0; i < n; ++i) f->dc[i] = 1.0;
  f->du[0] = 3.00;
  f->du[1] = 0.25;
  f->du[n - 2] = 0.25;
  f->dl[0] = 0.25;
  f->dl[n - 2] = 3.00;
  f->dl[n - 3] = 0.25;
  int info = orc_dgttrf(n, f->dl, f->dc, f->du, f->du2, f->ipiv);
  if (info) fprintf(stderr, "oracle: dgttrf info %d\n", info);
}

/* Open-boundary B of the GV2000 o6 filter, src/filter/filter_pen_core_o6_i_c_o.f90:203-240. */
static void lu_filter_open(orc_lu *f, int n, double alpha) {
  lu_alloc(f, n);
  for (int i = 0; i < n - 1; ++i) {
    f->dl[i] = alpha;
    f->du[i] = alpha;
  }
  for (int i = 0; i < n; ++i) f->dc[i] = 1.0;
  f->du[0] = 0.0;
  f->dl[n - 2] = 0.0;
  int info = orc_dgttrf(n, f->dl, f->dc, f->du, f->du2, f->ipiv);
  if (info) fprintf(stderr, "oracle: dgttrf info %d\n", info);
}

static void lu_solve(const orc_lu *f, double *q, int nlines) {
  const int n = f->n;
  orc_dgttrs(n, nlines, f->dl, f->dc, f->du, f->du2, f->ipiv, q, n);
  if (f->cyclic) {
    for (int l = 0; l < nlines; ++l) {
      double *x = q + (size_t)l * n;
      const double s = x[0] + f->fbrat * x[n - 1];
      for (int i = 0; i < n; ++i) x[i] = x[i] - f->fbnorm[i] * s;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Derivative cores.                                                        */

/* src/diff/diff_pen_core_o6_i_c_p.f90:87-145 (stencil 157-165: a = 14/9/2, b = 1/9/4). */
static void diff_periodic(int n, int nlines, double *q, double dx) {
  const double a = 14.00 / 9.0 / 2.0;
  const double b = 1.00 / 9.0 / 4.0;
  orc_lu f;
  lu_cyclic(&f, n, 1.0 / 3.0);
  double *t = xalloc(n);
  for (int l = 0; l < nlines; ++l) {
    double *x = q + (size_t)l * n;
    for (int i = 0; i < n; ++i) {
      const int ip1 = (i + 1) % n, ip2 = (i + 2) % n;
      const int im1 = (i - 1 + n) % n, im2 = (i - 2 + 2 * n) % n;
      t[i] = a * (x[ip1] - x[im1]) + b * (x[ip2] - x[im2]);
    }
    for (int i = 0; i < n; ++i) x[i] = t[i] / dx;
  }
  lu_solve(&f, q, nlines);
  free(t);
  lu_free(&f);
}

/* src/diff/diff_pen_core_o6_i_c_o.f90:85-134 (stencil 140-155). */
static void diff_open(int n, int nlines, double *q, double dx) {
  const double a = 14.00 / 9.0 / 2.0, b = 1.00 / 9.0 / 4.0;
  const double c = 17.00 / 6.0, d = 1.50, e = 1.00 / 6.0, f = 0.75;
  orc_lu lu;
  lu_diff_open(&lu, n);
  double *t = xalloc(n);
  for (int l = 0; l < nlines; ++l) {
    double *x = q + (size_t)l * n;
    t[0] = -c * x[0] + d * (x[1] + x[2]) - e * x[3];
    t[1] = -f * x[0] + f * x[2];
    for (int i = 2; i <= n - 3; ++i) t[i] = a * (x[i + 1] - x[i - 1]) + b * (x[i + 2] - x[i - 2]);
    t[n - 2] = f * x[n - 1] - f * x[n - 3];
    t[n - 1] = c * x[n - 1] - d * (x[n - 2] + x[n - 3]) + e * x[n - 4];
    for (int i = 0; i < n; ++i) x[i] = t[i] / dx;
  }
  lu_solve(&lu, q, nlines);
  free(t);
  lu_free(&lu);
}

void orc_diff_core(int scheme, int n, int nlines, double *q, double dx) {
  if (scheme == ORC_PERIODIC)
    diff_periodic(n, nlines, q, dx);
  else
    diff_open(n, nlines, q, dx);
}

/* 4th-order explicit cores: src/diff/diff_pen_core_o4_e_c_p.f90:53-82 and
 * diff_pen_core_o4_e_c_o.f90:53-82 (8 (f_{i+1} - f_{i-1}) - (f_{i+2} - f_{i-2}),
 * one-sided 5-point rows 1, 2, n-1, n on open lines; then / (12 dx)). */
void orc_diff_core_o4(int scheme, int n, int nlines, double *q, double dx) {
  double *t = xalloc(n);
  for (int l = 0; l < nlines; ++l) {
    double *x = q + (size_t)l * n;
    for (int i = 0; i < n; ++i) {
      if (scheme == ORC_PERIODIC || (i >= 2 && i <= n - 3)) {
        const int ip1 = (i + 1) % n, ip2 = (i + 2) % n;
        const int im1 = (i - 1 + n) % n, im2 = (i - 2 + 2 * n) % n;
        t[i] = 8.0 * (x[ip1] - x[im1]) - 1.0 * (x[ip2] - x[im2]);
      }
    }
    if (scheme != ORC_PERIODIC) {
      t[0] = -25.0 * x[0] + 48.0 * x[1] - 36.0 * x[2] + 16.0 * x[3] - 3.0 * x[4];
      t[1] = -3.0 * x[0] - 10.0 * x[1] + 18.0 * x[2] - 6.0 * x[3] + 1.0 * x[4];
      t[n - 2] = -1.0 * x[n - 5] + 6.0 * x[n - 4] - 18.0 * x[n - 3] + 10.0 * x[n - 2] + 3.0 * x[n - 1];
      t[n - 1] = 3.0 * x[n - 5] - 16.0 * x[n - 4] + 36.0 * x[n - 3] - 48.0 * x[n - 2] + 25.0 * x[n - 1];
    }
    for (int i = 0; i < n; ++i) x[i] = t[i] / (dx * 12.0);
  }
  free(t);
}

/* 2nd-order explicit open core, src/diff/diff_pen_core_o2_e_c_o.f90 diff_main:
 * (f+1 - f-1)/2 inside, one-sided rows 1, 2, n-1, n; then / dx. */
void orc_diff_core_o2(int n, int nlines, double *q, double dx) {
  double *t = xalloc(n);
  for (int l = 0; l < nlines; ++l) {
    double *x = q + (size_t)l * n;
    t[0] = -1.5 * x[0] + 2.0 * x[1] - 0.5 * x[2];
    t[1] = -0.5 * x[0] + 0.5 * x[2];
    for (int i = 2; i < n - 2; ++i) t[i] = 0.5 * (x[i + 1] - x[i - 1]);
    t[n - 2] = -0.5 * x[n - 3] + 0.5 * x[n - 1];
    t[n - 1] = +0.5 * x[n - 3] - 2.0 * x[n - 2] + 1.5 * x[n - 1];
    for (int i = 0; i < n; ++i) x[i] = t[i] / dx;
  }
  free(t);
}

/* 4th-order Pade cores: src/diff/diff_pen_core_o4_i_c_p_pade.f90:96-140 and
 * diff_pen_core_o4_i_c_o_pade.f90:81-124 (a (f_{i+1} - f_{i-1}), a = 1/2, with
 * the one-sided rows -a f1 + a f2 and a fn - a f(n-1) on open lines; / dx;
 * then the B solve). */
void orc_diff_core_pade(int scheme, int n, int nlines, double *q, double dx) {
  const double a = 1.0 / 2.0;
  orc_lu f;
  if (scheme == ORC_PERIODIC)
    lu_cyclic_d(&f, n, 1.0 / 6.0, 2.0 / 3.0);
  else
    lu_pade_open(&f, n);
  double *t = xalloc(n);
  for (int l = 0; l < nlines; ++l) {
    double *x = q + (size_t)l * n;
    for (int i = 0; i < n; ++i) {
      const int ip1 = (i + 1) % n, im1 = (i - 1 + n) % n;
      t[i] = a * (x[ip1] - x[im1]);
    }
    if (scheme != ORC_PERIODIC) {
      t[0] = -a * x[0] + a * (x[1]);
      t[n - 1] = a * x[n - 1] - a * (x[n - 2]);
    }
    for (int i = 0; i < n; ++i) x[i] = t[i] / dx;
  }
  lu_solve(&f, q, nlines);
  free(t);
  lu_free(&f);
}

/* ------------------------------------------------------------------------ */
/* Filter cores.                                                            */

/* Stencil a0..a3, src/filter/filter_pen_core_o6_i_c_p.f90:151-182. */
static void filter_stencil(double alpha, double s[4]) {
  double a0 = 11.0 / 16.0 + 5.0 * alpha / 8.0;
  double a1 = 15.0 / 32.0 + 17.0 * alpha / 16.0;
  a1 = a1 / 2.0;
  double a2 = -3.0 / 16.0 + 3.0 * alpha / 8.0;
  a2 = a2 / 2.0;
  double a3 = 1.0 / 32.0 - 1.0 * alpha / 16.0;
  a3 = a3 / 2.0;
  s[0] = a0;
  s[1] = a1;
  s[2] = a2;
  s[3] = a3;
}

/* src/filter/filter_pen_core_o6_i_c_p.f90:84-145. */
static void filter_periodic(int n, int nlines, double *q, double alpha) {
  double s[4];
  filter_stencil(alpha, s);
  orc_lu f;
  lu_cyclic(&f, n, alpha);
  double *t = xalloc(n);
  for (int l = 0; l < nlines; ++l) {
    double *x = q + (size_t)l * n;
    for (int i = 0; i < n; ++i) {
      const int p1 = (i + 1) % n, p2 = (i + 2) % n, p3 = (i + 3) % n;
      const int m1 = (i - 1 + n) % n, m2 = (i - 2 + 2 * n) % n, m3 = (i - 3 + 3 * n) % n;
      t[i] = s[0] * x[i] + s[1] * (x[p1] + x[m1]) + s[2] * (x[p2] + x[m2]) + s[3] * (x[p3] + x[m3]);
    }
    memcpy(x, t, n * sizeof(double));
  }
  lu_solve(&f, q, nlines);
  free(t);
  lu_free(&f);
}

/* Boundary rows of the open filter, src/filter/filter_pen_core_o6_i_c_o.f90:179-195. */
static void filter_open_boundary(double alpha, double cb[2][7]) {
  cb[0][0] = 1.0 / 64.0 + 31.0 * alpha / 32.0;
  cb[0][1] = 29.0 / 32.0 + 3.0 * alpha / 16.0;
  cb[0][2] = 15.0 / 64.0 + 17.0 * alpha / 32.0;
  cb[0][3] = -5.0 / 16.0 + 5.0 * alpha / 8.0;
  cb[0][4] = 15.0 / 64.0 - 15.0 * alpha / 32.0;
  cb[0][5] = -3.0 / 32.0 + 3.0 * alpha / 16.0;
  cb[0][6] = 1.0 / 64.0 - 1.0 * alpha / 32.0;
  cb[1][0] = -1.0 / 64.0 + 1.0 * alpha / 32.0;
  cb[1][1] = 3.0 / 32.0 + 13.0 * alpha / 16.0;
  cb[1][2] = 49.0 / 64.0 + 15.0 * alpha / 32.0;
  cb[1][3] = 5.0 / 16.0 + 3.0 * alpha / 8.0;
  cb[1][4] = -15.0 / 64.0 + 15.0 * alpha / 32.0;
  cb[1][5] = 3.0 / 32.0 - 3.0 * alpha / 16.0;
  cb[1][6] = -1.0 / 64.0 + 1.0 * alpha / 32.0;
}

/* src/filter/filter_pen_core_o6_i_c_o.f90:79-143. */
static void filter_open(int n, int nlines, double *q, double alpha) {
  double s[4], cb[2][7];
  filter_stencil(alpha, s);
  filter_open_boundary(alpha, cb);
  orc_lu f;
  lu_filter_open(&f, n, alpha);
  double *t = xalloc(n);
  for (int l = 0; l < nlines; ++l) {
    double *x = q + (size_t)l * n;
    for (int i = 3; i <= n - 4; ++i)
      t[i] = s[0] * x[i] + s[1] * (x[i + 1] + x[i - 1]) + s[2] * (x[i + 2] + x[i - 2]) +
             s[3] * (x[i + 3] + x[i - 3]);
    t[0] = x[0];
    t[n - 1] = x[n - 1];
    t[1] = 0.0;
    t[n - 2] = 0.0;
    for (int i = 0; i < 7; ++i) {
      t[1] = t[1] + cb[0][i] * x[i];
      t[n - 2] = t[n - 2] + cb[0][i] * x[n - 1 - i];
    }
    t[2] = 0.0;
    t[n - 3] = 0.0;
    for (int i = 0; i < 7; ++i) {
      t[2] = t[2] + cb[1][i] * x[i];
      t[n - 3] = t[n - 3] + cb[1][i] * x[n - 1 - i];
    }
    memcpy(x, t, n * sizeof(double));
  }
  lu_solve(&f, q, nlines);
  free(t);
  lu_free(&f);
}

void orc_filter_core(int scheme, int n, int nlines, double *q, double alpha) {
  if (scheme == ORC_PERIODIC)
    filter_periodic(n, nlines, q, alpha);
  else
    filter_open(n, nlines, q, alpha);
}

/* ------------------------------------------------------------------------ */
/* 3-D wrappers: gather each line along dir into a pencil (the single-rank
 * ftranspose/btranspose of src/parallelism/transpose_mpi_cartesian.f90:110-127,
 * 326-350, 566-594 only reorder lines, which the per-line algorithm ignores). */

static void gather_pencils(const orc_cfg *c, int dir, const double *in, double *pen) {
  const int n1 = c->n[0], n2 = c->n[1], n3 = c->n[2];
  const int nd = c->n[dir];
  size_t line = 0;
  if (dir == 0) {
    memcpy(pen, in, (size_t)n1 * n2 * n3 * sizeof(double));
    return;
  }
  (void)line;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < (dir == 2 ? n2 : n3); ++k)
    for (int i = 0; i < n1; ++i) {
      const size_t ln = (size_t)k * n1 + i;
      for (int p = 0; p < nd; ++p) {
        size_t idx = (dir == 1) ? (size_t)i + (size_t)n1 * (p + (size_t)n2 * k)
                                : (size_t)i + (size_t)n1 * (k + (size_t)n2 * p);
        pen[ln * nd + p] = in[idx];
      }
    }
}

static void scatter_pencils(const orc_cfg *c, int dir, const double *pen, double *out) {
  const int n1 = c->n[0], n2 = c->n[1], n3 = c->n[2];
  const int nd = c->n[dir];
  size_t line = 0;
  if (dir == 0) {
    memcpy(out, pen, (size_t)n1 * n2 * n3 * sizeof(double));
    return;
  }
  (void)line;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < (dir == 2 ? n2 : n3); ++k)
    for (int i = 0; i < n1; ++i) {
      const size_t ln = (size_t)k * n1 + i;
      for (int p = 0; p < nd; ++p) {
        size_t idx = (dir == 1) ? (size_t)i + (size_t)n1 * (p + (size_t)n2 * k)
                                : (size_t)i + (size_t)n1 * (k + (size_t)n2 * p);
        out[idx] = pen[ln * nd + p];
      }
    }
}

/* Optional hooks that run the reference's own compiled cores (the o6
 * derivative core with the signature of ref_diff in oracle/ref_harness.f90,
 * the o6 filter core with that of ref_filter) in place of the restated ones:
 * the CPU baseline times the reference's cores this way.  They are called on
 * all lines of a pencil at once, from one thread (the reference's cores keep
 * module state).  NULL (the default): the restatement. */
typedef void (*orc_diff_hook)(int scheme, double *q, int n, int nlines, double dx);
typedef void (*orc_filter_hook)(int scheme, double *q, int n, int nlines);
static orc_diff_hook g_diff_hook = NULL;
static orc_filter_hook g_filter_hook = NULL;

void orc_set_core_hooks(void *diff, void *filter) {
  g_diff_hook = (orc_diff_hook)diff;
  g_filter_hook = (orc_filter_hook)filter;
}

/* threads of the oracle's own parallel loops (OpenMP) */
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

/* src/diff/diff_i1_mpi_cartesian.f90:83-103 (and i2/i3). */
void orc_diff3d(const orc_cfg *c, int dir, double *out, const double *in, double h) {
  const size_t N = npts(c);
  double *pen = xalloc(N);
  gather_pencils(c, dir, in, pen);
  const int nd = c->n[dir];
  const long nl = (long)(N / nd);
  if (g_diff_hook && c->diff_kind[dir] == ORC_DIFF_O6_COMPACT) {
    g_diff_hook(c->scheme[dir], pen, nd, (int)nl, h);
    scatter_pencils(c, dir, pen, out);
    free(pen);
    return;
  }
  /* the lines are independent: each thread runs the core on its own share */
#pragma omp parallel
  {
    long l0 = 0, l1 = nl;
#ifdef _OPENMP
    const long nt = omp_get_num_threads(), t = omp_get_thread_num();
    l0 = nl * t / nt;
    l1 = nl * (t + 1) / nt;
#endif
    double *p = pen + (size_t)l0 * nd;
    const int m = (int)(l1 - l0);
    if (m > 0) {
      if (c->diff_kind[dir] == ORC_DIFF_O4_EXPLICIT)
        orc_diff_core_o4(c->scheme[dir], nd, m, p, h);
      else if (c->diff_kind[dir] == ORC_DIFF_O4_PADE)
        orc_diff_core_pade(c->scheme[dir], nd, m, p, h);
      else if (c->diff_kind[dir] == ORC_DIFF_O2_EXPLICIT)
        orc_diff_core_o2(nd, m, p, h);
      else
        orc_diff_core(c->scheme[dir], nd, m, p, h);
    }
  }
  scatter_pencils(c, dir, pen, out);
  free(pen);
}

/* src/filter/filter_i1_mpi_cartesian.f90:83-97 (and i2/i3). */
void orc_filter4d(const orc_cfg *c, int dir, double *q, int nv) {
  const size_t N = npts(c);
  double *pen = xalloc(N);
  for (int v = 0; v < nv; ++v) {
    gather_pencils(c, dir, q + v * N, pen);
    const int nd = c->n[dir];
    const long nl = (long)(N / nd);
    if (g_filter_hook) {
      g_filter_hook(c->scheme[dir], pen, nd, (int)nl);
      scatter_pencils(c, dir, pen, q + v * N);
      continue;
    }
#pragma omp parallel
    {
      long l0 = 0, l1 = nl;
#ifdef _OPENMP
      const long nt = omp_get_num_threads(), t = omp_get_thread_num();
      l0 = nl * t / nt;
      l1 = nl * (t + 1) / nt;
#endif
      if (l1 > l0)
        orc_filter_core(c->scheme[dir], nd, (int)(l1 - l0), pen + (size_t)l0 * nd, c->filter_alpha);
    }
    scatter_pencils(c, dir, pen, q + v * N);
  }
  free(pen);
}

/* ------------------------------------------------------------------------ */
/* Thermodynamic closures, src/equations/ns_rho_rhou_p_trafos.f90.          */

/* give_T :36-60  (T = (W/R) * p / rho) */
double orc_T(const orc_cfg *c, double rho, double p) {
  const double w_inv_r = c->W / c->R;
  return w_inv_r * p / rho;
}

/* give_mu :119-142 (Sutherland, S1 = 110.4).  The power law
 * mu0*(T/T0)**exp is the case-level override of the same function. */
double orc_mu(const orc_cfg *c, double T) {
  if (c->visc_law == ORC_VISC_POWER) return c->mu0 * pow(T / c->T0, c->visc_exp);
  const double S1 = 110.4;
  return c->mu0 * pow(T / c->T0, 1.5) * ((c->T0 + S1) / (T + S1));
}

/* give_lambda :146-171 (lambda = mu*cp/Pr, cp = gamma*(R/W)/(gamma-1)) */
double orc_lambda(const orc_cfg *c, double mu) {
  const double cp = c->gamma * (c->R / c->W) / (c->gamma - 1.0);
  return mu * cp / c->Pr;
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions.                                                     */

static int bc_type(const orc_cfg *c, int r) { return (int)lround(c->bc[r][0]); }

/* Visit the points of the box of BC row r (1-based inclusive indices). */
/* Box bounds clamped to the grid as modify_boundary_conditions_array does on
 * one rank (src/boundaries/boundary_conditions.f90:586-596: max(1, start),
 * min(n, end); a box outside the grid becomes empty), 0-based. */
static int box_lo(const orc_cfg *c, int r, int col) {
  const long v = lround(c->bc[r][col]);
  return (int)(v < 1 ? 0 : v - 1);
}
static int box_hi(const orc_cfg *c, int r, int col, int d) {
  const long v = lround(c->bc[r][col]);
  return (int)(v > c->n[d] ? c->n[d] - 1 : v - 1);
}
#define FOR_BOX(c, r, idx)                                                                      \
  for (int kk = box_lo(c, r, 5); kk <= box_hi(c, r, 6, 2); ++kk)                                \
    for (int jj = box_lo(c, r, 3); jj <= box_hi(c, r, 4, 1); ++jj)                              \
      for (int ii = box_lo(c, r, 1); ii <= box_hi(c, r, 2, 0); ++ii)                            \
        for (size_t idx = (size_t)ii + (size_t)(c)->n[0] * (jj + (size_t)(c)->n[1] * kk), _o = 1; _o; _o = 0)

/* src/boundaries/nonreflecting.f90:46-137 on one point (stride_v = distance
 * between variables).  npts points starting at q, contiguous. */
/* sc scales the change toward aim before the decomposition: the shape of
 * characteristic_q_k4d (nonreflecting.f90:180-185, type 203), 1 otherwise. */
static void char_q(int np, int sv, double *q, const double *aim, const double k[3], double gamma,
                   double sc) {
  for (int m = 0; m < np; ++m) {
    double *Q = q + m;
    const double *A = aim + m;
    double r = Q[0];
    double u1 = Q[sv] / Q[0], u2 = Q[2 * sv] / Q[0], u3 = Q[3 * sv] / Q[0];
    double p = Q[4 * sv];
    double c = sqrt(gamma * Q[4 * sv] / Q[0]);
    double dq1 = A[0] - r;
    double dq2 = A[sv] / A[0] - u1;
    double dq3 = A[2 * sv] / A[0] - u2;
    double dq4 = A[3 * sv] / A[0] - u3;
    double dq5 = A[4 * sv] - p;
    dq1 *= sc;
    dq2 *= sc;
    dq3 *= sc;
    dq4 *= sc;
    dq5 *= sc;
    const double kx = -k[0], ky = -k[1], kz = -k[2];
    double R1 = kx * dq1 + kz * dq3 - ky * dq4 - kx * dq5 / (c * c);
    double R2 = ky * dq1 - kz * dq2 + kx * dq4 - ky * dq5 / (c * c);
    double R3 = kz * dq1 + ky * dq2 - kx * dq3 - kz * dq5 / (c * c);
    double R4 = kx * dq2 + ky * dq3 + kz * dq4 + dq5 / (r * c);
    double R5 = -kx * dq2 - ky * dq3 - kz * dq4 + dq5 / (r * c);
    double lamS = kx * u1 + ky * u2 + kz * u3;
    double lamP = lamS + c, lamM = lamS - c;
    if (lamS < 0.0) R1 = R2 = R3 = 0.0;
    if (lamP < 0.0) R4 = 0.0;
    if (lamM < 0.0) R5 = 0.0;
    dq1 = kx * R1 + ky * R2 + kz * R3 + r / (2.0 * c) * R4 + r / (2.0 * c) * R5;
    dq2 = -kz * R2 + ky * R3 + 0.5 * kx * R4 - 0.5 * kx * R5;
    dq3 = kz * R1 - kx * R3 + 0.5 * ky * R4 - 0.5 * ky * R5;
    dq4 = -ky * R1 + kx * R2 + 0.5 * kz * R4 - 0.5 * kz * R5;
    dq5 = +0.5 * r * c * R4 + 0.5 * r * c * R5;
    r = r + dq1;
    u1 = u1 + dq2;
    u2 = u2 + dq3;
    u3 = u3 + dq4;
    p = p + dq5;
    Q[0] = r;
    Q[sv] = r * u1;
    Q[2 * sv] = r * u2;
    Q[3 * sv] = r * u3;
    Q[4 * sv] = p;
  }
}

void orc_characteristic_q_k1d(int np, int sv, double *q, const double *aim, const double k[3],
                              double gamma) {
  char_q(np, sv, q, aim, k, gamma, 1.0);
}

/* src/boundaries/nonreflecting.f90:453-554 (adjoint, target zero); sc: the
 * shape of characteristic_qs_k4d (:596-601, type 203), 1 otherwise. */
static void char_qs(int np, int sv, double *qs, const double *q, const double k[3], double gamma,
                    double sc) {
  for (int m = 0; m < np; ++m) {
    double *S = qs + m;
    const double *Q = q + m;
    double ds1 = -S[0] * sc, ds2 = -S[sv] * sc, ds3 = -S[2 * sv] * sc, ds4 = -S[3 * sv] * sc,
           ds5 = -S[4 * sv] * sc;
    const double kx = -k[0], ky = -k[1], kz = -k[2];
    const double gm1 = gamma - 1.0;
    double c = sqrt(gamma * (Q[4 * sv] / Q[0]));
    double g_c = gm1 / c;
    double u = Q[sv] / Q[0], v = Q[2 * sv] / Q[0], w = Q[3 * sv] / Q[0];
    double ku = kx * u + ky * v + kz * w;
    double kus = kx * ds2 + ky * ds3 + kz * ds4;
    double ruus = (ds1 + u * ds2 + v * ds3 + w * ds4) / c;
    double rups = ruus + ds5 / g_c;
    int Sin = ku < 0.0, Pin = ku + c < 0.0, Min = ku - c < 0.0;
    double drs = 0, dus = 0, dvs = 0, dws = 0, dps = 0, tmp;
    if (Sin) {
      drs = ds1 + ku * kus;
      dus = ds2 - kx * kus;
      dvs = ds3 - ky * kus;
      dws = ds4 - kz * kus;
      dps = -g_c * ruus;
    }
    if (Pin) {
      tmp = 0.5 * (kus + rups);
      drs = drs - ku * tmp;
      dus = dus + kx * tmp;
      dvs = dvs + ky * tmp;
      dws = dws + kz * tmp;
      dps = dps + g_c * tmp;
    }
    if (Min) {
      tmp = 0.5 * (kus - rups);
      drs = drs - ku * tmp;
      dus = dus + kx * tmp;
      dvs = dvs + ky * tmp;
      dws = dws + kz * tmp;
      dps = dps - g_c * tmp;
    }
    S[0] = S[0] + drs;
    S[sv] = S[sv] + dus;
    S[2 * sv] = S[2 * sv] + dvs;
    S[3 * sv] = S[3 * sv] + dws;
    S[4 * sv] = S[4 * sv] + dps;
  }
}

void orc_characteristic_qs_k1d(int np, int sv, double *qs, const double *q, const double k[3],
                               double gamma) {
  char_qs(np, sv, qs, q, k, gamma, 1.0);
}

static void bc_unsupported(int type, const char *where) {
  fprintf(stderr, "oracle: boundary type %d not implemented in %s\n", type, where);
  abort();
}

/* set_boundary_condition_q, src/boundaries/boundary_conditions.f90:52-136
 * (types 0, 201, 400, 401). */
/* characteristic_rhs_k1d, src/boundaries/nonreflecting.f90:237-341 (no
 * shape): dq = prim(aim) - prim(q) + (rhs dt, rhs_m dt / rho, rhs_5 dt), keep
 * the incoming waves, rhs <- dq / dt (momentum: times the updated rho). */
/* sh: the shape of characteristic_rhs_k4d (nonreflecting.f90:431-439, type
 * 206), which blends (1 - shape) rhs + shape (new rhs); null: no shape. */
static void char_rhs(int np, int sv, double *rhs, const double *q, const double *aim,
                     const double k[3], double gamma, double dt, const double *sh) {
  const double inv_dt = 1.0 / dt;
  for (int m = 0; m < np; ++m) {
    double *F = rhs + m;
    const double *Q = q + m, *A = aim + m;
    double r = Q[0];                                                /* give_r */
    const double u[3] = {Q[sv] / Q[0], Q[2 * sv] / Q[0], Q[3 * sv] / Q[0]}; /* give_u */
    const double c = sqrt(gamma * Q[4 * sv] / Q[0]);                /* give_c */
    double dq[5];
    dq[0] = A[0] - r; /* :274-278 */
    for (int d = 0; d < 3; ++d) dq[1 + d] = A[(1 + d) * sv] / A[0] - u[d];
    dq[4] = A[4 * sv] - Q[4 * sv];
    double tmp = dt / r; /* :280-285 */
    dq[0] = dq[0] + F[0] * dt;
    for (int d = 0; d < 3; ++d) dq[1 + d] = dq[1 + d] + F[(1 + d) * sv] * tmp;
    dq[4] = dq[4] + F[4 * sv] * dt;
    const double kx = -k[0], ky = -k[1], kz = -k[2]; /* :288-290 */
    double R1 = kx * dq[0] + kz * dq[2] - ky * dq[3] - kx * dq[4] / (c * c); /* :293-297 */
    double R2 = ky * dq[0] - kz * dq[1] + kx * dq[3] - ky * dq[4] / (c * c);
    double R3 = kz * dq[0] + ky * dq[1] - kx * dq[2] - kz * dq[4] / (c * c);
    double R4 = kx * dq[1] + ky * dq[2] + kz * dq[3] + dq[4] / (r * c);
    double R5 = -kx * dq[1] - ky * dq[2] - kz * dq[3] + dq[4] / (r * c);
    const double lamS = kx * u[0] + ky * u[1] + kz * u[2]; /* :299-310 */
    const double lamP = lamS + c, lamM = lamS - c;
    if (lamS < 0.0) { R1 = 0.0; R2 = 0.0; R3 = 0.0; }
    if (lamP < 0.0) R4 = 0.0;
    if (lamM < 0.0) R5 = 0.0;
    dq[0] = kx * R1 + ky * R2 + kz * R3 + r / (2.0 * c) * R4 + r / (2.0 * c) * R5; /* :313-318 */
    dq[1] = -kz * R2 + ky * R3 + 0.5 * kx * R4 - 0.5 * kx * R5;
    dq[2] = kz * R1 - kx * R3 + 0.5 * ky * R4 - 0.5 * ky * R5;
    dq[3] = -ky * R1 + kx * R2 + 0.5 * kz * R4 - 0.5 * kz * R5;
    dq[4] = +0.5 * r * c * R4 + 0.5 * r * c * R5;
    r = r + dq[0]; /* :321 */
    if (sh) {
      const double s = sh[m];
      tmp = s * inv_dt;
      F[0] = (1.0 - s) * F[0] + dq[0] * tmp;
      F[4 * sv] = (1.0 - s) * F[4 * sv] + dq[4] * tmp;
      tmp = tmp * r;
      for (int d = 0; d < 3; ++d) F[(1 + d) * sv] = (1.0 - s) * F[(1 + d) * sv] + dq[1 + d] * tmp;
      continue;
    }
    F[0] = dq[0] * inv_dt; /* :333-339 */
    tmp = inv_dt * r;
    for (int d = 0; d < 3; ++d) F[(1 + d) * sv] = dq[1 + d] * tmp;
    F[4 * sv] = dq[4] * inv_dt;
  }
}

void orc_characteristic_rhs_k1d(int np, int sv, double *rhs, const double *q, const double *aim,
                                const double k[3], double gamma, double dt) {
  char_rhs(np, sv, rhs, q, aim, k, gamma, dt, NULL);
}

/* init_boundary_conditions (boundary_conditions.f90:434-438): the last type
 * 203 row's column 8 picks where its rows act: -1 only in the call after the
 * RK step (outside_rhs), 1 only in the RHS's own call, 0 both. */
static int zone_applies(const orc_cfg *c, int outside) {
  int flag = 0;
  for (int r = 0; r < c->nbc; ++r)
    if ((int)lround(c->bc[r][0]) == 203) flag = (int)lround(c->bc[r][7]);
  return outside ? flag != 1 : flag != -1;
}

/* characteristic_zero_grad_init, src/boundaries/nonreflecting.f90:665-728:
 * offset of the inner neighbour (one point against the sign of each
 * component of the normalised kvector). */
static long zero_grad_shift(const orc_cfg *c, int r) {
  const double *kv = &c->bc[r][7];
  const double kn = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  long s[3];
  for (int d = 0; d < 3; ++d) {
    const double k = kv[d] / kn;
    s[d] = k > 0 ? -1 : (k < 0 ? 1 : 0);
  }
  return s[0] + (long)c->n[0] * (s[1] + (long)c->n[1] * s[2]);
}

/* outside: the call after the RK step (set_boundary_condition_q(qn,
 * outside_rhs=.true.)), which type 203 rows tell apart. */
static void bc_q_at(const orc_cfg *c, double *q, const double *qref, const double *shape,
                    int outside) {
  const size_t N = npts(c);
  for (int r = 0; r < c->nbc; ++r) {
    const int t = bc_type(c, r);
    if (t == 0 || t == 100 || t == 204 || t == 205 || t == 206 || t == 400) continue; /* RHS rows */
    if (t == 203) {
      /* characteristic_q_k4d(q, qref, kvector, shape) on the whole domain, :93-108 */
      if (!zone_applies(c, outside)) continue;
      const double *kv = c->kvector;
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) {
        const double k[3] = {kv[i], kv[N + i], kv[2 * N + i]};
        char_q(1, (int)N, q + i, qref + i, k, c->gamma, shape[i]);
      }
      continue;
    }
    if (t == 201) {
      const double k[3] = {c->bc[r][7], c->bc[r][8], c->bc[r][9]};
      FOR_BOX(c, r, idx) orc_characteristic_q_k1d(1, (int)N, q + idx, qref + idx, k, c->gamma);
    } else if (t == 202) {
      /* characteristic_q_k4d(q(box), qref(box), kvector(box,:)), boundary_conditions.f90:89-91 */
      const double *kv = c->kvector;
      FOR_BOX(c, r, idx) {
        const double k[3] = {kv[idx], kv[N + idx], kv[2 * N + idx]};
        orc_characteristic_q_k1d(1, (int)N, q + idx, qref + idx, k, c->gamma);
      }
    } else if (t == 200) {
      /* characteristic_q(q(box), q(box of inner neighbours), row(8:10)), boundary_conditions.f90:80-83 */
      const double k[3] = {c->bc[r][7], c->bc[r][8], c->bc[r][9]};
      const long sh = zero_grad_shift(c, r);
      FOR_BOX(c, r, idx) orc_characteristic_q_k1d(1, (int)N, q + idx, q + idx + sh, k, c->gamma);
    } else if (t == 401) {
      /* sponge_q with qref, src/boundaries/sponge_layer_core.f90:70-103 */
      for (int v = 0; v < NV; ++v)
        #pragma omp parallel for schedule(static)
        for (size_t i = 0; i < N; ++i)
          q[v * N + i] = q[v * N + i] - shape[i] * (q[v * N + i] - qref[v * N + i]);
    } else {
      bc_unsupported(t, "set_boundary_condition_q");
    }
  }
}

void orc_bc_q(const orc_cfg *c, double *q, const double *qref, const double *shape) {
  bc_q_at(c, q, qref, shape, 0);
}

/* set_boundary_condition_qs, boundary_conditions.f90:140-224 (types 0, 201, 400, 401). */
static void bc_qs_at(const orc_cfg *c, double *qs, const double *q, const double *shape,
                     int outside) {
  const size_t N = npts(c);
  for (int r = 0; r < c->nbc; ++r) {
    const int t = bc_type(c, r);
    if (t == 0 || t == 100 || t == 204 || t == 205 || t == 206 || t == 400) continue; /* RHS rows */
    if (t == 203) {
      /* characteristic_qs_k4d(qs, q, kvector, shape) on the whole domain, :181-194 */
      if (!zone_applies(c, outside)) continue;
      const double *kv = c->kvector;
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) {
        const double k[3] = {kv[i], kv[N + i], kv[2 * N + i]};
        char_qs(1, (int)N, qs + i, q + i, k, c->gamma, shape[i]);
      }
      continue;
    }
    if (t == 201) {
      const double k[3] = {c->bc[r][7], c->bc[r][8], c->bc[r][9]};
      FOR_BOX(c, r, idx) orc_characteristic_qs_k1d(1, (int)N, qs + idx, q + idx, k, c->gamma);
    } else if (t == 202) {
      /* characteristic_qs(qs(box), q(box), kvector(box,:)), boundary_conditions.f90:176-179 */
      const double *kv = c->kvector;
      FOR_BOX(c, r, idx) {
        const double k[3] = {kv[idx], kv[N + idx], kv[2 * N + idx]};
        orc_characteristic_qs_k1d(1, (int)N, qs + idx, q + idx, k, c->gamma);
      }
    } else if (t == 200) {
      /* characteristic_qs(qs(box), q(box of inner neighbours), row(8:10)), boundary_conditions.f90:167-170 */
      const double k[3] = {c->bc[r][7], c->bc[r][8], c->bc[r][9]};
      const long sh = zero_grad_shift(c, r);
      FOR_BOX(c, r, idx) orc_characteristic_qs_k1d(1, (int)N, qs + idx, q + idx + sh, k, c->gamma);
    } else if (t == 401) {
      /* sponge_qs, sponge_layer_core.f90:155-167 */
      for (int v = 0; v < NV; ++v)
        #pragma omp parallel for schedule(static)
        for (size_t i = 0; i < N; ++i) qs[v * N + i] = qs[v * N + i] - shape[i] * qs[v * N + i];
    } else {
      bc_unsupported(t, "set_boundary_condition_qs");
    }
  }
}

void orc_bc_qs(const orc_cfg *c, double *qs, const double *q, const double *shape) {
  bc_qs_at(c, qs, q, shape, 0);
}

/* set_boundary_condition_rhs, boundary_conditions.f90:228-298 (types 0, 100, 201, 204, 400, 401). */
void orc_bc_rhs(const orc_cfg *c, double *rhs, const double *q, const double *qref,
                const double *shape) {
  const size_t N = npts(c);
  for (int r = 0; r < c->nbc; ++r) {
    const int t = bc_type(c, r);
    if (t == 0 || t == 200 || t == 201 || t == 202 || t == 203 || t == 401) continue;
    if (t == 206) {
      /* characteristic_rhs_k4d(rhs, q, qref, kvector, shape) on the whole domain, :276-279 */
      const double *kv = c->kvector;
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) {
        const double k[3] = {kv[i], kv[N + i], kv[2 * N + i]};
        char_rhs(1, (int)N, rhs + i, q + i, qref + i, k, c->gamma, c->dt, shape + i);
      }
      continue;
    }
    if (t == 400) {
      /* sponge_rhs with qref, sponge_layer_core.f90:34-66 */
      for (int v = 0; v < NV; ++v)
        #pragma omp parallel for schedule(static)
        for (size_t i = 0; i < N; ++i)
          rhs[v * N + i] = rhs[v * N + i] - shape[i] * (q[v * N + i] - qref[v * N + i]);
    } else if (t == 100) {
      /* wall_slip_no_thermodynamics(rhs(box), nint(row(8))), wall_boundary_conditions.f90:13-23 */
      const int v = (int)lround(c->bc[r][7]) - 1;
      FOR_BOX(c, r, idx) rhs[v * N + idx] = 0.0;
    } else if (t == 205) {
      /* characteristic_rhs_k4d(rhs(box), q(box), qref(box), kvector(box,:)), boundary_conditions.f90:272-274 */
      const double *kv = c->kvector;
      FOR_BOX(c, r, idx) {
        const double k[3] = {kv[idx], kv[N + idx], kv[2 * N + idx]};
        orc_characteristic_rhs_k1d(1, (int)N, rhs + idx, q + idx, qref + idx, k, c->gamma, c->dt);
      }
    } else if (t == 204) {
      /* characteristic_rhs_k1d(rhs(box), q(box), qref(box), row(8:10)) */
      const double k[3] = {c->bc[r][7], c->bc[r][8], c->bc[r][9]};
      FOR_BOX(c, r, idx) orc_characteristic_rhs_k1d(1, (int)N, rhs + idx, q + idx, qref + idx, k, c->gamma, c->dt);
    } else {
      bc_unsupported(t, "set_boundary_condition_rhs");
    }
  }
}

/* set_boundary_condition_rhss (boundary_conditions.f90:302-377).  Type 100
 * zeroes its variable of the adjoint RHS on its box (:319-320).  Type 400:
 * called without q (ns_rho_rhou_p.f90:518, q == NULL here) it resolves to the
 * empty sponge_rhss (sponge_layer_core.f90:189-196); with q (the Euler
 * adjoint, euler_rho_rhou_p.f90:384) to sponge_rhss_correct (:107-151), on
 * qs = the adjoint state after set_boundary_condition_qs. */
static void bc_rhss(const orc_cfg *c, double *rhs, const double *qs, const double *q,
                    const double *shape) {
  const size_t N = npts(c);
  for (int r = 0; r < c->nbc; ++r) {
    const int t = bc_type(c, r);
    /* the cases 204-206 are commented out there (:339-353), so those rows
     * reach case default, which stops the run */
    if (t == 400 && q) {
      const double gm1 = c->gamma - 1.0;
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) {
        const double rho = q[i]; /* give_r, give_u */
        const double u1 = q[N + i] / q[i], u2 = q[2 * N + i] / q[i], u3 = q[3 * N + i] / q[i];
        double sp[5];
        sp[0] = 1.0 * qs[i] - u1 / rho * qs[N + i] - u2 / rho * qs[2 * N + i] - u3 / rho * qs[3 * N + i];
        sp[1] = 1.0 / rho * qs[N + i];
        sp[2] = 1.0 / rho * qs[2 * N + i];
        sp[3] = 1.0 / rho * qs[3 * N + i];
        sp[4] = gm1 * qs[4 * N + i];
        for (int v = 0; v < 5; ++v) rhs[v * N + i] = rhs[v * N + i] + shape[i] * sp[v];
      }
      continue;
    }
    if (t == 0 || t == 200 || t == 201 || t == 202 || t == 203 || t == 400 || t == 401) continue;
    if (t == 100) {
      /* wall_slip_no_thermodynamics(rhss(box), nint(row(8))), wall_boundary_conditions.f90:13-23 */
      const int v = (int)lround(c->bc[r][7]) - 1;
      FOR_BOX(c, r, idx) rhs[v * N + idx] = 0.0;
      continue;
    }
    bc_unsupported(t, "set_boundary_condition_rhss");
  }
}

/* src/boundaries/shape_functions.f90:45-63 (shape fcts) and :69-150 (composition). */
void orc_sponge_shape(const orc_cfg *c, const double x0[3], const double x1[3],
                      const double len[3], double amplitude, int cos_shape,
                      const double *X1, const double *X2, const double *X3, double *shape) {
  const double pi = 4.0 * atan(1.0);
  const double *X[3] = {X1, X2, X3};
  const int n1 = c->n[0], n2 = c->n[1], n3 = c->n[2];
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        const double xv[3] = {X[0][i], X[1][j], X[2][k]};
        double phi = 1.0;
        for (int d = 0; d < 3; ++d) {
          for (int side = 0; side < 2; ++side) {
            const double xo = side == 0 ? x0[d] + len[d] : x1[d] - len[d];
            double s = cos_shape ? 0.5 + 0.5 * cos((xv[d] - xo) * (pi / len[d]))
                                 : 1.0 - ((xv[d] - xo) / len[d]) * ((xv[d] - xo) / len[d]);
            if (side == 0 && xv[d] > xo) s = 1.0;
            if (side == 1 && xv[d] < xo) s = 1.0;
            phi = phi * s;
          }
        }
        phi = 1.0 - phi;
        shape[(size_t)i + (size_t)n1 * (j + (size_t)n2 * k)] = amplitude * phi;
      }
}

/* ------------------------------------------------------------------------ */
/* Euler right-hand side, src/equations/euler_rho_rhou_p.f90:30-122 (body :53-108): mass
 * -D(rho u_d); momentum -D(rho u_d u_i) - p_i; pressure
 * -gamma D(p u_d) + (gamma - 1) u.grad(p); force (no-op), BC rows. */
void orc_rhs_direct_euler(const orc_cfg *c, double *rhs, const double *q_in, const double *qref,
                          const double *shape) {
  const size_t N = npts(c);
  const double dxs[3] = {c->dx[0], c->dx[1], c->dx[2]};
  double *q = xalloc(5 * N), *u = xalloc(3 * N), *tmp = xalloc(N);
  double *a[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *pg[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *r[5];
  for (int v = 0; v < 5; ++v) r[v] = rhs + v * N;
  memcpy(q, q_in, 5 * N * sizeof(double)); /* :53 */
  orc_bc_q(c, q, qref, shape);              /* :56 */
  for (int m = 0; m < 3; ++m)               /* :59-63 */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) u[m * N + i] = q[(m + 1) * N + i] / q[i];
  for (int k = 0; k < 3; ++k) orc_diff3d(c, k, a[k], q + (k + 1) * N, dxs[k]); /* mass :66-69 */
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) r[0][i] = -a[0][i] - a[1][i] - a[2][i];
  for (int m = 0; m < 3; ++m) { /* momentum :73-95 */
    for (int k = 0; k < 3; ++k) {
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) tmp[i] = q[(k + 1) * N + i] * u[m * N + i];
      orc_diff3d(c, k, a[k], tmp, dxs[k]);
    }
    orc_diff3d(c, m, pg[m], q + 4 * N, dxs[m]);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) r[m + 1][i] = -a[0][i] - a[1][i] - a[2][i] - pg[m][i];
  }
  for (int k = 0; k < 3; ++k) { /* pressure :97-102 */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = q[4 * N + i] * u[k * N + i];
    orc_diff3d(c, k, a[k], tmp, dxs[k]);
  }
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    r[4][i] = -c->gamma * (a[0][i] + a[1][i] + a[2][i]);
    r[4][i] = r[4][i] + (c->gamma - 1.0) * (u[i] * pg[0][i] + u[N + i] * pg[1][i] + u[2 * N + i] * pg[2][i]);
  }
  orc_bc_rhs(c, rhs, q, qref, shape); /* force (no-op) :105, BC rows :108 */
  free(q); free(u); free(tmp);
  for (int k = 0; k < 3; ++k) { free(a[k]); free(pg[k]); }
}

/* Euler right-hand side on a deformed grid,
 * src/equations/euler_rho_rhou_p_grid_transform.f90:33-140: projected
 * velocities uh_d = sum_j M(d,j) u_j (:75-77), physical pressure gradient
 * p_j = sum_d M(d,j) D_d p (:80-85); mass -J sum_d D_d(rho uh_d); momentum
 * -J (sum_d D_d(rho u_i uh_d) + p_i); pressure
 * J (gamma - 1) (gamma/(gamma - 1) (-sum_d D_d(p uh_d)) + u.grad p)
 * (:118-128); force (no-op), sponge (empty), BC rows (:133-139).
 * With phi, euler_rho_rhou_p_ev_phi_grid_transform.f90:33-160: phi inside the
 * flux derivatives (phi rho uh_d, phi rho u_i uh_d, phi p uh_d) and 1/phi on
 * their sums; pressure J (-gamma/phi sum_d D_d(phi p uh_d) + (gamma - 1)
 * u.grad p); apply_volume_penalization is the reference's no-op. */
#define MET_E(Xm, a, b, N) ((Xm) + (size_t)(N) * ((a) + 3 * (b)))
static void euler_gt_impl(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                          const double *Xj, const double *phi, const double *inv_phi,
                          const double *qref, const double *shape) {
  const size_t N = npts(c);
  const double dxs[3] = {c->dx[0], c->dx[1], c->dx[2]};
  double *q = xalloc(5 * N), *u = xalloc(3 * N), *uh = xalloc(3 * N), *tmp = xalloc(N);
  double *a[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *pg[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *r[5];
  for (int v = 0; v < 5; ++v) r[v] = rhs + v * N;
  memcpy(q, q_in, 5 * N * sizeof(double)); /* :59 */
  orc_bc_q(c, q, qref, shape);              /* :62 */
  for (int m = 0; m < 3; ++m)               /* :65-69 */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) u[m * N + i] = q[(m + 1) * N + i] / q[i];
  for (int d = 0; d < 3; ++d) {             /* :75-77 */
    const double *m0 = MET_E(Xm, d, 0, N), *m1 = MET_E(Xm, d, 1, N), *m2 = MET_E(Xm, d, 2, N);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) uh[d * N + i] = m0[i] * u[i] + m1[i] * u[N + i] + m2[i] * u[2 * N + i];
  }
  for (int d = 0; d < 3; ++d) orc_diff3d(c, d, a[d], q + 4 * N, dxs[d]); /* :80-85 */
  for (int j = 0; j < 3; ++j) {
    const double *m0 = MET_E(Xm, 0, j, N), *m1 = MET_E(Xm, 1, j, N), *m2 = MET_E(Xm, 2, j, N);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) pg[j][i] = m0[i] * a[0][i] + m1[i] * a[1][i] + m2[i] * a[2][i];
  }
#define PHI_(i) (phi ? phi[i] : 1.0)
  for (int d = 0; d < 3; ++d) { /* mass :88-93 (ev: phi_rho uh) */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = phi ? phi[i] * q[i] * uh[d * N + i] : q[i] * uh[d * N + i];
    orc_diff3d(c, d, a[d], tmp, dxs[d]);
  }
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i)
    r[0][i] = phi ? Xj[i] * inv_phi[i] * (-a[0][i] - a[1][i] - a[2][i]) : Xj[i] * (-a[0][i] - a[1][i] - a[2][i]);
  for (int m = 0; m < 3; ++m) { /* momentum :96-117 (ev: phi_rho_u_m uh) */
    for (int d = 0; d < 3; ++d) {
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i)
        tmp[i] = phi ? PHI_(i) * q[(m + 1) * N + i] * uh[d * N + i] : q[i] * u[m * N + i] * uh[d * N + i];
      orc_diff3d(c, d, a[d], tmp, dxs[d]);
    }
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i)
      r[m + 1][i] = phi ? Xj[i] * (inv_phi[i] * (-a[0][i] - a[1][i] - a[2][i]) - pg[m][i])
                        : Xj[i] * (-a[0][i] - a[1][i] - a[2][i] - pg[m][i]);
  }
  for (int d = 0; d < 3; ++d) { /* pressure :120-128 (ev: phi_p uh) */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = PHI_(i) * q[4 * N + i] * uh[d * N + i];
    orc_diff3d(c, d, a[d], tmp, dxs[d]);
  }
#undef PHI_
  const double gt = c->gamma / (c->gamma - 1.0), gm1 = c->gamma - 1.0;
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    const double up = u[i] * pg[0][i] + u[N + i] * pg[1][i] + u[2 * N + i] * pg[2][i];
    if (phi) {
      double v = -c->gamma * inv_phi[i] * (a[0][i] + a[1][i] + a[2][i]);
      v = v + gm1 * up;
      r[4][i] = v * Xj[i];
    } else {
      double v = gt * (-a[0][i] - a[1][i] - a[2][i]);
      v = v + up;
      r[4][i] = Xj[i] * gm1 * v;
    }
  }
  orc_bc_rhs(c, rhs, q, qref, shape); /* :133-139 */
  free(q); free(u); free(uh); free(tmp);
  for (int k = 0; k < 3; ++k) { free(a[k]); free(pg[k]); }
}
#undef MET_E

void orc_rhs_direct_euler_gt(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                             const double *Xj, const double *qref, const double *shape) {
  euler_gt_impl(c, rhs, q_in, Xm, Xj, NULL, NULL, qref, shape);
}

void orc_rhs_direct_euler_ev_gt(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                                const double *Xj, const double *phi, const double *inv_phi,
                                const double *qref, const double *shape) {
  euler_gt_impl(c, rhs, q_in, Xm, Xj, phi, inv_phi, qref, shape);
}

/* ------------------------------------------------------------------------ */
/* Direct NS right-hand side, src/equations/ns_rho_rhou_p.f90:31-253.       */
void orc_rhs_direct(const orc_cfg *c, double *rhs, const double *q_in, const double *qref,
                    const double *shape, int heat_unit_spacing) {
  if (c->equations == ORC_EQ_EULER && c->penal_phi && c->X_metric) {
    /* => euler_rho_rhou_p_ev_phi_grid_transform */
    orc_rhs_direct_euler_ev_gt(c, rhs, q_in, c->X_metric, c->X_jacobian, c->penal_phi,
                               c->penal_inv_phi, qref, shape);
    return;
  }
  if (c->equations == ORC_EQ_EULER && c->penal_phi) {
    /* the reference has no Cartesian euler_rho_rhou_p_ev_phi module */
    fprintf(stderr, "oracle: no penalised Euler module for Cartesian grids\n");
    abort();
  }
  if (c->equations == ORC_EQ_EULER && c->X_metric) { /* => euler_rho_rhou_p_grid_transform */
    orc_rhs_direct_euler_gt(c, rhs, q_in, c->X_metric, c->X_jacobian, qref, shape);
    return;
  }
  if (c->equations == ORC_EQ_EULER) { /* governing_equations => euler_rho_rhou_p */
    orc_rhs_direct_euler(c, rhs, q_in, qref, shape);
    return;
  }
  if (c->X_metric && c->penal_phi) { /* => ns_rho_rhou_p_ev_phi_grid_transform */
    orc_rhs_direct_ev_gt(c, rhs, q_in, c->X_metric, c->X_jacobian, c->penal_phi, c->penal_inv_phi,
                         qref, shape);
    return;
  }
  if (c->X_metric) { /* governing_equations => ns_rho_rhou_p_grid_transform */
    orc_rhs_direct_gt(c, rhs, q_in, c->X_metric, c->X_jacobian, qref, shape);
    return;
  }
  if (c->penal_phi) { /* governing_equations => ns_rho_rhou_p_ev_phi */
    orc_rhs_direct_ev(c, rhs, q_in, c->penal_phi, c->penal_inv_phi, qref, shape);
    return;
  }
  const size_t N = npts(c);
  const double dx1 = c->dx[0], dx2 = c->dx[1], dx3 = c->dx[2];
  const double hx1 = heat_unit_spacing ? 1.0 : dx1;
  const double hx2 = heat_unit_spacing ? 1.0 : dx2;
  const double hx3 = heat_unit_spacing ? 1.0 : dx3;
  double *q = xalloc(5 * N), *u = xalloc(3 * N);
  double *ax = xalloc(N), *ay = xalloc(N), *az = xalloc(N);
  double *px = xalloc(N), *py = xalloc(N), *pz = xalloc(N);
  double *ux = xalloc(N), *uy = xalloc(N), *uz = xalloc(N);
  double *vx = xalloc(N), *vy = xalloc(N), *vz = xalloc(N);
  double *wx = xalloc(N), *wy = xalloc(N), *wz = xalloc(N);
  double *t11 = xalloc(N), *t12 = xalloc(N), *t13 = xalloc(N);
  double *t22 = xalloc(N), *t23 = xalloc(N), *t33 = xalloc(N);
  double *temp = xalloc(N), *mu = xalloc(N), *mudiv = xalloc(N), *lam = xalloc(N);
  double *tmp = xalloc(N);
  double *r1 = rhs, *r2 = rhs + N, *r3 = rhs + 2 * N, *r4 = rhs + 3 * N, *r5 = rhs + 4 * N;

  memcpy(q, q_in, 5 * N * sizeof(double)); /* :89 */
  orc_bc_q(c, q, qref, shape);              /* :92 */
  const double *q1 = q, *q2 = q + N, *q3 = q + 2 * N, *q4 = q + 3 * N, *q5 = q + 4 * N;
  double *u1 = u, *u2 = u + N, *u3 = u + 2 * N;
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) { /* :95-99 */
    u1[i] = q2[i] / q1[i];
    u2[i] = q3[i] / q1[i];
    u3[i] = q4[i] / q1[i];
  }
  /* :102-112 */
  orc_diff3d(c, 0, ux, u1, dx1);
  orc_diff3d(c, 1, uy, u1, dx2);
  orc_diff3d(c, 2, uz, u1, dx3);
  orc_diff3d(c, 0, vx, u2, dx1);
  orc_diff3d(c, 1, vy, u2, dx2);
  orc_diff3d(c, 2, vz, u2, dx3);
  orc_diff3d(c, 0, wx, u3, dx1);
  orc_diff3d(c, 1, wy, u3, dx2);
  orc_diff3d(c, 2, wz, u3, dx3);
  /* :115-128 */
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    temp[i] = orc_T(c, q1[i], q5[i]);
    mu[i] = orc_mu(c, temp[i]);
    mudiv[i] = -2.0 / 3.0 * mu[i] * (ux[i] + vy[i] + wz[i]);
    t11[i] = mudiv[i] + 2 * mu[i] * ux[i];
    t22[i] = mudiv[i] + 2 * mu[i] * vy[i];
    t33[i] = mudiv[i] + 2 * mu[i] * wz[i];
    t12[i] = mu[i] * (uy[i] + vx[i]);
    t13[i] = mu[i] * (uz[i] + wx[i]);
    t23[i] = mu[i] * (vz[i] + wy[i]);
  }
  /* mass :131-135 */
  orc_diff3d(c, 0, ax, q2, dx1);
  orc_diff3d(c, 1, ay, q3, dx2);
  orc_diff3d(c, 2, az, q4, dx3);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) r1[i] = -ax[i] - ay[i] - az[i];

  /* momentum :138-177 */
  const double *qd[3] = {q2, q3, q4};
  double *pd[3] = {px, py, pz};
  double *rm[3] = {r2, r3, r4};
  const double *tau[3][3] = {{t11, t12, t13}, {t12, t22, t23}, {t13, t23, t33}};
  const double dxs[3] = {dx1, dx2, dx3};
  for (int m = 0; m < 3; ++m) {
    const double *um = u + m * N;
    double *a[3] = {ax, ay, az};
    for (int d = 0; d < 3; ++d) {
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) tmp[i] = qd[d][i] * um[i];
      orc_diff3d(c, d, a[d], tmp, dxs[d]);
    }
    orc_diff3d(c, m, pd[m], q5, dxs[m]);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) rm[m][i] = -ax[i] - ay[i] - az[i] - pd[m][i];
    for (int d = 0; d < 3; ++d) orc_diff3d(c, d, a[d], tau[m][d], dxs[d]);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) rm[m][i] = rm[m][i] + ax[i] + ay[i] + az[i];
  }

  /* energy (pressure) :180-209 */
  {
    double *a[3] = {ax, ay, az};
    for (int d = 0; d < 3; ++d) {
      const double *ud = u + d * N;
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) tmp[i] = q5[i] * ud[i];
      orc_diff3d(c, d, a[d], tmp, dxs[d]);
    }
    const double gt = c->gamma / (c->gamma - 1.0);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      r5[i] = gt * (-ax[i] - ay[i] - az[i]);
      r5[i] = r5[i] + u1[i] * px[i] + u2[i] * py[i] + u3[i] * pz[i];
    }
    /* heat: Dx without h (dx = 1) in the reference, :189-197 */
    orc_diff3d(c, 0, ax, temp, hx1);
    orc_diff3d(c, 1, ay, temp, hx2);
    orc_diff3d(c, 2, az, temp, hx3);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) lam[i] = orc_lambda(c, mu[i]);
    const double hx[3] = {hx1, hx2, hx3};
    for (int d = 0; d < 3; ++d) {
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) tmp[i] = lam[i] * a[d][i];
      orc_diff3d(c, d, a[d], tmp, hx[d]);
    }
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      r5[i] = r5[i] + ax[i] + ay[i] + az[i];
      r5[i] = r5[i] + t11[i] * ux[i] + t12[i] * uy[i] + t13[i] * uz[i] + t12[i] * vx[i] +
              t22[i] * vy[i] + t23[i] * vz[i] + t13[i] * wx[i] + t23[i] * wy[i] + t33[i] * wz[i];
      r5[i] = (c->gamma - 1.0) * r5[i];
    }
  }
  /* force (no-op stub src/opt/force.f90), in-rhs sponge (empty, sponge_layer.f90:15) */
  orc_bc_rhs(c, rhs, q, qref, shape); /* :218 */

  free(q); free(u); free(ax); free(ay); free(az); free(px); free(py); free(pz);
  free(ux); free(uy); free(uz); free(vx); free(vy); free(vz); free(wx); free(wy); free(wz);
  free(t11); free(t12); free(t13); free(t22); free(t23); free(t33);
  free(temp); free(mu); free(mudiv); free(lam); free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Deformed grids.  X_metric(:,:,:,k,j) is stored at X_metric + N*(k + 3*j)  */
/* (0-based k, j; Fortran order, k fastest).                                  */
#define MET(Mp, k, j, N) ((Mp) + (N) * ((size_t)(k) + 3 * (size_t)(j)))

/* src/geometry/grid_metric.f90:16-188 (Visbal & Gaitonde's conservative form
 * of the metric, with the periodic-direction corrections c1..c3). */
void orc_calc_metric(const orc_cfg *c, const double *X, const double *X_calc, double *X_metric,
                     double *X_jacobian) {
  const size_t N = npts(c);
  const double c1 = c->scheme[0] == ORC_PERIODIC ? 1.0 : 0.0; /* :41-55 */
  const double c2 = c->scheme[1] == ORC_PERIODIC ? 1.0 : 0.0;
  const double c3 = c->scheme[2] == ORC_PERIODIC ? 1.0 : 0.0;
  const double dx1 = c->dx[0], dx2 = c->dx[1], dx3 = c->dx[2];
  double *Xt = xalloc(3 * N); /* :68-70 */
  const double cc[3] = {c1, c2, c3};
  for (int d = 0; d < 3; ++d)
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) Xt[d * N + i] = X[d * N + i] - cc[d] * X_calc[d * N + i];
  const double *X1 = Xt, *X2 = Xt + N, *X3 = Xt + 2 * N;
  double *d11 = xalloc(N), *d12 = xalloc(N), *d13 = xalloc(N); /* dx1dxi1, dx1dxi2, dx1dxi3 */
  double *d21 = xalloc(N), *d22 = xalloc(N), *d23 = xalloc(N);
  double *d31 = xalloc(N), *d32 = xalloc(N), *d33 = xalloc(N);
  double *t1 = xalloc(N), *t2 = xalloc(N), *pr = xalloc(N);
  /* :86-96 */
  orc_diff3d(c, 0, d11, X1, dx1);
  orc_diff3d(c, 0, d21, X2, dx1);
  orc_diff3d(c, 0, d31, X3, dx1);
  orc_diff3d(c, 1, d12, X1, dx2);
  orc_diff3d(c, 1, d22, X2, dx2);
  orc_diff3d(c, 1, d32, X3, dx2);
  orc_diff3d(c, 2, d13, X1, dx3);
  orc_diff3d(c, 2, d23, X2, dx3);
  orc_diff3d(c, 2, d33, X3, dx3);
#define PROD(a, b) for (size_t i = 0; i < N; ++i) pr[i] = (a)[i] * (b)[i]
  double *M;
  /* m11 :116-118 */
  PROD(d22, X3); orc_diff3d(c, 2, t1, pr, dx3);
  PROD(d23, X3); orc_diff3d(c, 1, t2, pr, dx2);
  M = MET(X_metric, 0, 0, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] + c2 * d33[i] + c3 * d22[i] + c2 * c3;
  /* m12 :124-126 */
  PROD(d13, X3); orc_diff3d(c, 1, t1, pr, dx2);
  PROD(d12, X3); orc_diff3d(c, 2, t2, pr, dx3);
  M = MET(X_metric, 0, 1, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] - c3 * d12[i];
  /* m13 :132-134 */
  PROD(d12, X2); orc_diff3d(c, 2, t1, pr, dx3);
  PROD(d13, X2); orc_diff3d(c, 1, t2, pr, dx2);
  M = MET(X_metric, 0, 2, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] - c2 * d13[i];
  /* m21 :139-141 */
  PROD(d23, X3); orc_diff3d(c, 0, t1, pr, dx1);
  PROD(d21, X3); orc_diff3d(c, 2, t2, pr, dx3);
  M = MET(X_metric, 1, 0, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] - c3 * d21[i];
  /* m22 :147-149 */
  PROD(d11, X3); orc_diff3d(c, 2, t1, pr, dx3);
  PROD(d13, X3); orc_diff3d(c, 0, t2, pr, dx1);
  M = MET(X_metric, 1, 1, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] + c1 * d33[i] + c3 * d11[i] + c1 * c3;
  /* m23 :155-157 */
  PROD(d13, X2); orc_diff3d(c, 0, t1, pr, dx1);
  PROD(d11, X2); orc_diff3d(c, 2, t2, pr, dx3);
  M = MET(X_metric, 1, 2, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] - c1 * d23[i];
  /* m31 :162-164 */
  PROD(X3, d21); orc_diff3d(c, 1, t1, pr, dx2);
  PROD(X3, d22); orc_diff3d(c, 0, t2, pr, dx1);
  M = MET(X_metric, 2, 0, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] - c2 * d31[i];
  /* m32 :169-171 */
  PROD(d12, X3); orc_diff3d(c, 0, t1, pr, dx1);
  PROD(d11, X3); orc_diff3d(c, 1, t2, pr, dx2);
  M = MET(X_metric, 2, 1, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] - c1 * d32[i];
  /* m33 :176-178 */
  PROD(d11, X2); orc_diff3d(c, 1, t1, pr, dx2);
  PROD(d12, X2); orc_diff3d(c, 0, t2, pr, dx1);
  M = MET(X_metric, 2, 2, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) M[i] = t1[i] - t2[i] + c1 * d22[i] + d11[i] * c2 + c1 * c2;
#undef PROD
  /* Jacobian :182-185 */
  const double *M31 = MET(X_metric, 2, 0, N), *M32 = MET(X_metric, 2, 1, N),
               *M33 = MET(X_metric, 2, 2, N);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    X_jacobian[i] = M31[i] * d13[i] + M32[i] * d23[i] + M33[i] * (d33[i] + c3);
    X_jacobian[i] = 1.0 / X_jacobian[i];
  }
  free(Xt); free(d11); free(d12); free(d13); free(d21); free(d22); free(d23);
  free(d31); free(d32); free(d33); free(t1); free(t2); free(pr);
}

/* Direct RHS on a deformed grid, src/equations/ns_rho_rhou_p_grid_transform.f90:31-330.
 * Gradients are formed in physical coordinates without the Jacobian,
 * fluxes with the contravariant velocity u_hat_k = sum_j M(k,j) u_j, and
 * every equation is multiplied by X_jacobian at the end; the heat flux's
 * outer derivatives take no h (:252-268). */
static void rhs_direct_gt_impl(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                               const double *Xj, const double *phi, const double *inv_phi,
                               const double *qref, const double *shape);

void orc_rhs_direct_gt(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                       const double *Xj, const double *qref, const double *shape) {
  rhs_direct_gt_impl(c, rhs, q_in, Xm, Xj, NULL, NULL, qref, shape);
}

/* Volume-penalised deformed-grid RHS, src/equations/ns_rho_rhou_p_ev_phi_grid_transform.f90:
 * the grid-transform RHS with phi inside every flux derivative (phi rho u_hat,
 * phi rho u_i u_hat, phi tau_proj, phi p u_hat, phi lambda J T_j :126-202) and
 * div_1_phi on the divergences (:210, :227, :292), not on the pressure
 * gradient, u.grad(p) or J tau:grad(u). */
void orc_rhs_direct_ev_gt(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                          const double *Xj, const double *phi, const double *inv_phi,
                          const double *qref, const double *shape) {
  rhs_direct_gt_impl(c, rhs, q_in, Xm, Xj, phi, inv_phi, qref, shape);
}

/* phi == NULL: ns_rho_rhou_p_grid_transform; else its penalised form. */
static void rhs_direct_gt_impl(const orc_cfg *c, double *rhs, const double *q_in, const double *Xm,
                               const double *Xj, const double *phi, const double *inv_phi,
                               const double *qref, const double *shape) {
  const size_t N = npts(c);
  const double dxs[3] = {c->dx[0], c->dx[1], c->dx[2]};
  double *q = xalloc(5 * N), *u = xalloc(3 * N), *uh = xalloc(3 * N);
  double *a[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *g[3][3]; /* g[i][j] = d(u_i)/dx_j without Jacobian */
  double *pg[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *tg[3] = {xalloc(N), xalloc(N), xalloc(N)};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g[i][j] = xalloc(N);
  double *tau[3][3];
  double *t11 = xalloc(N), *t12 = xalloc(N), *t13 = xalloc(N);
  double *t22 = xalloc(N), *t23 = xalloc(N), *t33 = xalloc(N);
  tau[0][0] = t11; tau[0][1] = t12; tau[0][2] = t13;
  tau[1][0] = t12; tau[1][1] = t22; tau[1][2] = t23;
  tau[2][0] = t13; tau[2][1] = t23; tau[2][2] = t33;
  double *temp = xalloc(N), *mu = xalloc(N), *mudiv = xalloc(N), *lam = xalloc(N);
  double *tmp = xalloc(N);
  double *r[5];
  for (int v = 0; v < 5; ++v) r[v] = rhs + v * N;

  memcpy(q, q_in, 5 * N * sizeof(double)); /* :102 */
  orc_bc_q(c, q, qref, shape);              /* :105 */
  for (int m = 0; m < 3; ++m) /* :108-112 */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) u[m * N + i] = q[(m + 1) * N + i] / q[i];
  /* gradient of f in physical coordinates without Jacobian: out_j = sum_k M(k,j) D_k f */
#define PHYS_GRAD(f, out)                                                                     \
  do {                                                                                        \
    for (int k_ = 0; k_ < 3; ++k_) orc_diff3d(c, k_, a[k_], (f), dxs[k_]);                    \
    for (int j_ = 0; j_ < 3; ++j_) {                                                          \
      const double *m0 = MET(Xm, 0, j_, N), *m1 = MET(Xm, 1, j_, N), *m2 = MET(Xm, 2, j_, N); \
      _Pragma("omp parallel for schedule(static)")                                            \
      for (size_t i = 0; i < N; ++i)                                                          \
        (out)[j_][i] = m0[i] * a[0][i] + m1[i] * a[1][i] + m2[i] * a[2][i];                   \
    }                                                                                         \
  } while (0)
  for (int m = 0; m < 3; ++m) PHYS_GRAD(u + m * N, g[m]); /* :116-135 */
  for (int k = 0; k < 3; ++k) { /* :139-141 */
    const double *m0 = MET(Xm, k, 0, N), *m1 = MET(Xm, k, 1, N), *m2 = MET(Xm, k, 2, N);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i)
      uh[k * N + i] = m0[i] * u[i] + m1[i] * u[N + i] + m2[i] * u[2 * N + i];
  }
  PHYS_GRAD(q + 4 * N, pg); /* :144-149 */
  /* friction :152-165 */
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    temp[i] = orc_T(c, q[i], q[4 * N + i]);
    mu[i] = orc_mu(c, temp[i]);
    mudiv[i] = -2.0 / 3.0 * mu[i] * (g[0][0][i] + g[1][1][i] + g[2][2][i]);
    t11[i] = mudiv[i] + 2 * mu[i] * g[0][0][i];
    t22[i] = mudiv[i] + 2 * mu[i] * g[1][1][i];
    t33[i] = mudiv[i] + 2 * mu[i] * g[2][2][i];
    t12[i] = mu[i] * (g[0][1][i] + g[1][0][i]);
    t13[i] = mu[i] * (g[0][2][i] + g[2][0][i]);
    t23[i] = mu[i] * (g[1][2][i] + g[2][1][i]);
  }
  /* heat :168-182 */
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) lam[i] = orc_lambda(c, mu[i]);
  PHYS_GRAD(temp, tg);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    lam[i] = phi ? phi[i] * lam[i] * Xj[i] : lam[i] * Xj[i]; /* ev: phi lambda J (:199) */
    tg[0][i] = lam[i] * tg[0][i];
    tg[1][i] = lam[i] * tg[1][i];
    tg[2][i] = lam[i] * tg[2][i];
  }
#undef PHYS_GRAD
  /* mass :184-190 */
  for (int k = 0; k < 3; ++k) {
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = (phi ? phi[i] * q[i] : q[i]) * uh[k * N + i];
    orc_diff3d(c, k, a[k], tmp, dxs[k]);
  }
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    r[0][i] = -a[0][i] - a[1][i] - a[2][i];
    r[0][i] = phi ? Xj[i] * inv_phi[i] * r[0][i] : Xj[i] * r[0][i];
  }
  /* momentum :192-241 */
  for (int m = 0; m < 3; ++m) {
    for (int k = 0; k < 3; ++k) {
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i)
        tmp[i] = (phi ? phi[i] * q[(m + 1) * N + i] : q[(m + 1) * N + i]) * uh[k * N + i];
      orc_diff3d(c, k, a[k], tmp, dxs[k]);
    }
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i)
      r[m + 1][i] = phi ? -a[0][i] - a[1][i] - a[2][i] : -a[0][i] - a[1][i] - a[2][i] - pg[m][i];
    for (int k = 0; k < 3; ++k) { /* tau_proj_mk = sum_j M(k,j) tau_mj */
      const double *m0 = MET(Xm, k, 0, N), *m1 = MET(Xm, k, 1, N), *m2 = MET(Xm, k, 2, N);
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) {
        tmp[i] = m0[i] * tau[m][0][i] + m1[i] * tau[m][1][i] + m2[i] * tau[m][2][i];
        if (phi) tmp[i] = phi[i] * tmp[i]; /* ev: phi tau_proj (:222-224) */
      }
      orc_diff3d(c, k, a[k], tmp, dxs[k]);
    }
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      r[m + 1][i] = r[m + 1][i] + a[0][i] + a[1][i] + a[2][i];
      if (phi) r[m + 1][i] = inv_phi[i] * r[m + 1][i] - pg[m][i]; /* ev (:227) */
      r[m + 1][i] = Xj[i] * r[m + 1][i];
    }
  }
  /* pressure :243-276 */
  for (int k = 0; k < 3; ++k) {
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = (phi ? phi[i] * q[4 * N + i] : q[4 * N + i]) * uh[k * N + i];
    orc_diff3d(c, k, a[k], tmp, dxs[k]);
  }
  const double gt = c->gamma / (c->gamma - 1.0);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    r[4][i] = gt * (-a[0][i] - a[1][i] - a[2][i]);
    if (!phi) r[4][i] = r[4][i] + u[i] * pg[0][i] + u[N + i] * pg[1][i] + u[2 * N + i] * pg[2][i];
  }
  for (int j = 0; j < 3; ++j) { /* heat flux in x_j (:252-268): outer Dx without h */
    for (int k = 0; k < 3; ++k) orc_diff3d(c, k, a[k], tg[j], 1.0);
    const double *m0 = MET(Xm, 0, j, N), *m1 = MET(Xm, 1, j, N), *m2 = MET(Xm, 2, j, N);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i)
      r[4][i] = r[4][i] + m0[i] * a[0][i] + m1[i] * a[1][i] + m2[i] * a[2][i];
  }
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) { /* :270-276 */
    if (phi) { /* ev: div_1_phi on the bracket, then u.grad(p) (:292-293) */
      r[4][i] = inv_phi[i] * r[4][i];
      r[4][i] = r[4][i] + u[i] * pg[0][i] + u[N + i] * pg[1][i] + u[2 * N + i] * pg[2][i];
    }
    r[4][i] = r[4][i] + Xj[i] * (t11[i] * g[0][0][i] + t12[i] * g[0][1][i] + t13[i] * g[0][2][i]);
    r[4][i] = r[4][i] + Xj[i] * (t12[i] * g[1][0][i] + t22[i] * g[1][1][i] + t23[i] * g[1][2][i]);
    r[4][i] = r[4][i] + Xj[i] * (t13[i] * g[2][0][i] + t23[i] * g[2][1][i] + t33[i] * g[2][2][i]);
    r[4][i] = Xj[i] * (c->gamma - 1.0) * r[4][i];
  }
  /* force (no-op), in-rhs sponge (empty), BC rhs :278-284 */
  orc_bc_rhs(c, rhs, q, qref, shape);

  free(q); free(u); free(uh);
  for (int k = 0; k < 3; ++k) { free(a[k]); free(pg[k]); free(tg[k]); }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) free(g[i][j]);
  free(t11); free(t12); free(t13); free(t22); free(t23); free(t33);
  free(temp); free(mu); free(mudiv); free(lam); free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Volume-penalised direct RHS, src/equations/ns_rho_rhou_p_ev_phi.f90:33-291:
 * the fluxes carry the porosity phi inside the derivatives, the divergences
 * are multiplied by div_1_phi (the stored 1/phi), the pressure gradient,
 * u.grad(p) and tau:grad(u) are not; both heat-flux derivatives take no h
 * (:212-220).  apply_volume_penalization (:237) is the reference's no-op
 * (src/immersed/volume_penalization.f90), force a no-op, then the BC rows. */
void orc_rhs_direct_ev(const orc_cfg *c, double *rhs, const double *q_in, const double *phi,
                       const double *inv_phi, const double *qref, const double *shape) {
  const size_t N = npts(c);
  const double dxs[3] = {c->dx[0], c->dx[1], c->dx[2]};
  double *q = xalloc(5 * N), *u = xalloc(3 * N), *pru = xalloc(3 * N), *php = xalloc(N);
  double *a[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *pg[3] = {xalloc(N), xalloc(N), xalloc(N)};
  double *g[3][3]; /* g[i][j] = d(u_i)/dx_j */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g[i][j] = xalloc(N);
  double *t11 = xalloc(N), *t12 = xalloc(N), *t13 = xalloc(N);
  double *t22 = xalloc(N), *t23 = xalloc(N), *t33 = xalloc(N);
  double *tau[3][3] = {{t11, t12, t13}, {t12, t22, t23}, {t13, t23, t33}};
  double *temp = xalloc(N), *mu = xalloc(N), *mudiv = xalloc(N), *plam = xalloc(N);
  double *tmp = xalloc(N);
  double *r[5];
  for (int v = 0; v < 5; ++v) r[v] = rhs + v * N;

  memcpy(q, q_in, 5 * N * sizeof(double)); /* :99 */
  orc_bc_q(c, q, qref, shape);              /* :102 */
  for (int m = 0; m < 3; ++m)               /* :105-117 */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      u[m * N + i] = q[(m + 1) * N + i] / q[i];
      pru[m * N + i] = phi[i] * q[(m + 1) * N + i];
    }
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) php[i] = phi[i] * q[4 * N + i];
  for (int m = 0; m < 3; ++m) /* :121-131 */
    for (int k = 0; k < 3; ++k) orc_diff3d(c, k, g[m][k], u + m * N, dxs[k]);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) { /* :134-147 */
    temp[i] = orc_T(c, q[i], q[4 * N + i]);
    mu[i] = orc_mu(c, temp[i]);
    mudiv[i] = -2.0 / 3.0 * mu[i] * (g[0][0][i] + g[1][1][i] + g[2][2][i]);
    t11[i] = mudiv[i] + 2 * mu[i] * g[0][0][i];
    t22[i] = mudiv[i] + 2 * mu[i] * g[1][1][i];
    t33[i] = mudiv[i] + 2 * mu[i] * g[2][2][i];
    t12[i] = mu[i] * (g[0][1][i] + g[1][0][i]);
    t13[i] = mu[i] * (g[0][2][i] + g[2][0][i]);
    t23[i] = mu[i] * (g[1][2][i] + g[2][1][i]);
  }
  for (int k = 0; k < 3; ++k) orc_diff3d(c, k, pg[k], q + 4 * N, dxs[k]); /* :149-151 */
  /* mass :153-159 */
  for (int k = 0; k < 3; ++k) orc_diff3d(c, k, a[k], pru + k * N, dxs[k]);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    r[0][i] = -a[0][i] - a[1][i] - a[2][i];
    r[0][i] = inv_phi[i] * r[0][i];
  }
  /* momentum :161-201 */
  for (int m = 0; m < 3; ++m) {
    for (int k = 0; k < 3; ++k) {
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) tmp[i] = pru[k * N + i] * u[m * N + i];
      orc_diff3d(c, k, a[k], tmp, dxs[k]);
    }
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) r[m + 1][i] = -a[0][i] - a[1][i] - a[2][i];
    for (int k = 0; k < 3; ++k) {
      #pragma omp parallel for schedule(static)
      for (size_t i = 0; i < N; ++i) tmp[i] = phi[i] * tau[m][k][i];
      orc_diff3d(c, k, a[k], tmp, dxs[k]);
    }
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      r[m + 1][i] = r[m + 1][i] + a[0][i] + a[1][i] + a[2][i];
      r[m + 1][i] = inv_phi[i] * r[m + 1][i] - pg[m][i];
    }
  }
  /* pressure :203-234 */
  for (int k = 0; k < 3; ++k) {
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = php[i] * u[k * N + i];
    orc_diff3d(c, k, a[k], tmp, dxs[k]);
  }
  const double gt = c->gamma / (c->gamma - 1.0);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) r[4][i] = gt * (-a[0][i] - a[1][i] - a[2][i]);
  for (int k = 0; k < 3; ++k) orc_diff3d(c, k, a[k], temp, 1.0); /* d T, no h */
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) plam[i] = phi[i] * orc_lambda(c, mu[i]);
  for (int k = 0; k < 3; ++k) {
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = plam[i] * a[k][i];
    orc_diff3d(c, k, a[k], tmp, 1.0);
  }
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    r[4][i] = r[4][i] + a[0][i] + a[1][i] + a[2][i];
    r[4][i] = inv_phi[i] * r[4][i];
    r[4][i] = r[4][i] + u[i] * pg[0][i] + u[N + i] * pg[1][i] + u[2 * N + i] * pg[2][i];
    r[4][i] = r[4][i] + t11[i] * g[0][0][i] + t12[i] * g[0][1][i] + t13[i] * g[0][2][i] +
              t12[i] * g[1][0][i] + t22[i] * g[1][1][i] + t23[i] * g[1][2][i] +
              t13[i] * g[2][0][i] + t23[i] * g[2][1][i] + t33[i] * g[2][2][i];
    r[4][i] = (c->gamma - 1.0) * r[4][i];
  }
  orc_bc_rhs(c, rhs, q, qref, shape); /* :237-246 */

  free(q); free(u); free(pru); free(php);
  for (int k = 0; k < 3; ++k) { free(a[k]); free(pg[k]); }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) free(g[i][j]);
  free(t11); free(t12); free(t13); free(t22); free(t23); free(t33);
  free(temp); free(mu); free(mudiv); free(plam); free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Adjoint (continuous Euler adjoint) RHS, ns_rho_rhou_p.f90:257-536.  The  */
/* grid-transform module's adjoint (ns_rho_rhou_p_grid_transform.f90:       */
/* 334-613) is the same body, without X_metric or X_jacobian, so a deformed */
/* grid uses this one as well.                                              */
void orc_rhs_adjoint(const orc_cfg *c, double *rhs, const double *qs_in, const double *q,
                     const double *g, const double *shape) {
  const size_t N = npts(c);
  const double gam = c->gamma, gm1 = c->gamma - 1.0;
  double *qs = xalloc(5 * N);
  double *u = xalloc(3 * N), *gdp = xalloc(N), *irho = xalloc(N);
  double *a = xalloc(N), *b = xalloc(N), *cc = xalloc(N), *d = xalloc(N), *e = xalloc(N);
  double *tmp = xalloc(N);
  memcpy(qs, qs_in, 5 * N * sizeof(double));
  orc_bc_qs(c, qs, q, shape); /* :290 */
  const double *q1 = q, *q5 = q + 4 * N;
  double *r[5];
  for (int v = 0; v < 5; ++v) r[v] = rhs + v * N;
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) {
    gdp[i] = gam * q5[i] / (q1[i] * gm1);
    irho[i] = 1.0 / q1[i];
  }
  memset(rhs, 0, 5 * N * sizeof(double));
  for (int m = 0; m < 3; ++m)
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) u[m * N + i] = q[(m + 1) * N + i] / q1[i];
  const double *u1 = u, *u2 = u + N, *u3 = u + 2 * N;
  const double *qs5 = qs + 4 * N;

  for (int dd = 0; dd < 3; ++dd) {
    const double h = c->dx[dd];
    orc_diff3d(c, dd, a, qs, h);
    orc_diff3d(c, dd, b, qs + N, h);
    orc_diff3d(c, dd, cc, qs + 2 * N, h);
    orc_diff3d(c, dd, d, qs + 3 * N, h);
    orc_diff3d(c, dd, e, qs + 4 * N, h);
    const double *ud = u + dd * N;
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      if (dd == 0) { /* :317-339 */
        r[0][i] = r[0][i] + (u1[i] * u1[i]) * b[i] + (u1[i] * u2[i]) * cc[i] +
                  (u1[i] * u3[i]) * d[i] + gdp[i] * u1[i] * e[i];
        r[1][i] = r[1][i] - a[i] - 2.0 * u1[i] * b[i] - u2[i] * cc[i] - u3[i] * d[i] - gdp[i] * e[i];
        r[2][i] = r[2][i] - u1[i] * cc[i];
        r[3][i] = r[3][i] - u1[i] * d[i];
        r[4][i] = r[4][i] - gm1 * b[i] - gam * u1[i] * e[i];
      } else if (dd == 1) { /* :407-428 */
        r[0][i] = r[0][i] + (u2[i] * u1[i]) * b[i] + (u2[i] * u2[i]) * cc[i] +
                  (u2[i] * u3[i]) * d[i] + gdp[i] * u2[i] * e[i];
        r[1][i] = r[1][i] - u2[i] * b[i];
        r[2][i] = r[2][i] - a[i] - u1[i] * b[i] - 2 * u2[i] * cc[i] - u3[i] * d[i] - gdp[i] * e[i];
        r[3][i] = r[3][i] - u2[i] * d[i];
        r[4][i] = r[4][i] - gm1 * cc[i] - gam * u2[i] * e[i];
      } else { /* :475-497 */
        r[0][i] = r[0][i] + (u3[i] * u1[i]) * b[i] + (u3[i] * u2[i]) * cc[i] +
                  (u3[i] * u3[i]) * d[i] + gdp[i] * u3[i] * e[i];
        r[1][i] = r[1][i] - u3[i] * b[i];
        r[2][i] = r[2][i] - u3[i] * cc[i];
        r[3][i] = r[3][i] - a[i] - u1[i] * b[i] - u2[i] * cc[i] - 2 * u3[i] * d[i] - gdp[i] * e[i];
        r[4][i] = r[4][i] - gm1 * d[i] - gam * u3[i] * e[i];
      }
    }
    /* -(A^T)^-1 (C^T qs)_x : :355-357, :443-445, :512-514 */
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) tmp[i] = -ud[i] * qs5[i];
    orc_diff3d(c, dd, a, tmp, h);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) r[4][i] = r[4][i] - gm1 * a[i];
    /* +(A^T)^-1 C~ c_x : :366-372 etc. */
    orc_diff3d(c, dd, a, q5, h);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) {
      r[0][i] = r[0][i] + qs5[i] * ud[i] * irho[i] * a[i];
      r[dd + 1][i] = r[dd + 1][i] - qs5[i] * irho[i] * a[i];
    }
  }
  if (g)
    for (size_t i = 0; i < 5 * N; ++i) rhs[i] = rhs[i] + g[i]; /* :512 */
  /* sponge_adjoint is the empty in-rhs sponge; rhss BCs without q */
  bc_rhss(c, rhs, qs, c->equations == ORC_EQ_EULER ? q : NULL, shape);
  free(qs); free(u); free(gdp); free(irho); free(a); free(b); free(cc); free(d); free(e);
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* RK4, src/time/rk_o4_e.f90:17-58.                                          */
void orc_time_stepper_direct(const orc_cfg *c, double *q1, const double *q0,
                             const double *qref, const double *shape, int heat_unit_spacing) {
  const size_t M = 5 * npts(c);
  const double dt = c->dt, dth = c->dt * (0.5);
  double *k1 = xalloc(M), *k2 = xalloc(M), *k3 = xalloc(M), *k4 = xalloc(M), *s = xalloc(M);
  orc_rhs_direct(c, k1, q0, qref, shape, heat_unit_spacing);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) s[i] = q0[i] + dth * k1[i];
  orc_rhs_direct(c, k2, s, qref, shape, heat_unit_spacing);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) s[i] = q0[i] + dth * k2[i];
  orc_rhs_direct(c, k3, s, qref, shape, heat_unit_spacing);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) s[i] = q0[i] + dt * k3[i];
  orc_rhs_direct(c, k4, s, qref, shape, heat_unit_spacing);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) q1[i] = q0[i] + dt / 6.0 * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
  free(k1); free(k2); free(k3); free(k4); free(s);
}

/* src/time/rk_o4_e.f90:62-105 (dt -> -dt; q0 and g frozen over the stages). */
void orc_time_stepper_adjoint(const orc_cfg *c, double *qs1, const double *qs0,
                              const double *q0, const double *g, const double *shape) {
  const size_t M = 5 * npts(c);
  const double dt = -c->dt, dth = -c->dt * (0.5);
  double *k1 = xalloc(M), *k2 = xalloc(M), *k3 = xalloc(M), *k4 = xalloc(M), *s = xalloc(M);
  orc_rhs_adjoint(c, k1, qs0, q0, g, shape);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) s[i] = qs0[i] + dth * k1[i];
  orc_rhs_adjoint(c, k2, s, q0, g, shape);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) s[i] = qs0[i] + dth * k2[i];
  orc_rhs_adjoint(c, k3, s, q0, g, shape);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) s[i] = qs0[i] + dt * k3[i];
  orc_rhs_adjoint(c, k4, s, q0, g, shape);
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < M; ++i) qs1[i] = qs0[i] + dt / 6.0 * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
  free(k1); free(k2); free(k3); free(k4); free(s);
}

/* Weighted filter, src/filter/filter_weighted_i1_mpi_cartesian.f90:15-54 with
 * filter_weighted_pen_core_o2_e_c_p.f90 filter_main (:75-108, periodic wrap of
 * rows 1 and n) and init_stencil (:112-128, a0 = -2/4, a1 = 1/4):
 * q_tilde = q; f = F(q); f = strength * weight * f; f = F(f); f = f / weight;
 * q = q_tilde - f. */
void orc_filter_weighted4d(const orc_cfg *c, int dir, double *q, int nv, const double *weight,
                           const double *strength) {
  const size_t N = npts(c);
  const int n1 = c->n[0], n2 = c->n[1], n = c->n[dir];
  const size_t st = dir == 0 ? 1 : (dir == 1 ? (size_t)n1 : (size_t)n1 * n2);
  const double a0 = -2.0 * 0.25, a1 = 1.0 * 0.25;
  double *f = xalloc(N), *g = xalloc(N);
  size_t *ip = (size_t *)malloc(N * sizeof(size_t)), *im = (size_t *)malloc(N * sizeof(size_t));
  #pragma omp parallel for schedule(static)
  for (size_t p = 0; p < N; ++p) {
    const int i = dir == 0 ? (int)(p % n1) : (dir == 1 ? (int)((p / n1) % n2) : (int)(p / ((size_t)n1 * n2)));
    ip[p] = i + 1 < n ? p + st : p - (size_t)(n - 1) * st;
    im[p] = i > 0 ? p - st : p + (size_t)(n - 1) * st;
  }
  for (int v = 0; v < nv; ++v) {
    double *x = q + (size_t)v * N;
    #pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) f[p] = a0 * x[p] + a1 * (x[ip[p]] + x[im[p]]);
    #pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p)
      f[p] = (strength ? strength[p] : 1.0) * (weight ? weight[p] : 1.0) * f[p];
    #pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) g[p] = a0 * f[p] + a1 * (f[ip[p]] + f[im[p]]);
    #pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) g[p] = g[p] / (weight ? weight[p] : 1.0);
    #pragma omp parallel for schedule(static)
    for (size_t p = 0; p < N; ++p) x[p] = x[p] - g[p];
  }
  free(f); free(g); free(ip); free(im);
}

/* Step-loop body, src/main/computation_direct.f90:67-98 / cruna_hpc.f90:171-242:
 * stepper -> qn = q1 -> BC on q (outside rhs) -> filters i1, i2, i3 at their
 * frequencies -> weighted filters i1, i2, i3 at theirs (cruna_hpc.f90:231-242;
 * the subset loop of computation_direct has none: leave wfilter_freq 0). */
void orc_run_direct(const orc_cfg *c, double *q, int nt0, int nsteps, const double *qref,
                    const double *shape, int heat_unit_spacing) {
  const size_t M = 5 * npts(c);
  double *q1 = xalloc(M);
  for (int nt = nt0 + 1; nt <= nt0 + nsteps; ++nt) {
    orc_time_stepper_direct(c, q1, q, qref, shape, heat_unit_spacing);
    memcpy(q, q1, M * sizeof(double));
    bc_q_at(c, q, qref, shape, 1);
    for (int d = 0; d < 3; ++d)
      if (c->filter_freq[d] > 0 && nt % c->filter_freq[d] == 0) orc_filter4d(c, d, q, 5);
    for (int d = 0; d < 3; ++d)
      if (c->wfilter_freq[d] > 0 && nt % c->wfilter_freq[d] == 0)
        orc_filter_weighted4d(c, d, q, 5, c->penal_phi, NULL);
  }
  free(q1);
}

/* The cruna_hpc loop with its convergence check, src/main/cruna_hpc.f90:171-242
 * (run.cfreq, run.ctol read at :81-82): after the stepper, at steps with
 * mod(nt, cfreq) == 0, residual = maxval(abs((qn - q1)/dt)) (:185-186); when
 * residual <= ctol the loop stops with q1 (the unfiltered new state the
 * reference stores as its snapshot, :191) and *nt_conv = nt, else qn = q1 and
 * the BC and filters follow as in orc_run_direct.  *res: the last residual
 * computed (-1 if none), *nt_conv: 0 if the tolerance was not reached. */
void orc_run_direct_conv(const orc_cfg *c, double *q, int nt0, int nsteps, int cfreq, double ctol,
                         const double *qref, const double *shape, int heat_unit_spacing,
                         double *res, int *nt_conv) {
  const size_t M = 5 * npts(c);
  double *q1 = xalloc(M);
  *res = -1.0;
  *nt_conv = 0;
  for (int nt = nt0 + 1; nt <= nt0 + nsteps; ++nt) {
    orc_time_stepper_direct(c, q1, q, qref, shape, heat_unit_spacing);
    if (cfreq > 0 && nt % cfreq == 0) {
      double m = 0.0;
      for (size_t i = 0; i < M; ++i) {
        const double d = fabs((q[i] - q1[i]) / c->dt);
        if (d > m) m = d;
      }
      *res = m;
      if (m <= ctol) {
        memcpy(q, q1, M * sizeof(double));
        *nt_conv = nt;
        break;
      }
    }
    memcpy(q, q1, M * sizeof(double));
    bc_q_at(c, q, qref, shape, 1);
    for (int d = 0; d < 3; ++d)
      if (c->filter_freq[d] > 0 && nt % c->filter_freq[d] == 0) orc_filter4d(c, d, q, 5);
    for (int d = 0; d < 3; ++d)
      if (c->wfilter_freq[d] > 0 && nt % c->wfilter_freq[d] == 0)
        orc_filter_weighted4d(c, d, q, 5, c->penal_phi, NULL);
  }
  free(q1);
}

/* src/main/computation_adjoint.f90:62-99 with g = 0 (calc_objective_g stub). */
void orc_run_adjoint(const orc_cfg *c, double *qs, const double *Q, int nsteps,
                     const double *shape) {
  const size_t M = 5 * npts(c);
  double *qs1 = xalloc(M);
  for (int nt = nsteps; nt >= 1; --nt) {
    const double *q0 = Q + (size_t)(nt - 1) * M;
    orc_time_stepper_adjoint(c, qs1, qs, q0, NULL, shape);
    memcpy(qs, qs1, M * sizeof(double));
    bc_qs_at(c, qs, q0, shape, 1);
    for (int d = 0; d < 3; ++d)
      if (c->afilter_freq[d] > 0 && nt % c->afilter_freq[d] == 0) orc_filter4d(c, d, qs, 5);
  }
  free(qs1);
}
