/*
 * dlu_oracle.c -- CPU oracle for the delayed-pivoting LU.  TEST INFRASTRUCTURE ONLY: it is the
 * checker for the HIP product path and the timed CPU baseline, never part of the product.
 *
 * Restated from the reference (paths relative to the reference repository):
 *   delayed_lu/delayed_dgetrf.cpp:129-251  (driver loop, panel restart, deferral)
 *   delayed_lu/delayed_dgetrf.cpp:32-47    (full-pivot search: idamax per column, strict '>')
 *   delayed_lu/delayed_dgetrf.cpp:61-85    (finish_panel: dtrsm + dgemm)
 *   delayed_lu/delayed_dgetrf.cpp:101-125  (materialize the accepted panel prefix)
 *   delayed_lu/delayed_dgetrf_column.cpp:30-45, 99-211 (the column-delay variant's fallback)
 *   delayed_lu/lapack_dgetrf.cpp:21-63     (0-based LAPACK wrapper, info mapping)
 *   delayed_lu/example.cpp:43-126          (reconstruction of P*A*Q from the compact output)
 * The BLAS call sequence is kept the same as the reference's because the oracle's job is to
 * reproduce its pivot decisions; the code itself is written fresh.
 */
#define _GNU_SOURCE
#include "dlu_oracle.h"

#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* CBLAS enumerator values (netlib cblas.h numbering, shared by MKL and OpenBLAS). */
enum {
  ORC_COL_MAJOR = 102,
  ORC_NO_TRANS = 111,
  ORC_TRANS = 112,
  ORC_LOWER = 122,
  ORC_UNIT = 132,
  ORC_LEFT = 141
};

typedef size_t (*fn_idamax)(int, const double *, int);
typedef void (*fn_dswap)(int, double *, int, double *, int);
typedef void (*fn_dscal)(int, double, double *, int);
typedef void (*fn_dger)(int, int, int, double, const double *, int, const double *, int,
                        double *, int);
typedef void (*fn_dtrsm)(int, int, int, int, int, int, int, double, const double *, int,
                         double *, int);
typedef void (*fn_dgemm)(int, int, int, int, int, int, double, const double *, int,
                         const double *, int, double, double *, int);
typedef void (*fn_dgetrf)(const int *, const int *, double *, const int *, int *, int *);
typedef void (*fn_sgetrf)(const int *, const int *, float *, const int *, int *, int *);
typedef void (*fn_set_threads)(int);
typedef void (*fn_dgemv)(int, int, int, int, double, const double *, int, const double *, int,
                         double, double *, int);
typedef void (*fn_dsyr2k)(int, int, int, int, int, double, const double *, int, const double *,
                          int, double, double *, int);
typedef void (*fn_dsyr)(int, int, int, double, const double *, int, double *, int);
typedef void (*fn_dsyr2)(int, int, int, double, const double *, int, const double *, int,
                         double *, int);

static struct {
  void *handle;
  fn_idamax idamax;
  fn_dswap dswap;
  fn_dscal dscal;
  fn_dger dger;
  fn_dtrsm dtrsm;
  fn_dgemm dgemm;
  fn_dgetrf dgetrf;
  fn_sgetrf sgetrf;
  fn_set_threads set_threads;
  fn_dgemv dgemv;   /* the symmetric factorization's routines: optional, checked on use */
  fn_dsyr2k dsyr2k;
  fn_dsyr dsyr;
  fn_dsyr2 dsyr2;
  char name[256];
} blas;

static void *sym(const char *prefix, const char *base) {
  char full[128];
  snprintf(full, sizeof full, "%s%s", prefix, base);
  return dlsym(blas.handle, full);
}

int dlu_oracle_init(const char *path, const char *prefix) {
  if (blas.handle != NULL) return 0;
  void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (h == NULL) return 1;
  blas.handle = h;
  blas.idamax = (fn_idamax)sym(prefix, "cblas_idamax");
  blas.dswap = (fn_dswap)sym(prefix, "cblas_dswap");
  blas.dscal = (fn_dscal)sym(prefix, "cblas_dscal");
  blas.dger = (fn_dger)sym(prefix, "cblas_dger");
  blas.dtrsm = (fn_dtrsm)sym(prefix, "cblas_dtrsm");
  blas.dgemm = (fn_dgemm)sym(prefix, "cblas_dgemm");
  blas.dgetrf = (fn_dgetrf)sym(prefix, "dgetrf_");
  blas.sgetrf = (fn_sgetrf)sym(prefix, "sgetrf_"); /* optional: FP32 partial-pivoting pin */
  blas.dgemv = (fn_dgemv)sym(prefix, "cblas_dgemv");
  blas.dsyr2k = (fn_dsyr2k)sym(prefix, "cblas_dsyr2k");
  blas.dsyr = (fn_dsyr)sym(prefix, "cblas_dsyr");
  blas.dsyr2 = (fn_dsyr2)sym(prefix, "cblas_dsyr2");
  if (!blas.idamax || !blas.dswap || !blas.dscal || !blas.dger || !blas.dtrsm || !blas.dgemm ||
      !blas.dgetrf) {
    dlclose(h);
    memset(&blas, 0, sizeof blas);
    return 2;
  }
  /* Library identification and thread control, whichever flavour this is. */
  void (*mkl_ver)(char *, int) = (void (*)(char *, int))dlsym(h, "mkl_get_version_string");
  char *(*ob_cfg)(void) = (char *(*)(void))sym(prefix, "openblas_get_config");
  if (mkl_ver != NULL) {
    mkl_ver(blas.name, (int)sizeof blas.name - 1);
    blas.set_threads = (fn_set_threads)dlsym(h, "MKL_Set_Num_Threads");
  } else if (ob_cfg != NULL) {
    snprintf(blas.name, sizeof blas.name, "OpenBLAS %s", ob_cfg());
    blas.set_threads = (fn_set_threads)sym(prefix, "openblas_set_num_threads");
  } else {
    snprintf(blas.name, sizeof blas.name, "unknown BLAS at %s", path);
  }
  for (char *p = blas.name; *p; ++p)
    if (*p == '\n') *p = ' ';
  return 0;
}

const char *dlu_oracle_blas_name(void) { return blas.name; }

int dlu_oracle_set_threads(int threads) {
  if (blas.set_threads != NULL && threads > 0) blas.set_threads(threads);
  return threads;
}

#define AT(a, lda, i, j) ((a)[(size_t)(i) + (size_t)(j) * (size_t)(lda)])

/* Largest |A(i,j)| over the active block A(k:m-1, k:n-1): idamax down each column, then a
 * strict '>' across columns, so ties keep the lowest column and, inside it, the lowest row
 * (delayed_dgetrf.cpp:32-47). */
static double active_block_max(int m, int n, const double *a, int lda, int k, int *row,
                               int *col) {
  double best = 0.0;
  *row = k;
  *col = k;
  for (int j = k; j < n; ++j) {
    const double *x = &AT(a, lda, k, j);
    const int r = k + (int)blas.idamax(m - k, x, 1);
    const double v = fabs(AT(a, lda, r, j));
    if (v > best) {
      best = v;
      *row = r;
      *col = j;
    }
  }
  return best;
}

/* Column-variant fallback (delayed_lu/delayed_dgetrf_column.cpp:30-45): walk the columns from
 * the right end down to k+1 and stop at the first whose idamax over rows [k, m) is nonzero. */
static double rightmost_nonzero_column(int m, int n, const double *a, int lda, int k, int *row,
                                       int *col) {
  *row = k;
  *col = k;
  for (int j = n - 1; j > k; --j) {
    const int r = k + (int)blas.idamax(m - k, &AT(a, lda, k, j), 1);
    const double v = fabs(AT(a, lda, r, j));
    if (v > 0.0) {
      *row = r;
      *col = j;
      return v;
    }
  }
  return 0.0;
}

/* Column pivot of step k: idamax over A(k:m-1, k).  Built with DLU_ORACLE_MUTANT_TIES (the
 * tests' planted mistake only), ties go to the highest row instead. */
static int column_pivot_row(int m, const double *colk) {
  int r = (int)blas.idamax(m, colk, 1);
#ifdef DLU_ORACLE_MUTANT_TIES
  for (int i = r + 1; i < m; ++i)
    if (fabs(colk[i]) == fabs(colk[r])) r = i;
#endif
  return r;
}

/* U12 = L11^{-1} A12 and A22 -= L21 U12 for the panel [begin, end) (delayed_dgetrf.cpp:61-85). */
static void close_panel(int m, int n, double *a, int lda, int begin, int end) {
  const int w = end - begin, right = n - end, below = m - end;
  if (w <= 0 || right <= 0) return;
  blas.dtrsm(ORC_COL_MAJOR, ORC_LEFT, ORC_LOWER, ORC_NO_TRANS, ORC_UNIT, w, right, 1.0,
             &AT(a, lda, begin, begin), lda, &AT(a, lda, begin, end), lda);
  if (below > 0)
    blas.dgemm(ORC_COL_MAJOR, ORC_NO_TRANS, ORC_NO_TRANS, below, right, w, -1.0,
               &AT(a, lda, end, begin), lda, &AT(a, lda, begin, end), lda, 1.0,
               &AT(a, lda, end, end), lda);
}

/* Bring the far columns [panel_end, n) up to date with the accepted prefix [begin, k) before a
 * full-pivot search (delayed_dgetrf.cpp:101-125). */
static void apply_prefix_to_far_columns(int m, int n, double *a, int lda, int begin, int k,
                                        int panel_end) {
  const int w = k - begin, far = n - panel_end, active = m - k;
  if (w <= 0 || far <= 0) return;
  blas.dtrsm(ORC_COL_MAJOR, ORC_LEFT, ORC_LOWER, ORC_NO_TRANS, ORC_UNIT, w, far, 1.0,
             &AT(a, lda, begin, begin), lda, &AT(a, lda, begin, panel_end), lda);
  if (active > 0)
    blas.dgemm(ORC_COL_MAJOR, ORC_NO_TRANS, ORC_NO_TRANS, active, far, w, -1.0,
               &AT(a, lda, k, begin), lda, &AT(a, lda, begin, panel_end), lda, 1.0,
               &AT(a, lda, k, panel_end), lda);
}

/* Instrumentation: relative gap between the two largest |x| of a column segment. */
static double top2_margin(const double *x, int len) {
  double t1 = 0.0, t2 = 0.0;
  for (int i = 0; i < len; ++i) {
    const double v = fabs(x[i]);
    if (v > t1) {
      t2 = t1;
      t1 = v;
    } else if (v > t2) {
      t2 = v;
    }
  }
  return t1 > 0.0 ? (t1 - t2) / t1 : 0.0;
}

/* The blocked driver of both reference variants; `column` selects the fallback search
 * (delayed_dgetrf_column.cpp:99-211 differs from delayed_dgetrf.cpp:129-251 only there). */
static void delayed_lu(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info, int nb,
                       double *margins, int *fallbacks, int column) {
  /* Argument rules of delayed_dgetrf.cpp:132-151. */
  *info = -1;
  if (fallbacks) *fallbacks = 0;
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return;
  const int steps = m < n ? m : n;
  if ((m > 0 && n > 0 && a == NULL) || (steps > 0 && (ipiv == NULL || jpiv == NULL))) return;
  if (nb <= 0) nb = 64;
  *info = steps;
  for (int k = 0; k < steps; ++k) {
    ipiv[k] = k;
    jpiv[k] = k;
    if (margins) margins[k] = -1.0;
  }

  int panel_begin = 0;
  int pivot_in_place = 0; /* the restarted panel's first pivot was already chosen and moved */
  while (panel_begin < steps) {
    const int panel_end = panel_begin + nb < steps ? panel_begin + nb : steps;
    int restarted = 0;
    for (int k = panel_begin; k < panel_end && !restarted; ++k) {
      if (pivot_in_place) {
        pivot_in_place = 0;
      } else {
        /* Column pivot first: idamax over A(k:m-1, k). */
        const double *colk = &AT(a, lda, k, k);
        int prow = k + column_pivot_row(m - k, colk);
        int pcol = k;
        const double alpha = fabs(AT(a, lda, prow, k));
        const int column_is_zero = !(alpha > 0.0); /* NaN reads as unacceptable too */
        if (!column_is_zero && margins) margins[k] = top2_margin(colk, m - k);
        if (column_is_zero) {
          if (fallbacks) ++*fallbacks;
          if (k > panel_begin) apply_prefix_to_far_columns(m, n, a, lda, panel_begin, k, panel_end);
          const double best = column ? rightmost_nonzero_column(m, n, a, lda, k, &prow, &pcol)
                                     : active_block_max(m, n, a, lda, k, &prow, &pcol);
          if (!(best > 0.0)) {
            *info = k; /* the whole active block is zero: [k, steps) is the delayed suffix */
            return;
          }
        }
        ipiv[k] = prow;
        jpiv[k] = pcol;
        if (prow != k) blas.dswap(n, &AT(a, lda, k, 0), lda, &AT(a, lda, prow, 0), lda);
        if (pcol != k) blas.dswap(m, &AT(a, lda, 0, k), 1, &AT(a, lda, 0, pcol), 1);
        if (column_is_zero && k > panel_begin) {
          /* [panel_begin, k) is now a finished short panel; reopen at k with the pivot set. */
          panel_begin = k;
          pivot_in_place = 1;
          restarted = 1;
          break;
        }
      }
      const int below = m - k - 1;
      if (below > 0) {
        const double recip = 1.0 / AT(a, lda, k, k); /* multiply by the reciprocal */
        blas.dscal(below, recip, &AT(a, lda, k + 1, k), 1);
      }
      const int in_panel_right = panel_end - k - 1;
      if (below > 0 && in_panel_right > 0)
        blas.dger(ORC_COL_MAJOR, below, in_panel_right, -1.0, &AT(a, lda, k + 1, k), 1,
                  &AT(a, lda, k, k + 1), lda, &AT(a, lda, k + 1, k + 1), lda);
    }
    if (restarted) continue;
    close_panel(m, n, a, lda, panel_begin, panel_end);
    panel_begin = panel_end;
  }
}

void dlu_oracle_delayed_getrf(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                              int *info, int nb, double *margins, int *fallbacks) {
  delayed_lu(m, n, a, lda, ipiv, jpiv, info, nb, margins, fallbacks, 0);
}

void dlu_oracle_delayed_getrf_column(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                     int *info, int nb, double *margins, int *fallbacks) {
  delayed_lu(m, n, a, lda, ipiv, jpiv, info, nb, margins, fallbacks, 1);
}

void dlu_oracle_lapack_getrf(int m, int n, double *a, int lda, int *ipiv, int *info) {
  *info = -1;
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return;
  const int steps = m < n ? m : n;
  if ((m > 0 && n > 0 && a == NULL) || (steps > 0 && ipiv == NULL)) return;
  int lapack_info = 0;
  int *one_based = (int *)malloc(sizeof(int) * (size_t)(steps > 0 ? steps : 1));
  if (one_based == NULL) return;
  blas.dgetrf(&m, &n, a, &lda, one_based, &lapack_info);
  for (int k = 0; k < steps; ++k) ipiv[k] = one_based[k] - 1;
  free(one_based);
  if (lapack_info == 0)
    *info = steps;
  else if (lapack_info > 0)
    *info = lapack_info - 1; /* first exactly-zero U diagonal, 0-based */
}

static void unblocked(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info,
                      int column);

int dlu_oracle_lapack_sgetrf(int m, int n, float *a, int lda, int *ipiv, int *info) {
  *info = -1;
  if (blas.sgetrf == NULL) return 1;
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return 0;
  const int steps = m < n ? m : n;
  int lapack_info = 0;
  int *one_based = (int *)malloc(sizeof(int) * (size_t)(steps > 0 ? steps : 1));
  if (one_based == NULL) return 2;
  blas.sgetrf(&m, &n, a, &lda, one_based, &lapack_info);
  for (int k = 0; k < steps; ++k) ipiv[k] = one_based[k] - 1;
  free(one_based);
  *info = lapack_info == 0 ? steps : (lapack_info > 0 ? lapack_info - 1 : -1);
  return 0;
}

void dlu_oracle_unblocked_ld(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info) {
  unblocked(m, n, a, lda, ipiv, jpiv, info, 0);
}

void dlu_oracle_unblocked_column_ld(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                    int *info) {
  unblocked(m, n, a, lda, ipiv, jpiv, info, 1);
}

static void unblocked(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info,
                      int column) {
  *info = -1;
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return;
  const int steps = m < n ? m : n;
  if ((m > 0 && n > 0 && a == NULL) || (steps > 0 && (ipiv == NULL || jpiv == NULL))) return;
  *info = steps;
  for (int k = 0; k < steps; ++k) ipiv[k] = jpiv[k] = k;
  long double *w = (long double *)malloc(sizeof(long double) * (size_t)m * (size_t)(n > 0 ? n : 1));
  if (w == NULL) {
    *info = -1;
    return;
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) w[(size_t)i + (size_t)j * m] = AT(a, lda, i, j);
#define W(i, j) w[(size_t)(i) + (size_t)(j) * (size_t)m]
  for (int k = 0; k < steps; ++k) {
    int pr = k, pc = k;
    long double best = 0.0L;
    for (int i = k; i < m; ++i) {
      const long double v = fabsl(W(i, k));
      if (v > best) {
        best = v;
        pr = i;
      }
    }
    if (!(best > 0.0L) && column) {
      /* rightmost column with a nonzero entry, its first largest |a| */
      for (int j = n - 1; j > k && !(best > 0.0L); --j)
        for (int i = k; i < m; ++i) {
          const long double v = fabsl(W(i, j));
          if (v > best) {
            best = v;
            pr = i;
            pc = j;
          }
        }
    } else if (!(best > 0.0L)) {
      for (int j = k; j < n; ++j)
        for (int i = k; i < m; ++i) {
          const long double v = fabsl(W(i, j));
          if (v > best) {
            best = v;
            pr = i;
            pc = j;
          }
        }
    }
    if (!(best > 0.0L)) { /* the whole active block is zero */
      *info = k;
      break;
    }
    ipiv[k] = pr;
    jpiv[k] = pc;
    if (pr != k)
      for (int j = 0; j < n; ++j) {
        const long double t = W(k, j);
        W(k, j) = W(pr, j);
        W(pr, j) = t;
      }
    if (pc != k)
      for (int i = 0; i < m; ++i) {
        const long double t = W(i, k);
        W(i, k) = W(i, pc);
        W(i, pc) = t;
      }
    const long double recip = 1.0L / W(k, k);
    for (int i = k + 1; i < m; ++i) W(i, k) *= recip;
    for (int j = k + 1; j < n; ++j) {
      const long double u = W(k, j);
      if (u == 0.0L) continue;
      for (int i = k + 1; i < m; ++i) W(i, j) -= W(i, k) * u;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) AT(a, lda, i, j) = (double)W(i, j);
#undef W
  free(w);
}

double dlu_oracle_backward_error(int m, int n, const double *original, const double *factor,
                                 int lda, const int *ipiv, const int *jpiv, int eliminated,
                                 int delayed) {
  const int steps = m < n ? m : n;
  const int e = eliminated < 0 ? 0 : (eliminated > steps ? steps : eliminated);
  double *lo = (double *)calloc((size_t)m * (size_t)(e > 0 ? e : 1), sizeof(double));
  double *up = (double *)calloc((size_t)(e > 0 ? e : 1) * (size_t)n, sizeof(double));
  double *prod = (double *)calloc((size_t)m * (size_t)(n > 0 ? n : 1), sizeof(double));
  int *rows = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
  int *cols = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  double result = -1.0;
  if (!lo || !up || !prod || !rows || !cols) goto done;
  /* Lbar columns [0,e): unit lower; Ubar rows [0,e): upper trapezoid. */
  for (int k = 0; k < e; ++k) {
    lo[(size_t)k + (size_t)k * m] = 1.0;
    for (int i = k + 1; i < m; ++i) lo[(size_t)i + (size_t)k * m] = AT(factor, lda, i, k);
    for (int j = k; j < n; ++j) up[(size_t)k + (size_t)j * e] = AT(factor, lda, k, j);
  }
  if (e > 0)
    blas.dgemm(ORC_COL_MAJOR, ORC_NO_TRANS, ORC_NO_TRANS, m, n, e, 1.0, lo, m, up, e, 0.0, prod,
               m);
  if (delayed) /* identity block of Lbar times the Schur complement rows of Ubar */
    for (int j = e; j < n; ++j)
      for (int i = e; i < m; ++i) prod[(size_t)i + (size_t)j * m] += AT(factor, lda, i, j);
  for (int i = 0; i < m; ++i) rows[i] = i;
  for (int j = 0; j < n; ++j) cols[j] = j;
  for (int k = 0; k < e; ++k) {
    int t = rows[k];
    rows[k] = rows[ipiv[k]];
    rows[ipiv[k]] = t;
    if (jpiv) {
      t = cols[k];
      cols[k] = cols[jpiv[k]];
      cols[jpiv[k]] = t;
    }
  }
  {
    long double err2 = 0.0L, nrm2 = 0.0L;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) {
        const double ref = AT(original, lda, rows[i], cols[j]);
        const long double d = (long double)prod[(size_t)i + (size_t)j * m] - ref;
        err2 += d * d;
        nrm2 += (long double)ref * ref;
      }
    const double big = (double)(m > n ? m : n);
    result = nrm2 > 0.0L ? (double)(sqrtl(err2) / (big * DBL_EPSILON * sqrtl(nrm2)))
                         : (err2 > 0.0L ? INFINITY : 0.0);
  }
done:
  free(lo);
  free(up);
  free(prod);
  free(rows);
  free(cols);
  return result;
}


/* ---- Delayed Bunch-Kaufman LDL^T (delayed-bunch-kaufman/) ----------------------------------
 * Restated from:
 *   delayed_sytrf.cpp:16-23          the Bunch-Kaufman constant (1 + sqrt 17) / 8 and the 2x2
 *                                    determinant tolerance 64 eps
 *   delayed_sytrf.cpp:28-67          lower-only symmetric read and symmetric interchange
 *   delayed_sytrf.cpp:71-106         active row maximum, stability test of a 2x2 block
 *   delayed_sytrf.cpp:108-173        1x1 / 2x2 elimination (dscal + dsyr; dsyr, dsyr2, dsyr)
 *   delayed_sytrf.cpp:177-282        the unblocked driver: pivot choice, deferral to the front of
 *                                    the delayed suffix
 *   delayed_sytrf_blocked.cpp:70-164 the lazy panel (S = A - L W^T, W = L D), dgemv
 *                                    materialization of a column or an active row, dsyr2k
 *                                    write-back over the whole trailing part
 *   delayed_sytrf_blocked.cpp:168-358 the blocked driver (panels of at most 64 columns that
 *                                    never split a 2x2 block)
 * Only the lower triangle is read or written.  perm[pos] is the original index of the node at
 * pos; piv_size 1 = 1x1 pivot, 2 = first column of a 2x2 pivot, 0 = its second column or a
 * delayed node; [0, info) eliminated, [info, n) delayed (their Schur complement in place). */

static double bk_alpha(void) { return (1.0 + sqrt(17.0)) / 8.0; }

static double bk_sym(const double *a, int lda, int r, int c) {
  return r >= c ? AT(a, lda, r, c) : AT(a, lda, c, r);
}

/* Symmetric interchange of nodes x and y in lower storage: row segments left of x, the part
 * between them (column x below x against row y), the diagonals, and the columns below y. */
static void bk_swap(int n, double *a, int lda, int x, int y, int *perm) {
  if (x == y) return;
  if (x > y) {
    const int t = x;
    x = y;
    y = t;
  }
  if (x > 0) blas.dswap(x, &AT(a, lda, x, 0), lda, &AT(a, lda, y, 0), lda);
  if (y - x - 1 > 0) blas.dswap(y - x - 1, &AT(a, lda, x + 1, x), 1, &AT(a, lda, y, x + 1), lda);
  {
    const double t = AT(a, lda, x, x);
    AT(a, lda, x, x) = AT(a, lda, y, y);
    AT(a, lda, y, y) = t;
  }
  if (n - y - 1 > 0) blas.dswap(n - y - 1, &AT(a, lda, y + 1, x), 1, &AT(a, lda, y + 1, y), 1);
  const int t = perm[x];
  perm[x] = perm[y];
  perm[y] = t;
}

/* D = [a b; b c] is accepted when its scaled determinant exceeds 64 eps times the square of its
 * scaled 1-norm (a zero or NaN block never is). */
static int bk_stable_2x2(double a, double b, double c) {
  double scale = fabs(a);
  if (fabs(b) > scale) scale = fabs(b);
  if (fabs(c) > scale) scale = fabs(c);
  if (!(scale > 0.0)) return 0;
  const double as = a / scale, bs = b / scale, cs = c / scale;
  const double det = as * cs - bs * bs;
  const double n1 = fabs(as) + fabs(bs), n2 = fabs(bs) + fabs(cs);
  const double nrm = n1 > n2 ? n1 : n2;
  return fabs(det) > 64.0 * DBL_EPSILON * nrm * nrm;
}

/* The Bunch-Kaufman choice at step k from the current column (col[k..active_end) of the Schur
 * complement) and, when it is needed, the candidate row's values over [k, active_end) (fetched
 * by row_fn).  Returns 0 defer, 1 for a 1x1 pivot (node *swap_to at k), 2 for a 2x2 pivot with
 * node *swap_to at k+1. */
typedef void (*bk_row_fn)(void *ctx, int row, double *row_values);
static int bk_choose(int k, int active_end, const double *col, int colmax_row, bk_row_fn row_fn,
                     void *ctx, double *row_values, int *swap_to) {
  const double alpha = bk_alpha();
  const double abs_akk = fabs(col[k]);
  const double colmax = active_end - k > 1 ? fabs(col[colmax_row]) : 0.0;
  *swap_to = k;
  if (!((abs_akk > colmax ? abs_akk : colmax) > 0.0)) return 0;
  if (abs_akk >= alpha * colmax) return 1;
  row_fn(ctx, colmax_row, row_values);
  double rowmax = 0.0;
  for (int c = k; c < active_end; ++c)
    if (c != colmax_row && fabs(row_values[c]) > rowmax) rowmax = fabs(row_values[c]);
  const double abs_app = fabs(row_values[colmax_row]);
  if (abs_akk >= alpha * colmax * (colmax / rowmax)) return 1;
  if (abs_app >= alpha * rowmax) {
#ifndef DLU_ORACLE_MUTANT_BK_NOSWAP
    *swap_to = colmax_row;
#endif /* built with DLU_ORACLE_MUTANT_BK_NOSWAP (tests only): the candidate's diagonal is taken
          without bringing the candidate to k, a mistake the long-double pin must see */
    return 1;
  }
  if (bk_stable_2x2(col[k], col[colmax_row], row_values[colmax_row])) {
    *swap_to = colmax_row;
    return 2;
  }
  return 0;
}

static int bk_args(int n, const double *a, int lda, const int *perm, const int *piv) {
  return !(n < 0 || lda < (n > 1 ? n : 1) || (n > 0 && (!a || !perm || !piv)));
}

/* The unblocked driver's row fetch: row p of the stored (always current) lower triangle. */
struct bk_plain_ctx {
  const double *a;
  int lda, k, active_end;
};
static void bk_plain_row(void *vctx, int row, double *out) {
  const struct bk_plain_ctx *c = (const struct bk_plain_ctx *)vctx;
  for (int j = c->k; j < c->active_end; ++j) out[j] = bk_sym(c->a, c->lda, row, j);
}

void dlu_oracle_delayed_sytrf(int n, double *a, int lda, int *perm, int *piv_size, int *info) {
  *info = -1;
  if (!bk_args(n, a, lda, perm, piv_size)) return;
  if (!blas.dsyr || !blas.dsyr2 || !blas.dscal || !blas.idamax || !blas.dswap) {
    *info = -2;
    return;
  }
  for (int i = 0; i < n; ++i) {
    perm[i] = i;
    piv_size[i] = 0;
  }
  double *col = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *row = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int k = 0, active_end = n;
  while (k < active_end) {
    for (int i = k; i < active_end; ++i) col[i] = AT(a, lda, i, k);
    const int p = active_end - k > 1 ? k + 1 + (int)blas.idamax(active_end - k - 1, &AT(a, lda, k + 1, k), 1) : k;
    struct bk_plain_ctx ctx = {a, lda, k, active_end};
    int to = k;
    const int kind = bk_choose(k, active_end, col, p, bk_plain_row, &ctx, row, &to);
    if (kind == 0) {
      bk_swap(n, a, lda, k, active_end - 1, perm);
      --active_end;
      continue;
    }
    if (kind == 1) {
      bk_swap(n, a, lda, k, to, perm);
      piv_size[k] = 1;
      const int t = n - k - 1;
      if (t > 0) {
        const double d = AT(a, lda, k, k);
        blas.dscal(t, 1.0 / d, &AT(a, lda, k + 1, k), 1);
        blas.dsyr(ORC_COL_MAJOR, ORC_LOWER, t, -d, &AT(a, lda, k + 1, k), 1,
                  &AT(a, lda, k + 1, k + 1), lda);
      }
      ++k;
      continue;
    }
    bk_swap(n, a, lda, k + 1, to, perm);
    piv_size[k] = 2;
    piv_size[k + 1] = 0;
    const int t = n - k - 2;
    if (t > 0) {
      const double d00 = AT(a, lda, k, k), d10 = AT(a, lda, k + 1, k);
      const double d11 = AT(a, lda, k + 1, k + 1);
      double sc = fabs(d00);
      if (fabs(d10) > sc) sc = fabs(d10);
      if (fabs(d11) > sc) sc = fabs(d11);
      const double s00 = d00 / sc, s10 = d10 / sc, s11 = d11 / sc;
      const double inv = 1.0 / (sc * (s00 * s11 - s10 * s10));
      double *l0 = &AT(a, lda, k + 2, k), *l1 = &AT(a, lda, k + 2, k + 1);
      for (int i = 0; i < t; ++i) {
        const double b0 = l0[i], b1 = l1[i];
        l0[i] = (s11 * b0 - s10 * b1) * inv;
        l1[i] = (s00 * b1 - s10 * b0) * inv;
      }
      double *c = &AT(a, lda, k + 2, k + 2);
      blas.dsyr(ORC_COL_MAJOR, ORC_LOWER, t, -d00, l0, 1, c, lda);
      blas.dsyr2(ORC_COL_MAJOR, ORC_LOWER, t, -d10, l0, 1, l1, 1, c, lda);
      blas.dsyr(ORC_COL_MAJOR, ORC_LOWER, t, -d11, l1, 1, c, lda);
    }
    k += 2;
  }
  *info = active_end;
  free(col);
  free(row);
}

/* The blocked driver's lazy Schur complement: column c over rows [c, n) and an active row, both
 * as the stored values minus the panel's L W^T (one dgemv each). */
struct bk_panel_ctx {
  const double *a;
  int lda, n, pb, pw, k, active_end;
  const double *w;
  int ldw;
};
static void bk_column(const struct bk_panel_ctx *c, int col, double *out) {
  for (int r = col; r < c->n; ++r) out[r] = AT(c->a, c->lda, r, col);
  if (c->pw > 0)
    blas.dgemv(ORC_COL_MAJOR, ORC_NO_TRANS, c->n - col, c->pw, -1.0, &AT(c->a, c->lda, col, c->pb),
               c->lda, c->w + col, c->ldw, 1.0, out + col, 1);
}
static void bk_panel_row(void *vctx, int row, double *out) {
  const struct bk_panel_ctx *c = (const struct bk_panel_ctx *)vctx;
  for (int j = c->k; j < c->active_end; ++j) out[j] = bk_sym(c->a, c->lda, row, j);
  if (c->pw > 0)
    blas.dgemv(ORC_COL_MAJOR, ORC_NO_TRANS, c->active_end - c->k, c->pw, -1.0,
               &AT(c->a, c->lda, c->k, c->pb), c->lda, c->w + row, c->ldw, 1.0, out + c->k, 1);
}
static void bk_swap_nodes(const struct bk_panel_ctx *c, double *a, int x, int y, int *perm,
                          double *w) {
  if (x == y) return;
  bk_swap(c->n, a, c->lda, x, y, perm);
  if (c->pw > 0) blas.dswap(c->pw, w + x, c->ldw, w + y, c->ldw);
}

void dlu_oracle_delayed_sytrf_blocked(int n, double *a, int lda, int *perm, int *piv_size,
                                      int *info, int nb) {
  *info = -1;
  if (!bk_args(n, a, lda, perm, piv_size) || nb < 2) return;
  if (!blas.dgemv || !blas.dsyr2k || !blas.idamax || !blas.dswap) {
    *info = -2;
    return;
  }
  for (int i = 0; i < n; ++i) {
    perm[i] = i;
    piv_size[i] = 0;
  }
  const int ldw = n > 1 ? n : 1;
  double *w = (double *)calloc((size_t)ldw * (size_t)nb, sizeof(double));
  double *c0 = (double *)calloc((size_t)ldw, sizeof(double));
  double *c1 = (double *)calloc((size_t)ldw, sizeof(double));
  double *rv = (double *)calloc((size_t)ldw, sizeof(double));
  int pb = 0, active_end = n;
  while (pb < active_end) {
    int k = pb, pw = 0;
    while (k < active_end && pw < nb) {
      struct bk_panel_ctx ctx = {a, lda, n, pb, pw, k, active_end, w, ldw};
      bk_column(&ctx, k, c0);
      const int p =
          active_end - k > 1 ? k + 1 + (int)blas.idamax(active_end - k - 1, c0 + k + 1, 1) : k;
      int to = k;
      const int kind = bk_choose(k, active_end, c0, p, bk_panel_row, &ctx, rv, &to);
      if (kind == 0) {
        bk_swap_nodes(&ctx, a, k, active_end - 1, perm, w);
        --active_end;
        continue;
      }
      if (kind == 2 && pw + 2 > nb) break; /* a 2x2 block never straddles two panels */
      if (kind == 1) {
        if (to != k) {
          bk_swap_nodes(&ctx, a, k, to, perm, w);
          bk_column(&ctx, k, c0);
        }
        const double d = c0[k];
        AT(a, lda, k, k) = d;
        piv_size[k] = 1;
        w[(size_t)k + (size_t)pw * ldw] = d;
        for (int r = k + 1; r < n; ++r) {
          AT(a, lda, r, k) = c0[r] / d;
          w[(size_t)r + (size_t)pw * ldw] = c0[r];
        }
        ++k;
        ++pw;
        continue;
      }
      bk_swap_nodes(&ctx, a, k + 1, to, perm, w);
      bk_column(&ctx, k, c0);
      bk_column(&ctx, k + 1, c1);
      const double d00 = c0[k], d10 = c0[k + 1], d11 = c1[k + 1];
      AT(a, lda, k, k) = d00;
      AT(a, lda, k + 1, k) = d10;
      AT(a, lda, k + 1, k + 1) = d11;
      piv_size[k] = 2;
      piv_size[k + 1] = 0;
      double sc = fabs(d00);
      if (fabs(d10) > sc) sc = fabs(d10);
      if (fabs(d11) > sc) sc = fabs(d11);
      const double s00 = d00 / sc, s10 = d10 / sc, s11 = d11 / sc;
      const double inv = 1.0 / (sc * (s00 * s11 - s10 * s10));
      w[(size_t)k + (size_t)pw * ldw] = d00;
      w[(size_t)(k + 1) + (size_t)pw * ldw] = d10;
      w[(size_t)k + (size_t)(pw + 1) * ldw] = d10;
      w[(size_t)(k + 1) + (size_t)(pw + 1) * ldw] = d11;
      for (int r = k + 2; r < n; ++r) {
        const double b0 = c0[r], b1 = c1[r];
        AT(a, lda, r, k) = (s11 * b0 - s10 * b1) * inv;
        AT(a, lda, r, k + 1) = (s00 * b1 - s10 * b0) * inv;
        w[(size_t)r + (size_t)pw * ldw] = b0;
        w[(size_t)r + (size_t)(pw + 1) * ldw] = b1;
      }
      k += 2;
      pw += 2;
    }
    /* Write the panel back: A22 -= 1/2 (L21 W21^T + W21 L21^T), lower, every trailing row. */
    if (k > pb && n - k > 0)
      blas.dsyr2k(ORC_COL_MAJOR, ORC_LOWER, ORC_NO_TRANS, n - k, k - pb, -0.5, &AT(a, lda, k, pb),
                  lda, w + k, ldw, 1.0, &AT(a, lda, k, k), lda);
    pb = k;
  }
  *info = active_end;
  free(w);
  free(c0);
  free(c1);
  free(rv);
}

/* An independent restatement of the unblocked driver in long double with no BLAS (the pin of
 * the two above): the same rules, its own argmax (first index of the largest magnitude, strict
 * '>') and its own updates. */
void dlu_oracle_sytrf_unblocked_ld(int n, double *a, int lda, int *perm, int *piv_size,
                                   int *info) {
  *info = -1;
  if (!bk_args(n, a, lda, perm, piv_size)) return;
  long double *m = (long double *)malloc(sizeof(long double) * (size_t)(n > 0 ? n : 1) *
                                         (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) m[(size_t)i + (size_t)j * n] = bk_sym(a, lda, i, j);
#define M(i, j) m[(size_t)(i) + (size_t)(j) * (size_t)n]
  for (int i = 0; i < n; ++i) {
    perm[i] = i;
    piv_size[i] = 0;
  }
  const long double alpha = (1.0L + sqrtl(17.0L)) / 8.0L;
  int k = 0, active_end = n;
  while (k < active_end) {
    int p = k;
    long double colmax = 0.0L;
    for (int i = k + 1; i < active_end; ++i)
      if (fabsl(M(i, k)) > colmax || p == k) {
        if (p == k || fabsl(M(i, k)) > colmax) {
          p = i;
          colmax = fabsl(M(i, k));
        }
      }
    const long double akk = fabsl(M(k, k));
    int kind = 0, to = k;
    if (!((akk > colmax ? akk : colmax) > 0.0L)) {
      kind = 0;
    } else if (akk >= alpha * colmax) {
      kind = 1;
    } else {
      long double rowmax = 0.0L;
      for (int j = k; j < active_end; ++j)
        if (j != p && fabsl(M(p, j)) > rowmax) rowmax = fabsl(M(p, j));
      const long double app = fabsl(M(p, p));
      if (akk >= alpha * colmax * (colmax / rowmax)) {
        kind = 1;
      } else if (app >= alpha * rowmax) {
        kind = 1;
        to = p;
      } else if (bk_stable_2x2((double)M(k, k), (double)M(p, k), (double)M(p, p))) {
        kind = 2;
        to = p;
      }
    }
    /* full symmetric interchange of nodes x and y in the dense long-double copy */
    int sx = -1, sy = -1;
    if (kind == 0) {
      sx = k;
      sy = active_end - 1;
    } else if (kind == 1) {
      sx = k;
      sy = to;
    } else {
      sx = k + 1;
      sy = to;
    }
    if (sx != sy) {
      for (int j = 0; j < n; ++j) {
        const long double t = M(sx, j);
        M(sx, j) = M(sy, j);
        M(sy, j) = t;
      }
      for (int i = 0; i < n; ++i) {
        const long double t = M(i, sx);
        M(i, sx) = M(i, sy);
        M(i, sy) = t;
      }
      const int t = perm[sx];
      perm[sx] = perm[sy];
      perm[sy] = t;
    }
    if (kind == 0) {
      --active_end;
      continue;
    }
    if (kind == 1) {
      piv_size[k] = 1;
      const long double d = M(k, k);
      for (int i = k + 1; i < n; ++i) M(i, k) /= d;
      for (int j = k + 1; j < n; ++j)
        for (int i = j; i < n; ++i) M(i, j) -= d * M(i, k) * M(j, k);
      for (int j = k + 1; j < n; ++j)
        for (int i = k + 1; i < j; ++i) M(i, j) = M(j, i);
      ++k;
      continue;
    }
    piv_size[k] = 2;
    piv_size[k + 1] = 0;
    const long double d00 = M(k, k), d10 = M(k + 1, k), d11 = M(k + 1, k + 1);
    const long double det = d00 * d11 - d10 * d10;
    for (int i = k + 2; i < n; ++i) {
      const long double b0 = M(i, k), b1 = M(i, k + 1);
      M(i, k) = (d11 * b0 - d10 * b1) / det;
      M(i, k + 1) = (d00 * b1 - d10 * b0) / det;
    }
    for (int j = k + 2; j < n; ++j)
      for (int i = j; i < n; ++i)
        M(i, j) -= M(i, k) * (d00 * M(j, k) + d10 * M(j, k + 1)) +
                   M(i, k + 1) * (d10 * M(j, k) + d11 * M(j, k + 1));
    for (int j = k + 2; j < n; ++j)
      for (int i = k + 2; i < j; ++i) M(i, j) = M(j, i);
    k += 2;
  }
  /* back to the caller's lower triangle */
  for (int j = 0; j < n; ++j)
    for (int i = j; i < n; ++i) AT(a, lda, i, j) = (double)M(i, j);
#undef M
  *info = active_end;
  free(m);
}
