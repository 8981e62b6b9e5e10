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
