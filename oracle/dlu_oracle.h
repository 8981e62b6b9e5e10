/*
 * dlu_oracle.h -- CPU oracle for the delayed-pivoting LU (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker, never the product: only tests/, the smoke entry and the
 * CPU-baseline leg of the benchmark load it.  The product path (libdlu.so) never links,
 * loads or calls anything here.
 *
 * What it restates (reference file:line, paths relative to the reference repository):
 *   dlu_oracle_delayed_getrf   delayed_lu/delayed_dgetrf.cpp:129-251 (blocked, nb=64 by default)
 *     - full-pivot search      delayed_lu/delayed_dgetrf.cpp:32-47
 *     - finish_panel           delayed_lu/delayed_dgetrf.cpp:61-85
 *     - materialize prefix     delayed_lu/delayed_dgetrf.cpp:101-125
 *   dlu_oracle_delayed_getrf_column
 *                              delayed_lu/delayed_dgetrf_column.cpp:99-211 (the column-delay
 *                              variant; fallback find_nonzero_column, :30-45)
 *   dlu_oracle_lapack_getrf    delayed_lu/lapack_dgetrf.cpp:21-63
 *   dlu_oracle_backward_error  delayed_lu/example.cpp:43-126 (reconstruction of P*A*Q from the
 *                              compact [L | U / Schur] storage), reported in Frobenius norm.
 *   dlu_oracle_unblocked_ld    an independent unblocked restatement in long double with no BLAS,
 *                              used to pin the blocked restatement's pivot sequence.
 *
 * All BLAS/LAPACK arithmetic goes through a library loaded at run time (MKL libmkl_rt, or the
 * OpenBLAS bundled with scipy, whose symbols carry a "scipy_" prefix): the reference calls
 * un-vendored CBLAS/LAPACK at exactly these points.
 */
#ifndef DLU_ORACLE_H
#define DLU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Load a BLAS/LAPACK library.  `path` is a shared-object path, `prefix` is prepended to every
 * symbol ("" for MKL, "scipy_" for scipy's OpenBLAS).  Returns 0 on success. */
int dlu_oracle_init(const char *path, const char *prefix);
/* Human-readable identification of the loaded library ("MKL 2021.4 ...", "OpenBLAS ..."). */
const char *dlu_oracle_blas_name(void);
/* Set the BLAS thread count (no-op when the library has no such call).  Returns the count set. */
int dlu_oracle_set_threads(int threads);

/* Blocked delayed LU, restating delayed_getrf.  nb <= 0 selects the reference's 64.
 * `margins` may be NULL; otherwise it receives, for each column-pivot step k in [0, info),
 * (|top1| - |top2|) / |top1| over the active column (instrumentation only; -1 for steps whose
 * pivot came from the full-pivot fallback). `fallbacks` may be NULL. */
void dlu_oracle_delayed_getrf(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                              int *info, int nb, double *margins, int *fallbacks);

/* The column-delay variant, restating delayed_getrf_column
 * (delayed_lu/delayed_dgetrf_column.cpp:99-211): a zero column falls back to the rightmost column
 * j > k with a nonzero entry in rows [k, m) (find_nonzero_column, :30-45).  Same arguments. */
void dlu_oracle_delayed_getrf_column(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                     int *info, int nb, double *margins, int *fallbacks);

/* LAPACK dgetrf_ with 0-based ipiv, restating lapack_getrf. */
void dlu_oracle_lapack_getrf(int m, int n, double *a, int lda, int *ipiv, int *info);

/* LAPACK sgetrf_ (FP32 partial pivoting) with 0-based ipiv: with no zero columns the FP32
 * delayed LU is FP32 partial pivoting, so this pins its pivots up to near-ties.  Returns 1 when
 * the loaded library has no sgetrf_. */
int dlu_oracle_lapack_sgetrf(int m, int n, float *a, int lda, int *ipiv, int *info);

/* Unblocked full restatement in long double on a copy (no BLAS).  `a` is overwritten by the
 * rounded-to-double factors. */
void dlu_oracle_unblocked_ld(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info);
/* The same for the column-delay variant. */
void dlu_oracle_unblocked_column_ld(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                    int *info);

/* ||P*A*Q - Lbar*Ubar||_F / (max(m,n) * eps * ||A||_F) for the compact output of any of the
 * routines above.  `jpiv` may be NULL (row pivoting only).  `eliminated` is info (the number
 * of eliminated steps); rows/columns at and after it hold the Schur complement when
 * `delayed` is nonzero.  Returns -1 on allocation failure. */
double dlu_oracle_backward_error(int m, int n, const double *original, const double *factor,
                                 int lda, const int *ipiv, const int *jpiv, int eliminated,
                                 int delayed);

/* Delayed Bunch-Kaufman LDL^T of the symmetric n x n matrix held in the lower triangle of `a`
 * (the upper triangle is neither read nor written): P^T A P = L D L^T with 1x1 and 2x2 blocks in
 * D, nodes that take neither deferred to the front of a delayed suffix [info, n) whose Schur
 * complement stays in place.  perm[pos] = original index of the node at pos; piv_size[pos] = 1
 * (1x1), 2 (first column of a 2x2), 0 (its second column, or delayed).  info = -1 on invalid
 * arguments, -2 when the BLAS lacks a routine.
 *   dlu_oracle_delayed_sytrf          delayed-bunch-kaufman/delayed_sytrf.cpp:177-282 (unblocked)
 *   dlu_oracle_delayed_sytrf_blocked  delayed-bunch-kaufman/delayed_sytrf_blocked.cpp:168-358
 *                                     (panels of at most nb columns; the reference uses 64)
 *   dlu_oracle_sytrf_unblocked_ld     an independent unblocked restatement in long double with
 *                                     no BLAS (pins the two above) */
void dlu_oracle_delayed_sytrf(int n, double *a, int lda, int *perm, int *piv_size, int *info);
void dlu_oracle_delayed_sytrf_blocked(int n, double *a, int lda, int *perm, int *piv_size,
                                      int *info, int nb);
void dlu_oracle_sytrf_unblocked_ld(int n, double *a, int lda, int *perm, int *piv_size,
                                   int *info);

#ifdef __cplusplus
}
#endif
#endif
