/*
 * dlu.h -- C ABI of libdlu, the MI355X (gfx950) delayed-pivoting LU.
 *
 * Drop-in boundary for the reference's GPU entry points
 *   delayed_lu_cuda/delayed_dgetrf_cuda.hpp:7-63 (status enum, device/host/native/hybrid entry
 *   points, status string), whose semantics are kept:
 *   - column-major storage, A(i,j) at a[i + j*lda], lda >= max(1,m);
 *   - in-place compact output: unit-lower L multipliers below the diagonal of the eliminated
 *     columns, U above, and the Schur complement of the delayed part in [info,m) x [info,n);
 *   - ipiv/jpiv: 0-based swap histories of length min(m,n) ("at step k swap row k with row
 *     ipiv[k]", likewise columns), identity for steps at and after info;
 *   - info = min(m,n) when nothing is delayed, info = k when the active block at step k is
 *     entirely zero, info = -1 whenever the returned status is not DLU_SUCCESS;
 *   - column pivoting first (largest |a|, lowest row on ties); a zero (or NaN-only) column falls
 *     back to the largest |a| of the whole active block (lowest column-major offset on ties);
 *   - ipiv, jpiv and info are HOST memory in every variant; the call runs on `stream` and
 *     synchronizes it before returning.
 * Names carry no vendor prefix; the ordinal values of dlu_status match the reference enum.
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 */
#ifndef DLU_H
#define DLU_H

#ifdef __cplusplus
extern "C" {
#endif

/* delayed_lu_cuda/delayed_dgetrf_cuda.hpp:7-13 (success, invalid_argument, allocation_failed,
 * cuda_failure, cublas_failure), same ordinals. */
typedef enum dlu_status {
  DLU_SUCCESS = 0,
  DLU_INVALID_ARGUMENT = 1,
  DLU_ALLOCATION_FAILED = 2,
  DLU_RUNTIME_FAILURE = 3, /* a HIP runtime call failed */
  DLU_KERNEL_FAILURE = 4   /* a kernel reported failure (e.g. a bounded wait expired) */
} dlu_status;

/* Replaces delayed_getrf_cuda_device (delayed_lu_cuda/delayed_dgetrf_cuda.hpp:24-26,
 * delayed_dgetrf_cuda.cu:403-447).  d_a is device memory.  Calls from several host threads on
 * one device run concurrently (up to 4 at a time, fewer while their panels would not fit on the
 * CUs together; further calls wait), as the reference's calls on different streams do. */
dlu_status dlu_getrf_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv, int *info,
                            void *stream);

/* Replaces delayed_getrf_cuda_host (delayed_dgetrf_cuda.hpp:35-37, .cu:449-488): copies lda*n
 * doubles to the device, factors there, copies the result back into a. */
dlu_status dlu_getrf_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info,
                          void *stream);

/* Replace delayed_getrf_cuda_native_{device,host} (delayed_dgetrf_cuda.hpp:40-46): aliases. */
dlu_status dlu_getrf_native_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                   int *info, void *stream);
dlu_status dlu_getrf_native_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                 int *info, void *stream);

/* Replace delayed_getrf_cuda_hybrid_{device,host} (delayed_dgetrf_cuda.hpp:55-61).  The
 * reference's CPU-panel/GPU-update path is not reproduced: the panel runs on the GPU here, so
 * these are aliases of the device path with the same signature and results. */
dlu_status dlu_getrf_hybrid_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                   int *info, void *stream);
dlu_status dlu_getrf_hybrid_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                 int *info, void *stream);

/* Replaces delayed_getrf_cuda_status_string (delayed_dgetrf_cuda.hpp:63). */
const char *dlu_status_string(dlu_status status);

/* The column-delay variant on the GPU: delayed_getrf_column (delayed_lu/delayed_dgetrf_column.hpp:19,
 * delayed_dgetrf_column.cpp:99-211), whose reference exists on the CPU only.  Same arguments,
 * storage and rules as dlu_getrf_device, except the fallback taken when the active part of
 * column k is zero: instead of the largest |a| of the whole active block, the rightmost column
 * j > k with a nonzero entry in rows [k, m) is brought to position k (its largest |a|, lowest row
 * on ties, becomes the pivot), so zero columns collect at the right end
 * (delayed_dgetrf_column.cpp:30-45).  The _host form copies lda*n doubles in and out, as
 * dlu_getrf_host does. */
dlu_status dlu_getrf_column_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                   int *info, void *stream);
dlu_status dlu_getrf_column_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                 int *info, void *stream);

/* No counterpart in the reference: dlu_getrf_device for callers that cannot allocate (code
 * recorded under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and rules as dlu_getrf_device; it calls no allocating routine.  Its scratch
 * is a fixed device array of about 2.7 MB per GPU, held until the process ends; min(m, n) must be
 * at most 32768 (else DLU_INVALID_ARGUMENT, info = -1).  The factorization runs in order on
 * `stream` (no look-ahead, so it is slower than dlu_getrf_device).  Every call on one GPU shares
 * that GPU's array: the library makes such calls take turns, and each call has finished its
 * device work when it returns. */
dlu_status dlu_getrf_device_noalloc(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                    int *info, void *stream);

/* No counterpart in the reference (whose calls free everything before returning): frees the
 * device memory, pinned host memory, streams and events libdlu keeps between calls for the
 * calling process's idle factorizations and the calling thread's solves.  Call it before the
 * process exits (the Python binding registers it with atexit) so that nothing of libdlu is left
 * for the HIP runtime to tear down; later calls allocate again as needed. */
void dlu_shutdown(void);

/* ---- Extensions with no counterpart in the reference ------------------------------------ */

/* Tuning knobs.  Zero-initialised fields take the defaults. */
typedef struct dlu_options {
  int nb;          /* panel width: 64, 128 or 256 (default 128) */
  int lookahead;   /* 1 (default): factor panel j+1 while panel j's far update runs; -1: off */
  long long spin_limit_ticks; /* bound on any in-kernel wait, 100 MHz ticks (default 2e8 = 2 s) */
  int profile;     /* nonzero: time every panel / solve / update launch with HIP events */
  int panel_cus;   /* CUs reserved for panel kernels while updates run on the rest: > 0 = that
                      many (whole XCDs), <= 0 = none (the default) */
  int rows_per_lane; /* panel rows held per lane: 0 = automatic, else 1 to 4 (the exchange-and-move
                        panel rounds 3 up to 4) */
  int merge_rows;    /* pairs of panels share one trailing GEMM (K = 2 nb) while at least this
                        many rows remain below them: 0 = automatic (3000), < 0 = never */
  int split_update;  /* the far trailing GEMM's last columns run on the panel stream once the
                        next panel is done, sized so both streams finish together: 0 = automatic
                        (only when panel_cus reserves CUs), < 0 = never, > 0 = always that many
                        columns */
  int fallback;      /* what a zero column falls back to: DLU_FALLBACK_FULL (delayed_getrf: the
                        largest |a| of the active block) or DLU_FALLBACK_COLUMN
                        (delayed_getrf_column: the rightmost nonzero column) */
  int panel_kernel;  /* DLU_PANEL_AUTO (0): one exchange per pivot step wherever the panel fits
                        it (up to 128 workgroups, nb 64 or 128), else the exchange-and-move
                        kernel; DLU_PANEL_MOVE: always the kernel that moves rows between
                        workgroups at every step (two exchanges per step); DLU_PANEL_LAG: one
                        exchange per step, an error where it does not fit */
  int defer_left;    /* row interchanges of the columns left of each panel: 0 (default) = once
                        at the end, composed (up to 16384 rows, until a full-pivot fallback),
                        < 0 = as each panel ends */
} dlu_options;

#define DLU_PANEL_AUTO 0
#define DLU_PANEL_MOVE 1
#define DLU_PANEL_LAG 2

#define DLU_FALLBACK_FULL 0
#define DLU_FALLBACK_COLUMN 1

/* What happened during one factorization. */
typedef struct dlu_report {
  int panels;       /* panel kernels launched (restarts included) */
  int fallbacks;    /* full-pivot fallbacks taken */
  /* The tie statistics below are kept when the caller passes a margins array to
   * dlu_getrf_device_ex (otherwise 0, 0 and 2.0): the bookkeeping costs time on every pivot step. */
  int near_ties;    /* column steps whose relative top-2 magnitude gap was below 1e-10 */
  int exact_ties;   /* column steps whose two largest magnitudes were exactly equal */
  double min_margin;/* smallest relative top-2 gap over the column steps (2.0 if none) */
  /* filled when dlu_options.profile is set: summed launch durations (event-timed on the stream
   * each kernel runs on) and the algorithmic flops of the trailing updates (2*M*N*K each) */
  int update_launches;
  double update_seconds;
  double update_flops;
  int large_update_launches;   /* the subset over at least 512 blocks of 128 x 128 of C */
  double large_update_seconds;
  double large_update_flops;
  int panel_launches;
  double panel_seconds;
  int solve_launches;
  double solve_seconds;
  /* panel kernel, workgroup 0, summed over all steps (profile only): local argmax, publish,
   * candidate exchange, reduction, pivot-row fetch, interchange+scale, rank-1 update, tail */
  double panel_phase_seconds[8];
  double large_update_bytes;   /* least memory traffic of the large updates: C read and written
                                  once, A and B read once (profile only) */
  /* full-pivot searches of the active block (profile only): launches, summed durations and the
   * bytes they must read, sizeof(T) * (m - k) * (n - k) each */
  int search_launches;
  double search_seconds;
  double search_bytes;
} dlu_report;

/* dlu_getrf_device with options and a report (either may be NULL).  `margins`, when not NULL,
 * is a HOST array of min(m,n) doubles receiving each column step's relative top-2 gap (-1 for
 * steps decided by a full-pivot fallback or never reached). */
dlu_status dlu_getrf_device_ex(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                               int *info, void *stream, const dlu_options *options,
                               dlu_report *report, double *margins);

/* The same factorization in IEEE binary32: the FP32 factor of the mixed-precision solver
 * (dlu_gesv_mixed_device).  Same semantics and pivoting rules as dlu_getrf_device, with every
 * product, multiplier and update computed in FP32 (FP32 MFMA for the trailing update). */
dlu_status dlu_sgetrf_device(int m, int n, float *d_a, int lda, int *ipiv, int *jpiv, int *info,
                             void *stream);
dlu_status dlu_sgetrf_device_ex(int m, int n, float *d_a, int lda, int *ipiv, int *jpiv,
                                int *info, void *stream, const dlu_options *options,
                                dlu_report *report, double *margins);

/* ---- Solve side (no counterpart in the reference, which factors only; its example checks the
 * factors by reconstruction, delayed_lu/example.cpp:43-126) ------------------------------------ */

/* Solve A X = B with the factors of dlu_getrf_device (P*A*Q = L*U): X = Q U^-1 L^-1 P B, the
 * row swaps applied to B in step order and the column swaps to the solution in reverse order.
 * Needs a complete elimination (square, info == n); a delayed suffix has no solve here.
 * d_b (n x nrhs, leading dimension ldb) is overwritten by X.  ipiv/jpiv are the HOST arrays the
 * factorization returned.  Synchronizes `stream` before returning. */
dlu_status dlu_getrs_device(int n, int nrhs, const double *d_lu, int lda, const int *ipiv,
                            const int *jpiv, int info, double *d_b, int ldb, void *stream);
dlu_status dlu_sgetrs_device(int n, int nrhs, const float *d_lu, int lda, const int *ipiv,
                             const int *jpiv, int info, float *d_b, int ldb, void *stream);

/* Per right-hand side j, in FP64 on the device:
 *   bwd[j] = ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)   (bwd: HOST array). */
dlu_status dlu_backward_error_device(int n, int nrhs, const double *d_a, int lda,
                                     const double *d_x, int ldx, const double *d_b, int ldb,
                                     double *bwd, void *stream);

typedef struct dlu_mixed_options {
  double tol;        /* target backward error of every column (default 1e-13) */
  int max_iter;      /* refinement sweeps at most (default 30) */
  int nb;            /* panel width of the factorization (0: library default) */
  int fp64_fallback; /* 0 (default): factor in FP64 when the FP32 factor leaves a delayed suffix;
                        < 0: report DLU_INVALID_ARGUMENT instead */
} dlu_mixed_options;

typedef struct dlu_mixed_report {
  int factor_info;   /* info of the (first) factorization */
  int iterations;    /* solve + residual sweeps */
  int converged;     /* 1 when every column met tol */
  int used_fp64;     /* 1 when the factors used were FP64 */
  double factor_seconds, refine_seconds, total_seconds;  /* event-timed on `stream` */
  double max_backward_error;
} dlu_mixed_report;

/* A X = B, FP64 in and out (A untouched).  dlu_gesv_mixed_device factors an FP32 copy of A with
 * the FP32 delayed LU, then refines in FP64: x += solve32(b - A x) until every column's backward
 * error (as in dlu_backward_error_device) is at most tol.  dlu_gesv_device factors an FP64 copy
 * and runs the same loop (normally one sweep).  bwd (HOST, nrhs doubles, may be NULL) receives the
 * final backward errors.  Synchronous on `stream`. */
dlu_status dlu_gesv_device(int n, int nrhs, const double *d_a, int lda, const double *d_b,
                           int ldb, double *d_x, int ldx, void *stream,
                           const dlu_mixed_options *options, dlu_mixed_report *report,
                           double *bwd);
dlu_status dlu_gesv_mixed_device(int n, int nrhs, const double *d_a, int lda, const double *d_b,
                                 int ldb, double *d_x, int ldx, void *stream,
                                 const dlu_mixed_options *options, dlu_mixed_report *report,
                                 double *bwd);

/* ---- Batched small matrices and serving (no counterpart in the reference) ------------------ */

#define DLU_BATCH_MAX_N 128  /* largest n of the one-workgroup-per-matrix kernel */
#define DLU_BATCH_MAX_RHS 8

/* One matrix of a batch.  Factor mode: factor A (same pivoting rules as dlu_getrf_device),
 * write the factors, pivots and info, and when info == n solve for the right-hand sides.
 * Solve-only mode: `a` holds factors from an earlier factor-mode item (ipiv/jpiv/info_in
 * likewise) and only the solve runs.  All pointers are device memory; lu may alias a and x may
 * alias b. */
typedef struct dlu_batch_item {
  const double *a;
  int lda;
  double *lu;
  int ldlu;
  int *ipiv;
  int *jpiv;
  int *info;
  int info_in;
  int n;
  int nrhs;
  const double *b;
  int ldb;
  double *x;
  int ldx;
  int solve_only;
} dlu_batch_item;

/* One launch, one workgroup per item (n <= DLU_BATCH_MAX_N, nrhs <= DLU_BATCH_MAX_RHS); the
 * item array is HOST memory.  Synchronous on `stream`.  Each item's results do not depend on
 * the other items of the batch. */
dlu_status dlu_batch_gesv_device(int count, const dlu_batch_item *items, void *stream);

/* A solve service on one device with continuous batching.  Requests (factor + solve of an
 * n x n FP64 system, or a solve with the resident factor of an earlier request) are admitted as
 * memory allows; those with n <= the batch cutoff are packed into one batched launch per round
 * (dlu_batch_gesv_device), while the others run through the full factorization on a second
 * worker, so both kinds overlap on the GPU.  Factors asked to be kept stay resident until the
 * HBM budget needs their memory (least recently used first).  Results do not depend on which
 * other requests were in flight. */
typedef struct dlu_server dlu_server;

typedef struct dlu_server_config {
  double hbm_budget_gib; /* resident factors + per-request buffers (default 32) */
  int batch_cutoff;      /* largest n for the batched kernel, <= DLU_BATCH_MAX_N (0: the
                            autotuned table's value, dlu_tuned_batch_cutoff(); < 0: timed at
                            creation on this device by running both paths) */
  int mixed;             /* 1: FP32 factor + FP64 refinement above the cutoff; 0: FP64 */
  double tol;            /* refinement target backward error (default 1e-13) */
  int max_batch;         /* most requests per batched launch (default 1024) */
  int large_workers;     /* requests above the cutoff factored at once, each on its own stream
                            (default 4, at most 4) */
  int eager_sweeps;      /* 0 (default): a later right-hand side of a resident factor (above
                            the cutoff, nrhs <= 128) replays the factor's refinement sweep from a
                            HIP graph recorded at its first reuse; 1: every sweep is launched
                            kernel by kernel.  The results are the same bit for bit. */
} dlu_server_config;

/* The service's streams are not ordered after the caller's: a, b and x must hold their final
 * contents (or allocations) before dlu_server_submit, e.g. after a stream synchronization. */
typedef struct dlu_request {
  long long id;        /* caller's id, unique among live requests */
  int n, nrhs;         /* nrhs <= DLU_BATCH_MAX_RHS for batched requests */
  const double *a;     /* device, the matrix (read until completion; also for reuse requests) */
  int lda;
  const double *b;     /* device right-hand sides */
  int ldb;
  double *x;           /* device solution */
  int ldx;
  long long factor_of; /* -1: factor a; else the id of an earlier request whose kept factor to
                          reuse (factored again from a if it was evicted) */
  int keep;            /* 1: keep this request's factor resident */
} dlu_request;

typedef struct dlu_completion {
  long long id;
  int status;            /* dlu_status */
  int info;              /* of the factorization used */
  int path;              /* 0 batched, 1 full FP64, 2 full FP32 + refinement */
  int iterations;        /* refinement sweeps (full paths) */
  int reused;            /* 1 when a resident factor served the request */
  double t_submit, t_start, t_done;  /* seconds on the server clock (dlu_server_now) */
  double backward_error; /* largest over the columns (full paths; -1 when not computed) */
  int graph_sweeps;      /* refinement sweeps replayed from a HIP graph (0: launched eagerly) */
} dlu_completion;

dlu_server *dlu_server_create(const dlu_server_config *config);
int dlu_server_batch_cutoff(const dlu_server *server);
dlu_status dlu_server_submit(dlu_server *server, const dlu_request *request);
/* Moves up to `max` completions into `out`; returns how many. */
int dlu_server_poll(dlu_server *server, dlu_completion *out, int max);
/* Blocks until every submitted request has completed. */
dlu_status dlu_server_drain(dlu_server *server);
double dlu_server_now(const dlu_server *server);
void dlu_server_destroy(dlu_server *server);

/* ---- Delayed Bunch-Kaufman LDL^T (delayed-bunch-kaufman/; CPU-only in the reference) ------- */

/* delayed_sytrf_blocked (delayed-bunch-kaufman/delayed_sytrf_blocked.hpp:16, .cpp:168-358) on the
 * GPU.  The symmetric n x n matrix is the lower triangle of the column-major d_a (the upper
 * triangle is neither read nor written).  On return P^T A P = L D L^T: D's 1x1 and 2x2 blocks on
 * the diagonal (a 2x2 block's off-diagonal at (k+1, k)), L's multipliers below them (unit
 * diagonal implied), and nodes that took neither a stable 1x1 nor a stable 2x2 pivot in the
 * delayed suffix [info, n), whose Schur complement stays in place (lower triangle).  perm[pos]
 * (host) = the original index of the node at pos; piv_size[pos] (host) = 1 for a 1x1 pivot, 2 for
 * the first column of a 2x2 pivot, 0 for its second column or a delayed node.  info (host) = the
 * number of eliminated nodes.  Panels of 64 columns, as the reference's.  The call synchronizes
 * `stream` before it returns.  The _host form copies lda*n doubles in and out. */
dlu_status dlu_sytrf_device(int n, double *d_a, int lda, int *perm, int *piv_size, int *info,
                            void *stream);
dlu_status dlu_sytrf_host(int n, double *a, int lda, int *perm, int *piv_size, int *info,
                          void *stream);
/* No counterpart in the reference: dlu_sytrf_device with panels of nb columns (2 <= nb <= 64),
 * for tests of the panel boundaries. */
dlu_status dlu_sytrf_device_nb(int n, double *d_a, int lda, int *perm, int *piv_size, int *info,
                               int nb, void *stream);

/* ---- Frontal-matrix steps (the design note's multifrontal use of the delayed suffix,
 * doc/（非）对称稠密矩阵的延迟主元LU方法.md:20-440; the reference implements only Step 1, as
 * delayed_getrf) ------------------------------------------------------------------------------ */

/* Steps 1 and 2 of an nf x nf front (device, column-major, leading dimension ldf) whose leading
 * ns x ns block is fully summed:
 *   Step 1: dlu_getrf_device on the fully summed block: P F_SS Q = [L_EE; L_DE] [U_EE U_ED],
 *           info = the number of eliminated steps e (the fully summed variables [e, ns) are
 *           delayed; F[e:ns, e:ns] holds their Schur complement S_DD);
 *   the interchanges extend over the front: the fully summed rows are interchanged across all nf
 *   columns and the fully summed columns down all nf rows;
 *   Step 2: U_EU = L_EE^-1 A_EU, L_UE = A_UE U_EE^-1, A_DU -= L_DE U_EU, A_UD -= L_UE U_ED,
 *           A_UU -= L_UE U_EU (doc :300-317).
 * On return F holds [[L\U_EE, U_ED, U_EU], [L_DE, S_DD, S_DU], [L_UE, S_UD, S_UU]]: the factors
 * of the eliminated variables and the contribution block F[e:nf, e:nf] for the parent.  ipiv and
 * jpiv (HOST, ns entries) are the fully summed block's swap histories, as dlu_getrf_device
 * returns them.  Synchronous on `stream`. */
dlu_status dlu_front_factor_device(int nf, int ns, double *d_f, int ldf, int *ipiv, int *jpiv,
                                   int *info, void *stream);

/* Step 3, the extend-add (doc :360-420): parent(map_r[i], map_c[j]) += child(i, j) for i < rows,
 * j < cols; child is a rows x cols contribution block (leading dimension ldc), map_r / map_c
 * (device int arrays) give each of its rows / columns a position in the parent front (leading
 * dimension ldp); the entries of each map must be distinct.  Several children are added by
 * successive calls.  Asynchronous on `stream`. */
dlu_status dlu_extend_add_device(int rows, int cols, const double *d_child, int ldc,
                                 const int *d_map_r, const int *d_map_c, double *d_parent,
                                 int ldp, void *stream);

/* Kernel-level entry point of the trailing update: C(MxN) -= A(MxK) * B(KxN), column-major
 * device matrices sharing leading dimension ld.  Asynchronous on `stream`. */
dlu_status dlu_gemm_sub_device(int M, int N, int K, double *C, const double *A, const double *B,
                               int ld, void *stream);

/* The same with a named tile configuration (0: 64 x 64, 1: 128 x 128, 2: 128 x 64, 3: 64 x 128;
 * < 0: the tuned table's choice), for the autotuner and its tests. */
dlu_status dlu_gemm_sub_device_config(int M, int N, int K, double *C, const double *A,
                                      const double *B, int ld, int config, void *stream);

/* The tile configuration the autotuned table (delayed-pivoting_amd/csrc/dlu_tuning.inc, written
 * by delayed-pivoting_amd/autotune.py) picks for an M x N x K trailing update; and the batch
 * cutoff it holds (the solve service's default). */
int dlu_tuned_gemm_config(int M, int N, int K);
int dlu_tuned_batch_cutoff(void);

#ifdef __cplusplus
}
#endif
#endif /* DLU_H */
