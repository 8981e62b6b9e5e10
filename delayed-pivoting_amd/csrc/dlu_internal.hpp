// Internal declarations shared by the HIP translation units of libdlu.
//
// Device-side state of one factorization lives in a Workspace (see dlu_getrf.cpp).  The control
// block below is written by kernels and read by later kernels (predication) and, once per panel,
// by the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "dlu.h"

namespace dlu {

// Control block written on the device.  One per workspace.
struct Control {
  int stop;          // panel kernel: steps [pb, stop) of the last panel are eliminated
  int zero_column;   // panel kernel: 1 when it stopped because column `stop` is zero
  int deferred;      // fallback: 1 when the whole active block at `stop` was zero (info = stop)
  int error;         // nonzero: a bounded wait expired (2) or an internal check failed (1)
  int fallbacks;     // number of full-pivot fallbacks taken so far
  int near_ties;     // column steps whose relative top-2 gap was below 1e-10
  int exact_ties;    // column steps whose two largest magnitudes were equal
  double min_margin; // smallest relative top-2 gap seen at a column step
  // full-pivot search result (fallback)
  double best_value;
  int best_row;
  int best_col;
  unsigned search_ticket;
  unsigned abort_flag;  // set by a panel workgroup whose wait expired; the others leave too
  int col_found;        // column-variant search: rightmost nonzero column so far (-1: none)
  // The done_seq of the last panel that eliminated all its columns (no zero column, nothing
  // deferred): the token the launches queued behind it on speculation test.
  unsigned long long go;
};

// A launch queued before the panel it follows has finished, on the guess that the panel
// eliminates all its columns: it does its work only when *go (Control::go) holds that panel's
// token, and returns at once otherwise.  go == nullptr: unconditional.
struct Pred {
  const unsigned long long *go;
  unsigned long long expect;
};
__device__ __forceinline__ bool skipped(const Pred &p) {
  return p.go != nullptr && *p.go != p.expect;
}

// What the host needs of a panel, as the panel publishes it into pinned, coherent host memory:
// one 64-bit word written by one system-scope store as the panel's last act (no fence: the
// fields travel together), which the host spins on.  Bits [0, 32) the stop, 32 zero column,
// 33 deferred, [34, 36) the error code (capped at 3), [36, 64) the low 28 bits of the launch's
// sequence number.  The rest of the control block (tie statistics, fallback count) is read once,
// at the end of the factorization.
struct PanelDone {
  unsigned long long word;
};
constexpr unsigned long long kDoneSeqMask = (1ull << 28) - 1;

__device__ __forceinline__ void publish_done(const Control *ctl, PanelDone *done,
                                             unsigned long long seq) {
  if (done == nullptr) return;
  const int err = ctl->error;
  const unsigned long long w =
      static_cast<unsigned long long>(static_cast<unsigned>(ctl->stop)) |
      (static_cast<unsigned long long>(ctl->zero_column != 0) << 32) |
      (static_cast<unsigned long long>(ctl->deferred != 0) << 33) |
      (static_cast<unsigned long long>(err < 0 ? 3 : (err > 3 ? 3 : err)) << 34) |
      ((seq & kDoneSeqMask) << 36);
  __hip_atomic_store(&done->word, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Exchange buffer geometry of the panel kernel (in 64-bit words).  Every word carries the step
// tag of the step that wrote it, so a reader can tell a fresh word from a stale one without any
// cache-maintenance fence.  A slot starts with the two header words (48-bit payload, 16-bit tag
// in the top bits: the candidate's |value| and row), then the second-largest |value| as two data
// words, padding, and the candidate row as data words (32-bit payload, tag in the high half).
// Tags run from 1 to kMaxStepTag; the host clears all exchange words before they would wrap.
constexpr int kSlotHeaderWords = 8;
constexpr unsigned kMaxStepTag = 0xFFFFu;
inline constexpr int slot_words(int nb) { return kSlotHeaderWords + 2 * nb; }
constexpr int kMaxPanelGroups = 256;
// The one-exchange-per-step panel kernel (dlu_panel_lag.hip) lays the same slot memory out as
// two step parities of kLagMaxGroups 8-word headers, then four step slots of kLagMaxGroups
// candidate rows (two words per element, nb <= 128); it takes at most kLagMaxGroups workgroups.
constexpr int kLagMaxGroups = 128;
static_assert(2 * kLagMaxGroups * 8 + 4 * kLagMaxGroups * 2 * 128 <=
                  2 * kMaxPanelGroups * (kSlotHeaderWords + 2 * 256),
              "the lagged panel's exchange fits the slot memory");

// 16 x 16 x 4 MFMA tiles in either precision (gfx950: v_mfma_f64_16x16x4_f64 and
// v_mfma_f32_16x16x4_f32).  Operand layout, lane l of 64: A (16 x 4) holds row l % 16, k l / 16;
// B (4 x 16) holds k l / 16, column l % 16.  The 16 x 16 result is four values per lane in
// column l % 16; their rows differ between the two instructions (acc_row).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int acc_row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int acc_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Full-pivot search order (delayed_lu/delayed_dgetrf.cpp:32-47): the larger magnitude, then the
// lower column-major offset in the active block; NaN never wins (both comparisons are false).
__device__ __forceinline__ bool better(double v, unsigned long long o, double bv,
                                       unsigned long long bo) {
  return v > bv || (v == bv && o < bo);
}

// The best (value, offset) of a 256-thread workgroup, in thread 0.
__device__ __forceinline__ void block_best(double &v, unsigned long long &o) {
  __shared__ double sv[4];
  __shared__ unsigned long long so[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off);
    const unsigned long long oo = __shfl_xor(o, off);
    if (better(ov, oo, v, o)) {
      v = ov;
      o = oo;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sv[wave] = v;
    so[wave] = o;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; ++i)
      if (better(sv[i], so[i], v, o)) {
        v = sv[i];
        o = so[i];
      }
  }
}

// The full-pivot argmax of an update's result, written by the update itself: workgroup b of the
// GEMM leaves its best (|value|, offset) at v[b], o[b]; offsets are column-major in the active
// block whose row 0 is the GEMM's row 0 and whose column 0 lies col_shift columns left of the
// GEMM's column 0, with `rows` rows.
struct AmaxOut {
  double *v;
  unsigned long long *o;
  int col_shift;
  int rows;
};
constexpr int kAmaxPartials = 1 << 16;  // most GEMM workgroups an update may leave partials for

// A panel's net row movement: [0, 2nb) destination rows, [2nb, 4nb) source rows, [4nb] count;
// [4nb+1] and [4nb+2] are the panel kernel's append counter and finished-workgroup count (zero
// between panels).  Two such lists alternate between consecutive panels so the far update of
// panel j can still read its list while panel j+1 writes the next one.
inline constexpr int swap_list_ints(int nb) { return 4 * nb + 4; }
// Row k exchange area per step parity: row k's values (at most 2 words per element), then the
// original index of the row whose content it holds.
inline constexpr int rowk_words(int nb) { return 2 * nb + 8; }

template <typename T>
struct PanelParams {
  T *a;
  int lda;
  int m;
  int pb;         // first column (and row) of the panel
  int pe;         // one past the last column of the panel
  int positioned; // the pivot of step pb is already in place (after a full-pivot fallback)
  int groups;     // number of workgroups (row blocks) taking part
  unsigned seq_base;
  int *ipiv;           // device ipiv (0-based swap history)
  double *margins;     // device per-step relative top-2 gap (may be null)
  Control *ctl;
  uint64_t *slots;     // [2][groups][slot_words(nb)]
  uint64_t *rowk;      // [2][rowk_words(nb)]
  int *swap_list;      // swap_list_ints(nb) ints for this panel (parity selected by the host)
  long long spin_limit;  // bounded wait, in s_memrealtime ticks (100 MHz)
  unsigned long long *phase_ticks;  // optional [8] per-phase tick totals (workgroup 0)
  PanelDone *done;                  // host-visible copy of the control block (may be null)
  unsigned long long done_seq;      // value of done->seq once this launch has published
  Pred pred;                        // a speculative launch: it factors nothing unless pred holds
};

// Launchers (return hipSuccess or the launch error).  T is float or double; both are
// instantiated in the translation unit that defines each launcher.
template <typename T>
hipError_t launch_panel(const PanelParams<T> &p, int nb, int rows_per_lane, hipStream_t s);
// The same panel with one exchange per step (rows tracked by logical position, the bulk update
// one step behind a leading wave); at most kLagMaxGroups workgroups, rows_per_lane <= 2 at
// nb = 256.  Leaves the matrix, ipiv, margins, swap list and control block as launch_panel does.
template <typename T>
hipError_t launch_panel_lag(const PanelParams<T> &p, int nb, int rows_per_lane, hipStream_t s);

// Right-side interchange + triangular solve: rows [pb, pb+w) of columns [c0, c1) become
// U12 = L11^{-1} * (rows after interchange); rows displaced below pb+w are written directly.
template <typename T>
hipError_t launch_swap_trsm(T *a, int lda, int pb, int w, int c0, int c1, const int *swap_list,
                            int nb, hipStream_t s, Pred pred = Pred{nullptr, 0});

// The first panel of a full pair (rows from b0, list l0) as launch_swap_trsm does it, then the
// second panel's movement (rows from b0 + nb, list l1) on the same columns, in one launch.
template <typename T>
hipError_t launch_swap_trsm_pair(T *a, int lda, int b0, int c0, int c1, const int *l0,
                                 const int *l1, int nb, hipStream_t s, Pred pred);

// C(M x N) -= A(M x K) * B(K x N), column-major with leading dimensions ldc, lda, ldb.
// gemm_uses_large_tiles names the launches the profiled report counts as large (at least 512
// blocks of 128 x 128 of C: the far updates while many rows remain).
inline bool gemm_uses_large_tiles(int M, int N) {
  return static_cast<long long>((M + 127) / 128) * ((N + 127) / 128) >= 512;
}
// The tile configuration comes from the autotuned table (dlu_tuning.inc) unless `config` names
// one (0..3: 64x64, 128x128, 128x64, 64x128).
template <typename T>
hipError_t launch_gemm_sub(T *c, int ldc, const T *a, int lda, const T *b, int ldb, int M, int N,
                           int K, hipStream_t s, int config = -1, Pred pred = Pred{nullptr, 0},
                           AmaxOut amax = AmaxOut{nullptr, nullptr, 0, 0});
// Workgroups of one launch of tile configuration `config` over an M x N result (the partials an
// argmax epilogue leaves).
int gemm_workgroups(int config, int M, int N);
// The configuration the table picks for an M x N x K update.
int tuned_gemm_config(int M, int N, int K);
constexpr int kGemmConfigs = 4;

// Row interchanges of the last panel applied to columns [c0, c1) (left of the panel).
template <typename T>
hipError_t launch_row_interchange(T *a, int lda, int c0, int c1, const int *swap_list, int nb,
                                  hipStream_t s, Pred pred = Pred{nullptr, 0});

// Deferred left interchanges (factorizations of at most kDeferMaxRows rows): compose the kept
// net row movements of `events` panels backward, writing the permutation into snaps[slot] at
// every e with snap_slot[e] = slot >= 0 (one workgroup; P and Pinv hold m ints each), then
// gather rows [col_r0[c], m) of each column c < ncols with col_slot[c] >= 0 by snaps[slot].
constexpr int kDeferMaxRows = 16384;
hipError_t launch_left_compose(int m, const int *lists, int stride, int nb, int events,
                               const int *snap_slot, int *snaps, hipStream_t s);
template <typename T>
hipError_t launch_left_gather(T *a, int lda, int m, int ncols, const int *col_slot,
                              const int *col_r0, const int *snaps, hipStream_t s);

// Identity swap histories and "unreached" margins for a new factorization.
hipError_t launch_init_steps(int *ipiv, int *jpiv, double *margins, int steps, hipStream_t s);

// Full-pivot search over the active block A(k:m-1, k:n-1); result in ctl->best_*.  With
// ext_count > 0 the columns from n on were searched by an update's argmax epilogue, which left
// ext_count partials (ext_v, ext_o); they join the reduction.
template <typename T>
hipError_t launch_active_search(const T *a, int lda, int m, int n, int k, double *partial_v,
                                unsigned long long *partial_o, Control *ctl, hipStream_t s,
                                const double *ext_v = nullptr,
                                const unsigned long long *ext_o = nullptr, int ext_count = 0);

// Column-variant fallback search (delayed_getrf_column): the rightmost column j in (k, n) with a
// nonzero entry in rows [k, m), and its largest |a| (lowest row on ties); result in ctl->best_*
// (best_value = 0 when every such column is zero).
template <typename T>
hipError_t launch_column_search(const T *a, int lda, int m, int n, int k, Control *ctl,
                                hipStream_t s);

// Tall panels (more rows than the register-resident panel kernel holds): one launch sequence
// per column step, every launch predicated on the control block.  begin: stop = pe (pb after a
// deferral); per step k: the column argmax (launch_active_search over one column), then
// tall_pivot (zero column: stop = k; else ipiv[k] and the whole-row interchange), then
// tall_update (multipliers and the rank-1 update of the panel's columns); finish: the panel's
// swap list (identity: rows already moved) and the host copy of the control block.
hipError_t launch_tall_begin(Control *ctl, int pb, int pe, hipStream_t s);
template <typename T>
hipError_t launch_tall_pivot(T *a, int lda, int n, int k, int *ipiv, Control *ctl, hipStream_t s);
template <typename T>
hipError_t launch_tall_update(T *a, int lda, int m, int k, int pe, const Control *ctl,
                              hipStream_t s);
hipError_t launch_tall_finish(Control *ctl, int *swap_list, int nb, int pb, PanelDone *done,
                              unsigned long long seq, hipStream_t s);

// Apply the full-pivot decision at step k (row swap, then column swap) or mark the deferral.
template <typename T>
hipError_t launch_fallback_apply(T *a, int lda, int m, int n, int k, int *ipiv, int *jpiv,
                                 Control *ctl, hipStream_t s);

// ---- Solve-side launchers (dlu_solve.hip) ----------------------------------------------------

// dst(i, j) = src(perm[i], j) for i < n, j < nrhs (column-major; perm on the device).
template <typename T>
hipError_t launch_gather_rows(const T *src, int lds, T *dst, int ldd, const int *perm, int n,
                              int nrhs, hipStream_t s);

// Device state of the one-launch triangular solves: the exchange words of solved rows (tagged
// with the epoch of the launch that wrote them; 16 columns x (1 or 2) words per row, rows of the
// padded right-hand side), a ticket counter that only grows (each launch takes ticket_base ..
// ticket_base + blocks - 1) and an error word set by an expired wait.
// A launch's ticket base and epoch in device memory.
struct ChainClock {
  unsigned long long base;
  unsigned epoch;
  unsigned pad;
};
constexpr int kChainSlots = 16;  // chain launches per clock advance
// Device state of chain launches replayed from a HIP graph: the ticket counter and error word
// (as in the eager state), the last epoch handed out, and the clocks of the next launches.
struct ChainClocks {
  unsigned long long ticket;
  int err;
  unsigned epoch;
  ChainClock slot[kChainSlots];
};
struct ChainState {
  unsigned long long *words;
  unsigned long long *ticket;
  unsigned long long ticket_base;
  unsigned epoch;
  const ChainClock *clock;  // not null: base and epoch are read from here at run time
  long long spin_limit;     // 100 MHz ticks
  int *err;
};

// X := T^{-1} X in one launch, T the unit lower (upper = false) or non-unit upper triangle of
// the n x n factor; X has nrhs <= 16 columns and ldx >= n rounded up to a multiple of 64 (the
// padding rows are scratch).  Consumes ceil(n / 64) tickets.
template <typename T>
hipError_t launch_trsv_chain(const T *lu, int lda, int n, T *x, int ldx, int nrhs, bool upper,
                             const ChainState &cs, hipStream_t s);

// One workgroup: the clocks of the next nslots chain launches of `blocks` tickets each go to
// st->slot[0, nslots), and zero[0, nzero) is zeroed (a fill kernel, so a HIP graph holds no
// memset node).  Clears `words` (nwords) when the epochs would wrap.
hipError_t launch_chain_clock(ChainClocks *st, int nslots, int blocks, unsigned long long *words,
                              size_t nwords, double *zero, int nzero, hipStream_t s);

// R = B - A X in FP64 (A n x n, X and B n x nrhs).
hipError_t launch_residual(const double *a, int lda, const double *x, int ldx, const double *b,
                           int ldb, double *r, int ldr, int n, int nrhs, hipStream_t s);

// out[0] = max_i sum_j |A(i, j)| (the infinity norm of an n x n matrix); work holds 32*n
// doubles; out[0] must be zeroed first.
hipError_t launch_norm_inf_matrix(const double *a, int lda, int n, double *work, double *out,
                                  hipStream_t s);

// out[j] = max_i |X(i, j)| for j < nrhs (out must be zeroed first; accumulates with max).
hipError_t launch_norm_inf_columns(const double *x, int ldx, int n, int nrhs, double *out,
                                   hipStream_t s);

// Precision conversions (column-major, n x ncol): dst = float(src); x += double(d).
hipError_t launch_to_float(const double *src, int lds, float *dst, int ldd, int n, int ncol,
                           hipStream_t s);
template <typename T>
hipError_t launch_add_correction(double *x, int ldx, const T *d, int ldd, int n, int ncol,
                                 hipStream_t s);

// ---- Batched small matrices (dlu_batch.hip) ----------------------------------------------------
constexpr int kBatchMaxRhs = DLU_BATCH_MAX_RHS;
typedef dlu_batch_item BatchItem;
// One workgroup per item; d_items is a device array of `count` items, every n <= nmax <= 128.
hipError_t launch_batch_gesv(const BatchItem *d_items, int count, int nmax, hipStream_t s);
// a[i] in [-1, 1), deterministic in (seed, i).
hipError_t launch_fill_uniform(double *a, size_t count, unsigned long long seed, hipStream_t s);

// ---- Host drivers shared between translation units -------------------------------------------

// A factorization kept on the device for later solves (buffers owned by the caller).
struct DeviceFactor {
  int n = 0;
  bool fp32 = false;
  int info = -1;
  void *lu = nullptr;     // n x n, leading dimension n, float or double
  int *d_perm = nullptr;  // net row order of P b, then net order of Q y (n each)
};
dlu_status make_factor(int n, const double *d_a, int lda, bool fp32, int nb, float *lu32,
                       double *lu64, int *d_perm, DeviceFactor &f, int *first_info,
                       hipStream_t s);
dlu_status solve_refine(const DeviceFactor &f, int nrhs, const double *d_a, int lda,
                        const double *d_b, int ldb, double *d_x, int ldx, double tol,
                        int max_iter, hipStream_t s, int *iterations, double *bwd,
                        bool *converged);
// The refinement sweep of solve_refine for one factor, right-hand-side count and matrix,
// recorded once as a HIP graph with buffers of its own, and replayed for each later solve:
// b is copied in, the graph runs once per sweep (the host reads the backward errors between
// sweeps, as solve_refine does), and x is copied out.  Results equal solve_refine's bit for bit
// (the same kernels on the same values).  The factor and the matrix must stay in place while
// the plan lives; one solve at a time per plan.
struct RefinePlan;
constexpr int kPlanMaxRhs = kChainSlots / 2 * 16;
dlu_status refine_plan_create(const DeviceFactor &f, int nrhs, const double *d_a, int lda,
                              hipStream_t s, RefinePlan **out, size_t *bytes);
bool refine_plan_matches(const RefinePlan *p, int nrhs, const double *d_a, int lda);
dlu_status refine_plan_solve(RefinePlan *p, const double *d_b, int ldb, double *d_x, int ldx,
                             double tol, int max_iter, hipStream_t s, int *iterations,
                             double *bwd, bool *converged);
// Frees the plan after the work queued on s (the stream of its last use).
void refine_plan_destroy(RefinePlan *p, hipStream_t s);
dlu_status to_status(hipError_t e);
// Process teardown before the HIP runtime goes: the idle factorization workspaces of every
// device, and the calling thread's solve-side scratch (each serving worker frees its own).
void release_pools();
void release_thread_solve_cache();
// Device bytes held by the calling thread's solve-side scratch on `device`.
size_t thread_solve_cache_bytes(int device);
template <typename T>
dlu_status getrf_device(int m, int n, T *d_a, int lda, int *ipiv, int *jpiv, int *info,
                        hipStream_t stream, const dlu_options *opt, dlu_report *rep,
                        double *margins);

}  // namespace dlu
