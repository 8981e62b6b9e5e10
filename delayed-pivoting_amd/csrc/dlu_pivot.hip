// Pivot-side kernels of the delayed-pivoting LU on gfx950: the full-pivot search over the active
// block, the interchanges it decides, and the batched row interchange applied left of a panel.
//
// Reference behaviour restated here:
//   full-pivot search, lowest column-major offset wins ties
//       delayed_lu/delayed_dgetrf.cpp:32-47; delayed_lu_cuda/delayed_dgetrf_cuda.cu:13-20, 24-85
//   deferral when the active block is entirely zero       delayed_lu/delayed_dgetrf.cpp:189-192
//   row swap over all n columns, then column swap over all m rows   :202-209
//   NaN is never selected (both comparisons are false)    delayed_lu_cuda/delayed_dgetrf_cuda.cu:46
#include "dlu_internal.hpp"

#include <climits>

namespace dlu {
namespace {

// The active block is cut into items of one column x 512 rows; every wave of the grid takes
// items in increasing order (item = column * segments + segment), each lane with 8 independent
// coalesced loads in flight.  A lane therefore meets its elements in increasing column-major
// offset, so a strict > keeps the lowest offset among its own equal maxima; the reductions
// across lanes, waves and workgroups apply the full rule (larger magnitude, then lower offset).
// The last workgroup to finish reduces the partials.  Magnitudes are compared in double (exact
// for float entries); NaN compares false and is never selected.
template <typename T>
__global__ __launch_bounds__(256) void active_search_kernel(const T *__restrict__ a, int lda,
                                                           int m, int n, int k,
                                                           double *partial_v,
                                                           unsigned long long *partial_o,
                                                           Control *ctl,
                                                           const double *__restrict__ ext_v,
                                                           const unsigned long long *__restrict__ ext_o,
                                                           int ext_count) {
  constexpr int U = 8, SEG = 64 * U;
  const int rows = m - k, cols = n - k;
  const int nseg = (rows + SEG - 1) / SEG;
  const long long items = static_cast<long long>(cols) * nseg;
  const int lane = threadIdx.x & 63;
  const long long waves = static_cast<long long>(gridDim.x) * 4;
  double best = 0.0;  // a zero entry never displaces the initial candidate
  unsigned long long bo = 0ull;
  for (long long it = static_cast<long long>(blockIdx.x) * 4 + (threadIdx.x >> 6); it < items;
       it += waves) {
    const int col = static_cast<int>(it / nseg), seg = static_cast<int>(it % nseg);
    const T *cp = a + static_cast<size_t>(k) + static_cast<size_t>(k + col) * lda;
    const unsigned long long base = static_cast<unsigned long long>(col) * rows;
    const int r0 = seg * SEG + lane;
    T x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {  // clamped, unconditional loads
      const int r = r0 + 64 * u;
      x[u] = cp[r < rows ? r : 0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + 64 * u;
      const double v = r < rows ? fabs(static_cast<double>(x[u])) : 0.0;
      const bool take = v > best;
      bo = take ? base + static_cast<unsigned long long>(r) : bo;
      best = take ? v : best;
    }
  }
  // Partials an update's argmax epilogue left for the columns right of [k, n) (their offsets
  // continue this block's column-major numbering), shared out over the grid.
  for (int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x); i < ext_count;
       i += static_cast<int>(gridDim.x) * 256) {
    const double v = ext_v[i];
    const unsigned long long o = ext_o[i];
    if (better(v, o, best, bo)) {
      best = v;
      bo = o;
    }
  }
  block_best(best, bo);
  __shared__ bool last;
  if (threadIdx.x == 0) {
    partial_v[blockIdx.x] = best;
    partial_o[blockIdx.x] = bo;
    __threadfence();
    last = atomicAdd(&ctl->search_ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  best = 0.0;
  bo = 0ull;
  {
    // Every partial load is issued before any is used: one memory round trip, not one per
    // partial (at most 2048 partials, 8 per thread).
    constexpr int PER = 2048 / 256;
    double pv[PER];
    unsigned long long po[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int i = threadIdx.x + 256 * j;
      const int ic = i < static_cast<int>(gridDim.x) ? i : 0;
      pv[j] = __hip_atomic_load(&partial_v[ic], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      po[j] = __hip_atomic_load(&partial_o[ic], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const bool take = threadIdx.x + 256 * j < gridDim.x && better(pv[j], po[j], best, bo);
      bo = take ? po[j] : bo;
      best = take ? pv[j] : best;
    }
  }
  __syncthreads();
  block_best(best, bo);
  if (threadIdx.x == 0) {
    ctl->best_value = best;
    ctl->best_row = k + static_cast<int>(bo % static_cast<unsigned long long>(rows));
    ctl->best_col = k + static_cast<int>(bo / static_cast<unsigned long long>(rows));
    ctl->search_ticket = 0u;
  }
}

// Column-variant fallback search at step k (delayed_lu/delayed_dgetrf_column.cpp:30-45): the
// reference walks the columns right to left and stops at the first whose idamax over rows
// [k, m) is nonzero.  Here workgroup b of G examines columns n-1-b, n-1-b-G, ... (right to left)
// and stops at the first column that lies left of a nonzero column already found; a workgroup
// stops reading a column at its first nonzero entry (8 coalesced loads in flight per lane).  The
// rightmost nonzero column is kept with an atomic max, so the answer does not depend on
// dispatch order; G stays small (at most 256) so the closing ticket costs few atomics.  The
// last workgroup to finish then takes that column's largest |a| (lowest row on ties, as idamax)
// and clears the search state.  NaN compares false: a column of NaN and zeros counts as zero.
template <typename T>
__global__ __launch_bounds__(256) void column_search_kernel(const T *__restrict__ a, int lda,
                                                           int m, int n, int k, Control *ctl) {
  constexpr int U = 8;
  const int rows = m - k;
  for (int j = n - 1 - static_cast<int>(blockIdx.x); j > k; j -= static_cast<int>(gridDim.x)) {
    // Thread 0 reads the column found so far; the barrier's reduction makes the decision the
    // same for every thread of the workgroup (all of them meet the barriers below).
    const bool beaten =
        threadIdx.x == 0 &&
        __hip_atomic_load(&ctl->col_found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= j;
    if (__syncthreads_or(beaten)) break;
    const T *cp = a + static_cast<size_t>(k) + static_cast<size_t>(j) * lda;
    bool hit = false;
    for (int r0 = 0; r0 < rows; r0 += 256 * U) {
      T x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + 256 * u + static_cast<int>(threadIdx.x);
        x[u] = cp[r < rows ? r : 0];
      }
      bool nz = false;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + 256 * u + static_cast<int>(threadIdx.x);
        nz = nz || (r < rows && fabs(static_cast<double>(x[u])) > 0.0);
      }
      if (__syncthreads_or(nz)) {
        if (threadIdx.x == 0) atomicMax(&ctl->col_found, j);
        hit = true;
        break;
      }
    }
    if (hit) break;  // columns further left cannot win
  }
  __shared__ bool last;
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&ctl->search_ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  const int c = __hip_atomic_load(&ctl->col_found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  double best = 0.0;
  unsigned long long bo = 0ull;
  if (c > k) {
    const T *cp = a + static_cast<size_t>(k) + static_cast<size_t>(c) * lda;
    for (int r0 = 0; r0 < rows; r0 += 256 * U) {
      T x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + 256 * u + static_cast<int>(threadIdx.x);
        x[u] = cp[r < rows ? r : 0];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {  // a thread meets its rows in increasing order
        const int r = r0 + 256 * u + static_cast<int>(threadIdx.x);
        const double v = r < rows ? fabs(static_cast<double>(x[u])) : 0.0;
        const bool take = v > best;
        bo = take ? static_cast<unsigned long long>(r) : bo;
        best = take ? v : best;
      }
    }
  }
  block_best(best, bo);
  if (threadIdx.x == 0) {
    ctl->best_value = best;
    ctl->best_row = k + static_cast<int>(bo);
    ctl->best_col = c > k ? c : k;
    ctl->search_ticket = 0u;
    ctl->col_found = -1;
  }
}

// ---- Tall panels ---------------------------------------------------------------------------
// A panel with more rows than the register-resident panel kernel can hold is factored one
// column step at a time: the column argmax (active_search_kernel over the single column k),
// tall_pivot_kernel, tall_update_kernel.  Every launch is predicated on ctl->stop, so a zero
// column ends the panel on the device (stop = k) and the remaining launches return at once.
// Rows are interchanged across all n columns as the reference does (delayed_dgetrf.cpp:202-205),
// so the panel's swap list is the identity.

__global__ void tall_begin_kernel(Control *ctl, int pb, int pe) {
  if (threadIdx.x == 0) {
    ctl->stop = ctl->deferred ? pb : pe;  // a speculative restart after a deferral does nothing
    ctl->zero_column = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void tall_pivot_kernel(T *a, int lda, int n, int k, int *ipiv,
                                                        Control *ctl) {
  if (k >= ctl->stop) return;
  const double best = ctl->best_value;
  if (!(best > 0.0)) {  // zero (or NaN-only) column: the panel stops at k
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->stop = k;
      ctl->zero_column = 1;
    }
    return;
  }
  const int r = ctl->best_row;
  if (blockIdx.x == 0 && threadIdx.x == 0) ipiv[k] = r;
  if (r == k) return;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
    T *pk = a + static_cast<size_t>(k) + static_cast<size_t>(j) * lda;
    T *pr = a + static_cast<size_t>(r) + static_cast<size_t>(j) * lda;
    const T t = *pk;
    *pk = *pr;
    *pr = t;
  }
}

// Multipliers of column k (times the reciprocal of the pivot, delayed_dgetrf.cpp:221-228) and the
// rank-1 update of the panel's columns (k, pe) (:230-239), one row per thread.
template <typename T>
__global__ __launch_bounds__(256) void tall_update_kernel(T *a, int lda, int m, int k, int pe,
                                                         const Control *ctl) {
  if (k >= ctl->stop) return;
  const int i = k + 1 + static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (i >= m) return;
  const T *rowk = a + static_cast<size_t>(k);
  T *rowi = a + static_cast<size_t>(i);
  const T recip = T(1) / rowk[static_cast<size_t>(k) * lda];
  const T l = rowi[static_cast<size_t>(k) * lda] * recip;
  rowi[static_cast<size_t>(k) * lda] = l;
  for (int j = k + 1; j < pe; ++j)
    rowi[static_cast<size_t>(j) * lda] =
        fma(-l, rowk[static_cast<size_t>(j) * lda], rowi[static_cast<size_t>(j) * lda]);
}

// The panel's swap list: rows [pb, stop) in place (their interchanges are done), nothing
// displaced; then the host's copy of the control block.
__global__ void tall_finish_kernel(Control *ctl, int *swap_list, int nb, int pb, PanelDone *done,
                                   unsigned long long seq) {
  const int steps = ctl->stop - pb;
  for (int i = threadIdx.x; i < steps; i += blockDim.x) {
    swap_list[i] = pb + i;
    swap_list[2 * nb + i] = pb + i;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    swap_list[4 * nb] = steps;
    swap_list[4 * nb + 1] = 0;
    swap_list[4 * nb + 2] = 0;
    __threadfence();
    publish_done(ctl, done, seq);
  }
}

// Step k of a full-pivot fallback in one launch: record the decision, swap rows k and best_row
// over all n columns, then columns k and best_col over all m rows.  The two swaps meet only in
// the four elements of rows {k, r} x columns {k, c}; one thread does those, rows first, as the
// reference does, and the grid does every other element (disjoint sets, no ordering needed).
// An all-zero active block marks the deferral instead (info = k).
template <typename T>
__global__ __launch_bounds__(256) void fallback_apply_kernel(T *a, int lda, int m, int n, int k,
                                                            int *ipiv, int *jpiv, Control *ctl) {
  const double best = ctl->best_value;
  if (!(best > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->deferred = 1;
      ctl->stop = k;
    }
    return;
  }
  const int r = ctl->best_row, c = ctl->best_col;
  auto at = [&](int i, int j) -> T & {
    return a[static_cast<size_t>(i) + static_cast<size_t>(j) * lda];
  };
  auto swap = [](T &x, T &y) {
    const T t = x;
    x = y;
    y = t;
  };
  const int t = blockIdx.x * 256 + threadIdx.x, stride = gridDim.x * 256;
  if (t == 0) {
    ipiv[k] = r;
    jpiv[k] = c;
    ctl->fallbacks += 1;
    if (r != k) {
      swap(at(k, k), at(r, k));
      if (c != k) swap(at(k, c), at(r, c));
    }
    if (c != k) {
      swap(at(k, k), at(k, c));
      if (r != k) swap(at(r, k), at(r, c));
    }
  }
  if (r != k)
    for (int j = t; j < n; j += stride)
      if (j != k && j != c) swap(at(k, j), at(r, j));
  if (c != k)
    for (int i = t; i < m; i += stride)
      if (i != k && i != r) swap(at(i, k), at(i, c));
}

// Net row movement of a panel applied to columns [c0, c1): one wave per column gathers every
// source row, then writes every destination row.
template <typename T, int NB>
__global__ __launch_bounds__(256) void row_interchange_kernel(T *__restrict__ a, int lda,
                                                             int c0, int c1,
                                                             const int *__restrict__ swap_list,
                                                             Pred pred) {
  if (skipped(pred)) return;
  constexpr int PER = (2 * NB + 63) / 64;
  const int cnt = swap_list[4 * NB];
  const int lane = threadIdx.x & 63;
  const int col = c0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  const bool active = col < c1;
  T *cp = a + static_cast<size_t>(active ? col : c0) * lda;
  T vals[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int i = lane + 64 * s;
    vals[s] = (active && i < cnt) ? cp[swap_list[2 * NB + i]] : T(0);
  }
  __syncthreads();  // every read of this column has completed before any write
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    const int i = lane + 64 * s;
    if (active && i < cnt) cp[swap_list[i]] = vals[s];
  }
}

// Deferred left interchanges.  The panels' net row movements are kept in order (list e: row
// dst[k] receives the content of row src[k]); after the last panel one workgroup composes,
// from the last movement backward, the permutation P (row i of a block closed before movement
// e ends up with the content row P(i) holds now: P <- q_e o P, with q_e(dst) = src) and copies
// P out wherever a block needs it.  Each list moves a set of rows onto itself, so the update of
// P is two parallel phases through its inverse.  Rows below a block's last row are the only
// ones the later movements touch.  P and its inverse live in LDS (at most kDeferMaxRows rows),
// and the next movement's list is loaded while the current one is applied, so a movement costs
// two LDS phases and two barriers.
__global__ __launch_bounds__(1024) void left_compose_kernel(int m, const int *__restrict__ lists,
                                                           int stride, int nb, int events,
                                                           const int *__restrict__ snap_slot,
                                                           int *__restrict__ snaps) {
  __shared__ int P[kDeferMaxRows], Pinv[kDeferMaxRows];
  const int t = threadIdx.x;
  for (int i = t; i < m; i += blockDim.x) {
    P[i] = i;
    Pinv[i] = i;
  }
  // Entry k of a list is read by thread k (at most 2 nb <= 512 entries).
  const bool reader = t < 2 * nb;
  auto load = [&](int e, int &cnt, int &dst, int &src) {
    const int *L = lists + static_cast<size_t>(e - 1) * stride;
    cnt = L[4 * nb];
    dst = reader ? L[t] : 0;
    src = reader ? L[2 * nb + t] : 0;
  };
  int cnt = 0, dst = 0, src = 0;
  if (events >= 1) load(events, cnt, dst, src);
  __syncthreads();
  for (int e = events; e >= 1; --e) {
    int ncnt = 0, ndst = 0, nsrc = 0;
    if (e > 1) load(e - 1, ncnt, ndst, nsrc);
    const int slot = snap_slot[e];
    if (slot >= 0) {
      int *out = snaps + static_cast<size_t>(slot) * m;
      if ((m & 3) == 0) {  // 16-byte stores (m rows of ints, slot-aligned to 16 bytes)
        for (int i = t; i < m / 4; i += blockDim.x)
          reinterpret_cast<int4 *>(out)[i] = reinterpret_cast<const int4 *>(P)[i];
      } else {
        for (int i = t; i < m; i += blockDim.x) out[i] = P[i];
      }
    }
    const bool mine = t < cnt;
    const int at = mine ? Pinv[dst] : 0;
    __syncthreads();
    if (mine) {
      P[at] = src;
      Pinv[src] = at;
    }
    __syncthreads();
    cnt = ncnt;
    dst = ndst;
    src = nsrc;
  }
}

// One workgroup per column of the deferred blocks: rows [r0, m) of the column, staged in LDS,
// are written back in the order of the block's composed permutation.
template <typename T>
__global__ __launch_bounds__(1024) void left_gather_kernel(T *a, int lda, int m,
                                                          const int *__restrict__ col_slot,
                                                          const int *__restrict__ col_r0,
                                                          const int *__restrict__ snaps) {
  __shared__ T seg[kDeferMaxRows];
  const int c = blockIdx.x;
  const int slot = col_slot[c];
  if (slot < 0) return;
  const int r0 = col_r0[c];
  T *cp = a + static_cast<size_t>(c) * lda;
  for (int i = r0 + threadIdx.x; i < m; i += blockDim.x) seg[i - r0] = cp[i];
  __syncthreads();
  const int *Pc = snaps + static_cast<size_t>(slot) * m;
  for (int i = r0 + threadIdx.x; i < m; i += blockDim.x) cp[i] = seg[Pc[i] - r0];
}

__global__ void init_steps_kernel(int *ipiv, int *jpiv, double *margins, int steps) {
  for (int k = blockIdx.x * 256 + threadIdx.x; k < steps; k += gridDim.x * 256) {
    ipiv[k] = k;
    jpiv[k] = k;
    margins[k] = -1.0;
  }
}

}  // namespace

hipError_t launch_init_steps(int *ipiv, int *jpiv, double *margins, int steps, hipStream_t s) {
  const int g = (steps + 255) / 256 < 1024 ? (steps + 255) / 256 : 1024;
  hipLaunchKernelGGL(init_steps_kernel, dim3(g > 0 ? g : 1), dim3(256), 0, s, ipiv, jpiv, margins,
                     steps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_active_search(const T *a, int lda, int m, int n, int k, double *partial_v,
                                unsigned long long *partial_o, Control *ctl, hipStream_t s,
                                const double *ext_v, const unsigned long long *ext_o,
                                int ext_count) {
  // Four waves per workgroup, about four items per wave when the block is large; at most as
  // many workgroups as there are partial slots.
  const long long items = static_cast<long long>(n - k) * ((m - k + 511) / 512);
  const long long want = (items + 15) / 16;
  const int grid = want < 1 ? 1 : (want > 2048 ? 2048 : static_cast<int>(want));
  hipLaunchKernelGGL(active_search_kernel<T>, dim3(grid), dim3(256), 0, s, a, lda, m, n, k,
                     partial_v, partial_o, ctl, ext_v, ext_o, ext_count);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_column_search(const T *a, int lda, int m, int n, int k, Control *ctl,
                                hipStream_t s) {
  const int cols = n - k - 1;
  const int grid = cols < 1 ? 1 : (cols > 256 ? 256 : cols);
  hipLaunchKernelGGL(column_search_kernel<T>, dim3(grid), dim3(256), 0, s, a, lda, m, n, k, ctl);
  return hipGetLastError();
}

hipError_t launch_tall_begin(Control *ctl, int pb, int pe, hipStream_t s) {
  hipLaunchKernelGGL(tall_begin_kernel, dim3(1), dim3(64), 0, s, ctl, pb, pe);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tall_pivot(T *a, int lda, int n, int k, int *ipiv, Control *ctl, hipStream_t s) {
  const int g = (n + 255) / 256 < 512 ? (n + 255) / 256 : 512;
  hipLaunchKernelGGL(tall_pivot_kernel<T>, dim3(g > 0 ? g : 1), dim3(256), 0, s, a, lda, n, k,
                     ipiv, ctl);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tall_update(T *a, int lda, int m, int k, int pe, const Control *ctl,
                              hipStream_t s) {
  const int rows = m - k - 1;
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(tall_update_kernel<T>, dim3((rows + 255) / 256), dim3(256), 0, s, a, lda, m,
                     k, pe, ctl);
  return hipGetLastError();
}

hipError_t launch_tall_finish(Control *ctl, int *swap_list, int nb, int pb, PanelDone *done,
                              unsigned long long seq, hipStream_t s) {
  hipLaunchKernelGGL(tall_finish_kernel, dim3(1), dim3(256), 0, s, ctl, swap_list, nb, pb, done,
                     seq);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fallback_apply(T *a, int lda, int m, int n, int k, int *ipiv, int *jpiv,
                                 Control *ctl, hipStream_t s) {
  const int len = m > n ? m : n;
  const int g = (len + 255) / 256 < 1024 ? (len + 255) / 256 : 1024;
  hipLaunchKernelGGL(fallback_apply_kernel<T>, dim3(g > 0 ? g : 1), dim3(256), 0, s, a, lda, m, n,
                     k, ipiv, jpiv, ctl);
  return hipGetLastError();
}

hipError_t launch_left_compose(int m, const int *lists, int stride, int nb, int events,
                               const int *snap_slot, int *snaps, hipStream_t s) {
  if (m > kDeferMaxRows || 2 * nb > 1024) return hipErrorInvalidValue;  // the kernel's LDS / threads
  hipLaunchKernelGGL(left_compose_kernel, dim3(1), dim3(1024), 0, s, m, lists, stride, nb, events,
                     snap_slot, snaps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_left_gather(T *a, int lda, int m, int ncols, const int *col_slot,
                              const int *col_r0, const int *snaps, hipStream_t s) {
  if (ncols <= 0) return hipSuccess;
  if (m > kDeferMaxRows) return hipErrorInvalidValue;
  hipLaunchKernelGGL(left_gather_kernel<T>, dim3(ncols), dim3(1024), 0, s, a, lda, m, col_slot,
                     col_r0, snaps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_row_interchange(T *a, int lda, int c0, int c1, const int *swap_list, int nb,
                                  hipStream_t s, Pred pred) {
  if (c1 <= c0) return hipSuccess;
  const dim3 grid((c1 - c0 + 3) / 4), block(256);
  switch (nb) {
    case 64: hipLaunchKernelGGL((row_interchange_kernel<T, 64>), grid, block, 0, s, a, lda, c0, c1, swap_list, pred); break;
    case 128: hipLaunchKernelGGL((row_interchange_kernel<T, 128>), grid, block, 0, s, a, lda, c0, c1, swap_list, pred); break;
    case 256: hipLaunchKernelGGL((row_interchange_kernel<T, 256>), grid, block, 0, s, a, lda, c0, c1, swap_list, pred); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define DLU_PIVOT_INSTANTIATE(T)                                                                 \
  template hipError_t launch_active_search<T>(const T *, int, int, int, int, double *,          \
                                              unsigned long long *, Control *, hipStream_t,     \
                                              const double *, const unsigned long long *, int); \
  template hipError_t launch_fallback_apply<T>(T *, int, int, int, int, int *, int *, Control *, \
                                               hipStream_t);                                     \
  template hipError_t launch_row_interchange<T>(T *, int, int, int, const int *, int, hipStream_t, \
                                                Pred);                                          \
  template hipError_t launch_left_gather<T>(T *, int, int, int, const int *, const int *,         \
                                            const int *, hipStream_t);                            \
  template hipError_t launch_column_search<T>(const T *, int, int, int, int, Control *,          \
                                              hipStream_t);                                     \
  template hipError_t launch_tall_pivot<T>(T *, int, int, int, int *, Control *, hipStream_t);   \
  template hipError_t launch_tall_update<T>(T *, int, int, int, int, const Control *, hipStream_t);
DLU_PIVOT_INSTANTIATE(float)
DLU_PIVOT_INSTANTIATE(double)
#undef DLU_PIVOT_INSTANTIATE

}  // namespace dlu
