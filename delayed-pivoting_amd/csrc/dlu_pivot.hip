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

__device__ __forceinline__ bool better(double v, unsigned long long o, double bv,
                                       unsigned long long bo) {
  return v > bv || (v == bv && o < bo);
}

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
                                                           Control *ctl) {
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
                                                             const int *__restrict__ swap_list) {
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
                                unsigned long long *partial_o, Control *ctl, hipStream_t s) {
  // Four waves per workgroup, about four items per wave when the block is large; at most as
  // many workgroups as there are partial slots.
  const long long items = static_cast<long long>(n - k) * ((m - k + 511) / 512);
  const long long want = (items + 15) / 16;
  const int grid = want < 1 ? 1 : (want > 2048 ? 2048 : static_cast<int>(want));
  hipLaunchKernelGGL(active_search_kernel<T>, dim3(grid), dim3(256), 0, s, a, lda, m, n, k,
                     partial_v, partial_o, ctl);
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

template <typename T>
hipError_t launch_row_interchange(T *a, int lda, int c0, int c1, const int *swap_list, int nb,
                                  hipStream_t s) {
  if (c1 <= c0) return hipSuccess;
  const dim3 grid((c1 - c0 + 3) / 4), block(256);
  switch (nb) {
    case 64: hipLaunchKernelGGL((row_interchange_kernel<T, 64>), grid, block, 0, s, a, lda, c0, c1, swap_list); break;
    case 128: hipLaunchKernelGGL((row_interchange_kernel<T, 128>), grid, block, 0, s, a, lda, c0, c1, swap_list); break;
    case 256: hipLaunchKernelGGL((row_interchange_kernel<T, 256>), grid, block, 0, s, a, lda, c0, c1, swap_list); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define DLU_PIVOT_INSTANTIATE(T)                                                                 \
  template hipError_t launch_active_search<T>(const T *, int, int, int, int, double *,          \
                                              unsigned long long *, Control *, hipStream_t);    \
  template hipError_t launch_fallback_apply<T>(T *, int, int, int, int, int *, int *, Control *, \
                                               hipStream_t);                                     \
  template hipError_t launch_row_interchange<T>(T *, int, int, int, const int *, int, hipStream_t);
DLU_PIVOT_INSTANTIATE(float)
DLU_PIVOT_INSTANTIATE(double)
#undef DLU_PIVOT_INSTANTIATE

}  // namespace dlu
