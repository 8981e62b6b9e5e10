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

// One workgroup per group of columns; the last workgroup to finish reduces the partials.
// Magnitudes are compared in double (exact for float entries).
template <typename T>
__global__ __launch_bounds__(256) void active_search_kernel(const T *__restrict__ a, int lda,
                                                           int m, int n, int k,
                                                           double *partial_v,
                                                           unsigned long long *partial_o,
                                                           Control *ctl) {
  const int rows = m - k, cols = n - k;
  double best = 0.0;  // a zero entry never displaces the initial candidate
  unsigned long long bo = 0ull;
  for (int col = blockIdx.x; col < cols; col += gridDim.x) {
    const T *cp = a + static_cast<size_t>(k) + static_cast<size_t>(k + col) * lda;
    const unsigned long long base = static_cast<unsigned long long>(col) * rows;
    for (int r = threadIdx.x; r < rows; r += 256) {
      const double x = fabs(static_cast<double>(cp[r]));
      const unsigned long long o = base + r;
      if (better(x, o, best, bo)) {
        best = x;
        bo = o;
      }
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
  for (int i = threadIdx.x; i < static_cast<int>(gridDim.x); i += 256) {
    const double v = __hip_atomic_load(&partial_v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long o =
        __hip_atomic_load(&partial_o[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (better(v, o, best, bo)) {
      best = v;
      bo = o;
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

// Step k of a full-pivot fallback, part 1: record the decision and swap rows k and best_row over
// all n columns.  An all-zero active block marks the deferral instead (info = k).
template <typename T>
__global__ __launch_bounds__(256) void fallback_rows_kernel(T *a, int lda, int n, int k,
                                                           int *ipiv, int *jpiv, Control *ctl) {
  const double best = ctl->best_value;
  if (!(best > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->deferred = 1;
      ctl->stop = k;
    }
    return;
  }
  const int r = ctl->best_row;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ipiv[k] = r;
    jpiv[k] = ctl->best_col;
    ctl->fallbacks += 1;
  }
  if (r == k) return;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
    T *x = a + static_cast<size_t>(k) + static_cast<size_t>(j) * lda;
    T *y = a + static_cast<size_t>(r) + static_cast<size_t>(j) * lda;
    const T t = *x;
    *x = *y;
    *y = t;
  }
}

// Part 2: swap columns k and best_col over all m rows (after the row swap).
template <typename T>
__global__ __launch_bounds__(256) void fallback_cols_kernel(T *a, int lda, int m, int k,
                                                           const Control *ctl) {
  if (ctl->deferred || !(ctl->best_value > 0.0)) return;
  const int c = ctl->best_col;
  if (c == k) return;
  T *x = a + static_cast<size_t>(k) * lda;
  T *y = a + static_cast<size_t>(c) * lda;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
    const T t = x[i];
    x[i] = y[i];
    y[i] = t;
  }
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
  const int cols = n - k;
  const int grid = cols < 2048 ? (cols > 0 ? cols : 1) : 2048;
  hipLaunchKernelGGL(active_search_kernel<T>, dim3(grid), dim3(256), 0, s, a, lda, m, n, k,
                     partial_v, partial_o, ctl);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fallback_apply(T *a, int lda, int m, int n, int k, int *ipiv, int *jpiv,
                                 Control *ctl, hipStream_t s) {
  const int gr = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  hipLaunchKernelGGL(fallback_rows_kernel<T>, dim3(gr > 0 ? gr : 1), dim3(256), 0, s, a, lda, n,
                     k, ipiv, jpiv, ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int gc = (m + 255) / 256 < 1024 ? (m + 255) / 256 : 1024;
  hipLaunchKernelGGL(fallback_cols_kernel<T>, dim3(gc > 0 ? gc : 1), dim3(256), 0, s, a, lda, m,
                     k, ctl);
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
