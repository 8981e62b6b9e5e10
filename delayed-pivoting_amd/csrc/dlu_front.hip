// Frontal-matrix steps around the delayed-pivoting LU on gfx950: the front update that follows
// the factorization of the fully summed block, and the extend-add of a contribution block into
// its parent front.
//
// Reference behaviour restated here (doc/（非）对称稠密矩阵的延迟主元LU方法.md):
//   Step 1: delayed LU of the fully summed block; E eliminated, D delayed        :20-226
//   Step 2.1: U_EU = L_EE^{-1} A_EU,  L_UE = A_UE U_EE^{-1}                       :300-307
//   Step 2.2: A_DU -= L_DE U_EU,  A_UD -= L_UE U_ED                               :308-313
//   Step 2.3: A_UU -= L_UE U_EU                                                   :314-317
//   Step 3: the contribution block [[S_DD, S_DU], [S_UD, S_UU]] with its global row and column
//           indices goes to the parent, which scatter-adds it by those indices
//           (extend-add); delayed variables join the parent's candidate set       :319-440
// The front is an nf x nf column-major matrix whose leading ns x ns block is fully summed.  The
// row interchanges of Step 1 span the fully summed rows across the whole front width and its
// column interchanges the fully summed columns down the whole front height (as the reference
// LU's interchanges span the whole matrix, delayed_lu/delayed_dgetrf.cpp:202-209).
#include "dlu.h"
#include "dlu_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace dlu {
namespace {

// The row interchanges of a swap history (step k swaps rows k and ipiv[k], k < e) replayed on
// columns [c0, c1): one thread per column, in step order.
__global__ __launch_bounds__(256) void replay_row_swaps_kernel(double *a, int lda, int c0, int c1,
                                                              const int *__restrict__ ipiv,
                                                              int e) {
  const int col = c0 + static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (col >= c1) return;
  double *cp = a + static_cast<size_t>(col) * lda;
  for (int k = 0; k < e; ++k) {
    const int r = ipiv[k];
    if (r != k) {
      const double t = cp[k];
      cp[k] = cp[r];
      cp[r] = t;
    }
  }
}

// The column interchanges of a swap history replayed on rows [r0, r1): one thread per row (the
// threads of a wave touch consecutive rows of a column together).
__global__ __launch_bounds__(256) void replay_col_swaps_kernel(double *a, int lda, int r0, int r1,
                                                              const int *__restrict__ jpiv,
                                                              int e) {
  const int row = r0 + static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (row >= r1) return;
  for (int k = 0; k < e; ++k) {
    const int c = jpiv[k];
    if (c != k) {
      double *x = a + row + static_cast<size_t>(k) * lda;
      double *y = a + row + static_cast<size_t>(c) * lda;
      const double t = *x;
      *x = *y;
      *y = t;
    }
  }
}

// X := X U^{-1} for a block of w <= 64 columns: U is w x w upper triangular (non-unit), X has
// `rows` rows.  U goes to LDS; each thread solves one row of X left to right in registers.
constexpr int kRightBlock = 64;
__global__ __launch_bounds__(256) void trsm_right_upper_kernel(double *x, int ldx, int rows,
                                                              const double *__restrict__ u,
                                                              int ldu, int w) {
  __shared__ double su[kRightBlock][kRightBlock + 1];
  for (int i = threadIdx.x; i < w * w; i += 256) {
    const int r = i % w, c = i / w;
    su[r][c] = r <= c ? u[r + static_cast<size_t>(c) * ldu] : 0.0;
  }
  __syncthreads();
  const int row = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (row >= rows) return;
  double v[kRightBlock];
#pragma unroll
  for (int c = 0; c < kRightBlock; ++c)
    v[c] = c < w ? x[row + static_cast<size_t>(c) * ldx] : 0.0;
#pragma unroll
  for (int c = 0; c < kRightBlock; ++c) {
    if (c < w) {
      double s = v[c];
#pragma unroll
      for (int i = 0; i < c; ++i) s = fma(-v[i], su[i][c], s);
      v[c] = s / su[c][c];
    }
  }
#pragma unroll
  for (int c = 0; c < kRightBlock; ++c)
    if (c < w) x[row + static_cast<size_t>(c) * ldx] = v[c];
}

// parent(map_r[i], map_c[j]) += child(i, j).  Within one launch the entries of each map are
// distinct, so no two threads touch one element.
__global__ __launch_bounds__(256) void extend_add_kernel(int rows, int cols,
                                                        const double *__restrict__ child,
                                                        int ldc, const int *__restrict__ map_r,
                                                        const int *__restrict__ map_c,
                                                        double *parent, int ldp) {
  const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (i >= rows) return;
  const int pr = map_r[i];
  for (int j = blockIdx.y; j < cols; j += gridDim.y) {
    const int pc = map_c[j];
    parent[pr + static_cast<size_t>(pc) * ldp] += child[i + static_cast<size_t>(j) * ldc];
  }
}

inline double *at(double *a, int lda, int i, int j) {
  return a + static_cast<size_t>(i) + static_cast<size_t>(j) * static_cast<size_t>(lda);
}

// Step 2 of a front whose fully summed block has been factored (e eliminated steps; ipiv/jpiv
// are device copies of its swap histories).  Everything runs in order on stream s.
hipError_t front_update(int nf, int ns, int e, double *f, int ldf, const int *d_ipiv,
                        const int *d_jpiv, hipStream_t s) {
  const int nu = nf - ns;
  if (nu <= 0 || e <= 0) return hipSuccess;
  hipError_t err = hipSuccess;
  // The fully summed rows' interchanges across the update columns, and the fully summed columns'
  // interchanges down the update rows.
  hipLaunchKernelGGL(replay_row_swaps_kernel, dim3((nu + 255) / 256), dim3(256), 0, s, f, ldf, ns,
                     nf, d_ipiv, e);
  hipLaunchKernelGGL(replay_col_swaps_kernel, dim3((nu + 255) / 256), dim3(256), 0, s, f, ldf, ns,
                     nf, d_jpiv, e);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  // Step 2.1: U_EU = L_EE^{-1} A_EU, blocked by 128 rows (the swap+TRSM kernel solves a block in
  // place, the trailing-update GEMM carries it down).
  constexpr int kLeft = 128;
  for (int b = 0; b < e && err == hipSuccess; b += kLeft) {
    const int w = std::min(kLeft, e - b);
    err = launch_swap_trsm<double>(f, ldf, b, w, ns, nf, nullptr, kLeft, s);
    if (err == hipSuccess && b + w < e)
      err = launch_gemm_sub<double>(at(f, ldf, b + w, ns), ldf, at(f, ldf, b + w, b), ldf,
                                    at(f, ldf, b, ns), ldf, e - b - w, nu, w, s);
  }
  // L_UE = A_UE U_EE^{-1}, left-looking by 64 columns.
  for (int j = 0; j < e && err == hipSuccess; j += kRightBlock) {
    const int w = std::min(kRightBlock, e - j);
    if (j > 0)
      err = launch_gemm_sub<double>(at(f, ldf, ns, j), ldf, at(f, ldf, ns, 0), ldf,
                                    at(f, ldf, 0, j), ldf, nu, w, j, s);
    if (err == hipSuccess) {
      hipLaunchKernelGGL(trsm_right_upper_kernel, dim3((nu + 255) / 256), dim3(256), 0, s,
                         at(f, ldf, ns, j), ldf, nu, at(f, ldf, j, j), ldf, w);
      err = hipGetLastError();
    }
  }
  // Steps 2.2 and 2.3: [A_DU; A_UU] -= [L_DE; L_UE] U_EU, then A_UD -= L_UE U_ED.
  if (err == hipSuccess)
    err = launch_gemm_sub<double>(at(f, ldf, e, ns), ldf, at(f, ldf, e, 0), ldf,
                                  at(f, ldf, 0, ns), ldf, nf - e, nu, e, s);
  if (err == hipSuccess && ns > e)
    err = launch_gemm_sub<double>(at(f, ldf, ns, e), ldf, at(f, ldf, ns, 0), ldf,
                                  at(f, ldf, 0, e), ldf, nu, ns - e, e, s);
  return err;
}

}  // namespace
}  // namespace dlu

extern "C" {

dlu_status dlu_front_factor_device(int nf, int ns, double *d_f, int ldf, int *ipiv, int *jpiv,
                                   int *info, void *stream) {
  if (info == nullptr) return DLU_INVALID_ARGUMENT;
  *info = -1;
  if (nf < 0 || ns < 0 || ns > nf || ldf < std::max(1, nf) || (nf > 0 && d_f == nullptr) ||
      (ns > 0 && (ipiv == nullptr || jpiv == nullptr)))
    return DLU_INVALID_ARGUMENT;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (ns == 0) {
    *info = 0;
    return DLU_SUCCESS;
  }
  // Step 1: the delayed LU of the fully summed block, in place (leading dimension ldf).
  const dlu_status st =
      dlu::getrf_device<double>(ns, ns, d_f, ldf, ipiv, jpiv, info, s, nullptr, nullptr, nullptr);
  if (st != DLU_SUCCESS) return st;
  const int e = *info;
  if (nf == ns || e == 0) return DLU_SUCCESS;
  int *d_piv = nullptr;
  hipError_t err = hipMallocAsync(reinterpret_cast<void **>(&d_piv), 2 * sizeof(int) * ns, s);
  if (err != hipSuccess) {
    *info = -1;
    return DLU_ALLOCATION_FAILED;
  }
  err = hipMemcpyAsync(d_piv, ipiv, sizeof(int) * ns, hipMemcpyHostToDevice, s);
  if (err == hipSuccess)
    err = hipMemcpyAsync(d_piv + ns, jpiv, sizeof(int) * ns, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = dlu::front_update(nf, ns, e, d_f, ldf, d_piv, d_piv + ns, s);
  const hipError_t e2 = hipFreeAsync(d_piv, s);
  if (err == hipSuccess) err = e2;
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err != hipSuccess) {
    *info = -1;
    return dlu::to_status(err);
  }
  return DLU_SUCCESS;
}

dlu_status dlu_extend_add_device(int rows, int cols, const double *d_child, int ldc,
                                 const int *d_map_r, const int *d_map_c, double *d_parent,
                                 int ldp, void *stream) {
  if (rows < 0 || cols < 0 || ldc < std::max(1, rows) || ldp < 1) return DLU_INVALID_ARGUMENT;
  if (rows == 0 || cols == 0) return DLU_SUCCESS;
  if (!d_child || !d_map_r || !d_map_c || !d_parent) return DLU_INVALID_ARGUMENT;
  hipLaunchKernelGGL(dlu::extend_add_kernel, dim3((rows + 255) / 256, std::min(cols, 4096)),
                     dim3(256), 0,
                     static_cast<hipStream_t>(stream), rows, cols, d_child, ldc, d_map_r, d_map_c,
                     d_parent, ldp);
  return dlu::to_status(hipGetLastError());
}

}  // extern "C"
