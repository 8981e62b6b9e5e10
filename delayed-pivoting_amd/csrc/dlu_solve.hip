// Solve-side kernels on gfx950: the triangular solves with the delayed-pivoting factors, the
// permutation gathers around them, the FP64 residual and norms of iterative refinement, and the
// precision conversions of the mixed-precision solver.
//
// The factorization (delayed_lu/delayed_dgetrf.cpp:129-251) leaves P*A*Q = L*U with P and Q the
// products of the recorded row and column swaps.  A*x = b is then solved as
//   y = U^{-1} L^{-1} (P b),   x = Q y,
// the row swaps applied to b in step order and the column swaps to y in reverse step order.
// The reference has no solve of its own; its example checks the factors by reconstruction
// (delayed_lu/example.cpp:43-126).
#include "dlu_internal.hpp"

namespace dlu {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void gather_rows_kernel(const T *__restrict__ src, int lds,
                                                         T *__restrict__ dst, int ldd,
                                                         const int *__restrict__ perm, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) dst[static_cast<size_t>(i) + static_cast<size_t>(j) * ldd] =
      src[static_cast<size_t>(perm[i]) + static_cast<size_t>(j) * lds];
}

// ---------------------------------------------------------------------------------------------
// X(r0 : r0+w, c) := D^{-1} X(r0 : r0+w, c) for a strip of 16 right-hand sides c, with D the
// w x w diagonal block of the factor at (r0, r0): unit lower (L) or non-unit upper (U), w <= 128.
// Every workgroup forms the inverses of D's 16 x 16 diagonal sub-blocks itself (one thread per
// column, by substitution), then runs block substitution on MFMA: X_b := inv(D_bb) X_b, and the
// rows still to be solved (below for L, above for U) lose D(rows, b) X_b, with that slice of D
// staged in LDS.  Rows beyond w are zero-padded and never written back.
// ---------------------------------------------------------------------------------------------
template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void block_trsv_kernel(const T *__restrict__ lu, int lda,
                                                        int r0, int w, T *__restrict__ x,
                                                        int ldx, int nrhs) {
  typedef Mfma<T> MF;
  constexpr int NB = 128, NS = 16;
  __shared__ T X[NB][NS + 2];
  __shared__ T invD[NB / 16][16][17];
  __shared__ T Ds[NB][17];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int cs = blockIdx.x * NS;
  const int ncol = (nrhs - cs) < NS ? (nrhs - cs) : NS;
  const int nblk = (w + 15) / 16;
  const int rows = nblk * 16;
  const T *D = lu + static_cast<size_t>(r0) + static_cast<size_t>(r0) * lda;

  for (int idx = t; idx < rows * NS; idx += 256) {
    const int i = idx % rows, col = idx / rows;
    X[i][col] = (i < w && col < ncol)
                    ? x[static_cast<size_t>(r0 + i) + static_cast<size_t>(cs + col) * ldx]
                    : T(0);
  }
  if (t < 16 * nblk) {
    const int blk = t >> 4, col = t & 15;
    const int b0 = blk * 16;
    const int nr = (w - b0) < 16 ? (w - b0) : 16;
    T v[16];
    if (!UPPER) {  // unit lower: forward substitution for column `col` of the inverse
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        T sum = (r == col) ? T(1) : T(0);
        if (r > col && r < nr) {
#pragma unroll
          for (int q = 0; q < 16; ++q)
            if (q >= col && q < r)
              sum -= D[static_cast<size_t>(b0 + r) + static_cast<size_t>(b0 + q) * lda] * v[q];
        }
        v[r] = (r < nr) ? sum : T(0);
      }
    } else {  // non-unit upper: back substitution
#pragma unroll
      for (int r = 15; r >= 0; --r) {
        T val = T(0);
        if (r < nr && col < nr && r <= col) {
          T sum = (r == col) ? T(1) : T(0);
#pragma unroll
          for (int q = 0; q < 16; ++q)
            if (q > r && q <= col)
              sum -= D[static_cast<size_t>(b0 + r) + static_cast<size_t>(b0 + q) * lda] * v[q];
          val = sum / D[static_cast<size_t>(b0 + r) + static_cast<size_t>(b0 + r) * lda];
        }
        v[r] = val;
      }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) invD[blk][r][col] = v[r];
  }
  __syncthreads();

  const int lr = lane & 15, lk = lane >> 4;
  for (int step = 0; step < nblk; ++step) {
    const int blk = UPPER ? nblk - 1 - step : step;
    const int b0 = blk * 16;
    const int other = UPPER ? b0 : rows - (b0 + 16);  // rows still to be solved
    const int obase = UPPER ? 0 : b0 + 16;
    for (int idx = t; idx < other * 16; idx += 256) {
      const int rr = idx % other, q = idx / other;
      const int row = obase + rr, colq = b0 + q;
      Ds[rr][q] = (row < w && colq < w)
                      ? D[static_cast<size_t>(row) + static_cast<size_t>(colq) * lda]
                      : T(0);
    }
    if (wave == 0) {
      typename MF::acc_t d = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int k4 = 0; k4 < 16; k4 += 4)
        d = MF::mma(invD[blk][lr][k4 + lk], X[b0 + k4 + lk][lr], d);
#pragma unroll
      for (int r = 0; r < 4; ++r) X[b0 + MF::acc_row(lane, r)][lr] = d[r];
    }
    __syncthreads();
    for (int tb = wave; tb < other / 16; tb += 4) {
      const int rt = obase + 16 * tb;
      typename MF::acc_t d;
#pragma unroll
      for (int r = 0; r < 4; ++r) d[r] = X[rt + MF::acc_row(lane, r)][lr];
#pragma unroll
      for (int k4 = 0; k4 < 16; k4 += 4)
        d = MF::mma(-Ds[16 * tb + lr][k4 + lk], X[b0 + k4 + lk][lr], d);
#pragma unroll
      for (int r = 0; r < 4; ++r) X[rt + MF::acc_row(lane, r)][lr] = d[r];
    }
    __syncthreads();
  }

  for (int idx = t; idx < w * ncol; idx += 256) {
    const int i = idx % w, col = idx / w;
    x[static_cast<size_t>(r0 + i) + static_cast<size_t>(cs + col) * ldx] = X[i][col];
  }
}

// ---------------------------------------------------------------------------------------------
// R = B - A X, FP64, for a 16 x 16 tile of R per workgroup.  The four waves take interleaved
// k-quarters of the product on v_mfma_f64_16x16x4_f64 and combine their partial tiles in LDS in
// a fixed order, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void residual_kernel(const double *__restrict__ a, int lda,
                                                      const double *__restrict__ xm, int ldx,
                                                      const double *__restrict__ b, int ldb,
                                                      double *__restrict__ r, int ldr, int n,
                                                      int nrhs) {
  typedef Mfma<double> MF;
  __shared__ double part[4][64][4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
  const int lr = lane & 15, lk = lane >> 4;
  const int row = r0 + lr, col = c0 + lr;
  const bool row_ok = row < n, col_ok = col < nrhs;
  const double *ap = a + static_cast<size_t>(row_ok ? row : 0);
  const double *xp = xm + static_cast<size_t>(col_ok ? col : 0) * ldx;
  MF::acc_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int k = 4 * wave; k < n; k += 16) {
    const int kk = k + lk;
    const bool ok = kk < n;
    const double av = (row_ok && ok) ? ap[static_cast<size_t>(kk) * lda] : 0.0;
    const double xv = (col_ok && ok) ? xp[kk] : 0.0;
    acc = MF::mma(av, xv, acc);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) part[wave][lane][q] = acc[q];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double s = ((part[0][lane][q] + part[1][lane][q]) + part[2][lane][q]) +
                       part[3][lane][q];
      const int orow = r0 + MF::acc_row(lane, q);
      if (orow < n && col_ok) {
        const size_t ob = static_cast<size_t>(orow) + static_cast<size_t>(col) * ldb;
        r[static_cast<size_t>(orow) + static_cast<size_t>(col) * ldr] = b[ob] - s;
      }
    }
  }
}

// Max of non-negative doubles through their bit patterns (which order like the values).
__device__ __forceinline__ void atomic_max_nonneg(double *out, double v) {
  atomicMax(reinterpret_cast<unsigned long long *>(out),
            static_cast<unsigned long long>(__double_as_longlong(v)));
}

__device__ __forceinline__ double block_max(double v) {
  __shared__ double sm[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// Row sums of |A| over column chunk blockIdx.y (fixed summation order within a chunk).
constexpr int kNormChunks = 32;
__global__ __launch_bounds__(256) void row_abs_sum_kernel(const double *__restrict__ a, int lda,
                                                         int n, double *__restrict__ part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int per = (n + kNormChunks - 1) / kNormChunks;
  const int j0 = blockIdx.y * per, j1 = (j0 + per) < n ? (j0 + per) : n;
  if (i >= n) return;
  double s = 0.0;
  for (int j = j0; j < j1; ++j) s += fabs(a[static_cast<size_t>(i) + static_cast<size_t>(j) * lda]);
  part[static_cast<size_t>(blockIdx.y) * n + i] = s;
}
__global__ __launch_bounds__(256) void row_sum_max_kernel(const double *__restrict__ part, int n,
                                                         double *out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  double s = 0.0;
  if (i < n)
    for (int c = 0; c < kNormChunks; ++c) s += part[static_cast<size_t>(c) * n + i];
  s = block_max(s);
  if (threadIdx.x == 0) atomic_max_nonneg(out, s);
}

__global__ __launch_bounds__(256) void col_abs_max_kernel(const double *__restrict__ x, int ldx,
                                                         int n, double *out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  const double v = i < n ? fabs(x[static_cast<size_t>(i) + static_cast<size_t>(j) * ldx]) : 0.0;
  const double m = block_max(v);
  if (threadIdx.x == 0) atomic_max_nonneg(out + j, m);
}

__global__ __launch_bounds__(256) void to_float_kernel(const double *__restrict__ src, int lds,
                                                      float *__restrict__ dst, int ldd, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const size_t j = blockIdx.y;
  if (i < n) dst[i + j * ldd] = static_cast<float>(src[i + j * lds]);
}

template <typename T>
__global__ __launch_bounds__(256) void add_correction_kernel(double *__restrict__ x, int ldx,
                                                            const T *__restrict__ d, int ldd,
                                                            int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const size_t j = blockIdx.y;
  if (i < n) x[i + j * ldx] += static_cast<double>(d[i + j * ldd]);
}

}  // namespace

template <typename T>
hipError_t launch_gather_rows(const T *src, int lds, T *dst, int ldd, const int *perm, int n,
                              int nrhs, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_rows_kernel<T>, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, src,
                     lds, dst, ldd, perm, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_block_trsv(const T *lu, int lda, int r0, int w, T *x, int ldx, int nrhs,
                             bool upper, hipStream_t s) {
  if (w <= 0 || nrhs <= 0) return hipSuccess;
  if (w > 128) return hipErrorInvalidValue;
  const dim3 grid((nrhs + 15) / 16), block(256);
  if (upper)
    hipLaunchKernelGGL((block_trsv_kernel<T, true>), grid, block, 0, s, lu, lda, r0, w, x, ldx,
                       nrhs);
  else
    hipLaunchKernelGGL((block_trsv_kernel<T, false>), grid, block, 0, s, lu, lda, r0, w, x, ldx,
                       nrhs);
  return hipGetLastError();
}

hipError_t launch_residual(const double *a, int lda, const double *x, int ldx, const double *b,
                           int ldb, double *r, int ldr, int n, int nrhs, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(residual_kernel, dim3((n + 15) / 16, (nrhs + 15) / 16), dim3(256), 0, s, a,
                     lda, x, ldx, b, ldb, r, ldr, n, nrhs);
  return hipGetLastError();
}

hipError_t launch_norm_inf_matrix(const double *a, int lda, int n, double *work, double *out,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(row_abs_sum_kernel, dim3((n + 255) / 256, kNormChunks), dim3(256), 0, s, a,
                     lda, n, work);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(row_sum_max_kernel, dim3((n + 255) / 256), dim3(256), 0, s, work, n, out);
  return hipGetLastError();
}

hipError_t launch_norm_inf_columns(const double *x, int ldx, int n, int nrhs, double *out,
                                   hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(col_abs_max_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, x, ldx, n,
                     out);
  return hipGetLastError();
}

hipError_t launch_to_float(const double *src, int lds, float *dst, int ldd, int n, int ncol,
                           hipStream_t s) {
  if (n <= 0 || ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(to_float_kernel, dim3((n + 255) / 256, ncol), dim3(256), 0, s, src, lds, dst,
                     ldd, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_add_correction(double *x, int ldx, const T *d, int ldd, int n, int ncol,
                                 hipStream_t s) {
  if (n <= 0 || ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(add_correction_kernel<T>, dim3((n + 255) / 256, ncol), dim3(256), 0, s, x,
                     ldx, d, ldd, n);
  return hipGetLastError();
}
template hipError_t launch_add_correction<float>(double *, int, const float *, int, int, int,
                                                 hipStream_t);
template hipError_t launch_add_correction<double>(double *, int, const double *, int, int, int,
                                                  hipStream_t);

template hipError_t launch_gather_rows<float>(const float *, int, float *, int, const int *, int,
                                              int, hipStream_t);
template hipError_t launch_gather_rows<double>(const double *, int, double *, int, const int *,
                                               int, int, hipStream_t);
template hipError_t launch_block_trsv<float>(const float *, int, int, int, float *, int, int, bool,
                                             hipStream_t);
template hipError_t launch_block_trsv<double>(const double *, int, int, int, double *, int, int,
                                              bool, hipStream_t);

}  // namespace dlu
