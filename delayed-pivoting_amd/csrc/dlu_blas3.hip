// Level-3 kernels of the delayed-pivoting LU on gfx950: the FP64 MFMA trailing update and the
// triangular solve that forms the U block row.
//
// Reference behaviour restated here (delayed_lu/delayed_dgetrf.cpp):
//   finish_panel: U12 = L11^{-1} A12 (unit lower), A22 -= L21 U12        :61-85
//   materialize_partial_panel: the same two steps for an accepted prefix  :101-125
//   row interchanges span the columns right of the panel                  :202-205
#include "dlu_internal.hpp"

namespace dlu {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------------
// C(M x N) -= A(M x K) * B(K x N), column-major, one leading dimension for all three operands.
//
// Workgroup tile BM x BN, 256 threads = 4 waves in a 2 x 2 grid, each wave owns a
// (BM/2) x (BN/2) block made of 16 x 16 v_mfma_f64_16x16x4_f64 tiles.  K advances in chunks of
// 16 staged through double-buffered LDS; the next chunk's global loads are in flight while the
// current chunk feeds the MFMA pipe.
//
// v_mfma_f64_16x16x4_f64 operand layout (lane l of 64):
//   A (16 x 4): one double, row l % 16, k l / 16
//   B (4 x 16): one double, k l / 16, column l % 16
//   D (16 x 16): four doubles r = 0..3, row l / 16 + 4 * r, column l % 16
// ---------------------------------------------------------------------------------------------
template <int BM, int BN>
__global__ __launch_bounds__(256) void gemm_sub_kernel(double *__restrict__ C,
                                                      const double *__restrict__ A,
                                                      const double *__restrict__ B, int lda,
                                                      int M, int N, int K, int tiles_m,
                                                      int tiles_n) {
  constexpr int BK = 16;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int BKP = BK + 2;  // padded k stride of the B tile (16-byte aligned rows)
  __shared__ double As[2][BK][BM];
  __shared__ double Bs[2][BN][BKP];

  // Tile schedule.  Workgroups are dealt to the 8 XCDs round-robin; give each XCD a contiguous
  // range of tiles, then walk that range in groups of 8 tile rows so neighbours share L2 data.
  const int nt = tiles_m * tiles_n;
  const int b = blockIdx.x;
  const int q = nt / 8, rem = nt % 8, x = b % 8;
  const int tid = x * q + (x < rem ? x : rem) + b / 8;
  constexpr int GROUP = 8;
  const int in_group = GROUP * tiles_n;
  const int first_m = (tid / in_group) * GROUP;
  const int gsize = (tiles_m - first_m) < GROUP ? (tiles_m - first_m) : GROUP;
  const int tm = first_m + (tid % in_group) % gsize;
  const int tn = (tid % in_group) / gsize;
  const int row0 = tm * BM, col0 = tn * BN;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  constexpr int A_PER = BM * BK / 256;  // doubles loaded per thread
  constexpr int A_TPC = BM / A_PER;     // threads per A column
  constexpr int B_PER = BN * BK / 256;
  constexpr int B_TPC = BK / B_PER;
  const int a_k = threadIdx.x / A_TPC, a_r = (threadIdx.x % A_TPC) * A_PER;
  const int b_n = threadIdx.x / B_TPC, b_k = (threadIdx.x % B_TPC) * B_PER;

  double ra[A_PER], rb[B_PER];
  auto load = [&](int k0) {
    const int ka = k0 + a_k;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int row = row0 + a_r + i;
      ra[i] = (row < M && ka < K) ? A[static_cast<size_t>(row) + static_cast<size_t>(ka) * lda] : 0.0;
    }
    const int col = col0 + b_n;
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int kb = k0 + b_k + i;
      rb[i] = (kb < K && col < N) ? B[static_cast<size_t>(kb) + static_cast<size_t>(col) * lda] : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_PER; ++i) As[buf][a_k][a_r + i] = ra[i];
#pragma unroll
    for (int i = 0; i < B_PER; ++i) Bs[buf][b_n][b_k + i] = rb[i];
  };

  double4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  const int nk = (K + BK - 1) / BK;
  load(0);
  stash(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[cur][kk + (lane >> 4)][wm * WM + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[cur][wn * WN + 16 * j + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) stash(cur ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = col0 + wn * WN + 16 * j + (lane & 15);
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * WM + 16 * i + (lane >> 4) + 4 * r;
        if (row < M) {
          double *cp = C + static_cast<size_t>(row) + static_cast<size_t>(col) * lda;
          *cp = *cp - acc[i][j][r];
        }
      }
    }
}

// ---------------------------------------------------------------------------------------------
// Row interchange + unit-lower triangular solve on a strip of NS columns.
//
// swap_list holds the panel's net row movement as (dst, src) pairs: entries [0, w) are rows
// pb .. pb+w-1 in order, later entries are rows >= pb+w that received a row from the panel.
// The strip's touched rows are gathered (all reads before any write), displaced rows are written
// back in place, and rows [pb, pb+w) are solved in LDS:
//   X := L11^{-1} X  with L11 = A(pb:pb+w, pb:pb+w) unit lower,
// by 16-row blocks: the diagonal 16 x 16 blocks are inverted explicitly (every workgroup does
// it for itself), then X_b := inv(L_bb) X_b and X_rest -= L(rest, b) X_b.
// ---------------------------------------------------------------------------------------------
template <int NB, int NS>
__global__ __launch_bounds__(256) void swap_trsm_kernel(double *__restrict__ a, int lda, int pb,
                                                       int w, int c0, int c1,
                                                       const int *__restrict__ swap_list) {
  __shared__ double X[NS][NB + 1];
  __shared__ double E[NS][NB + 1];
  __shared__ double invL[NB / 16][16][17];
  const int t = threadIdx.x;
  const int cnt = swap_list[4 * NB];
  const int *dst = swap_list;
  const int *src = swap_list + 2 * NB;
  const int cs = c0 + blockIdx.x * NS;
  const int ncol = (c1 - cs) < NS ? (c1 - cs) : NS;
  const int nblk = (w + 15) / 16;
  const double *L = a + static_cast<size_t>(pb) + static_cast<size_t>(pb) * lda;

  // Inverses of the diagonal blocks: thread (blk, col) solves L_bb x = e_col.
  if (t < 16 * nblk) {
    const int blk = t >> 4, col = t & 15;
    const int r0 = blk * 16;
    const int nr = (w - r0) < 16 ? (w - r0) : 16;
    double x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      double s = (r == col) ? 1.0 : 0.0;
      if (r > col && r < nr) {
#pragma unroll
        for (int q = 0; q < 16; ++q)
          if (q >= col && q < r)
            s -= L[static_cast<size_t>(r0 + r) + static_cast<size_t>(r0 + q) * lda] * x[q];
      }
      x[r] = (r < nr) ? s : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) invL[blk][r][col] = x[r];
  }

  // Gather every touched row of the strip before anything is written.
  for (int idx = t; idx < cnt * ncol; idx += 256) {
    const int i = idx % cnt, col = idx / cnt;
    const double val = a[static_cast<size_t>(src[i]) + static_cast<size_t>(cs + col) * lda];
    if (i < w)
      X[col][i] = val;
    else
      E[col][i - w] = val;
  }
  __syncthreads();
  for (int idx = t; idx < (cnt - w) * ncol; idx += 256) {
    const int i = idx % (cnt - w), col = idx / (cnt - w);
    a[static_cast<size_t>(dst[w + i]) + static_cast<size_t>(cs + col) * lda] = E[col][i];
  }

  for (int blk = 0; blk < nblk; ++blk) {
    const int r0 = blk * 16;
    const int nr = (w - r0) < 16 ? (w - r0) : 16;
    // X_b := inv(L_bb) X_b
    double y = 0.0;
    const int rr = t & 15, cc = t >> 4;  // 16 rows x 16 columns of the strip
    if (cc < ncol && rr < nr) {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q <= rr && q < nr) y += invL[blk][rr][q] * X[cc][r0 + q];
    }
    __syncthreads();
    if (cc < ncol && rr < nr) X[cc][r0 + rr] = y;
    __syncthreads();
    // rows below the block: X_r -= L(r, b) X_b
    const int col = t % NS, rg = t / NS;
    constexpr int RG = 256 / NS;
    if (col < ncol) {
      for (int r = r0 + 16 + rg; r < w; r += RG) {
        double s = X[col][r];
#pragma unroll
        for (int q = 0; q < 16; ++q)
          if (q < nr) s -= L[static_cast<size_t>(r) + static_cast<size_t>(r0 + q) * lda] * X[col][r0 + q];
        X[col][r] = s;
      }
    }
    __syncthreads();
  }

  for (int idx = t; idx < w * ncol; idx += 256) {
    const int i = idx % w, col = idx / w;
    a[static_cast<size_t>(pb + i) + static_cast<size_t>(cs + col) * lda] = X[col][i];
  }
}

template <int BM, int BN>
hipError_t launch_gemm_tile(double *c, const double *a, const double *b, int lda, int M, int N,
                            int K, hipStream_t s) {
  const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  hipLaunchKernelGGL((gemm_sub_kernel<BM, BN>), dim3(tm * tn), dim3(256), 0, s, c, a, b, lda, M,
                     N, K, tm, tn);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm_sub(double *c, const double *a, const double *b, int lda, int M, int N,
                           int K, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  const long long big = static_cast<long long>((M + 127) / 128) * ((N + 127) / 128);
  if (big >= 512) return launch_gemm_tile<128, 128>(c, a, b, lda, M, N, K, s);
  return launch_gemm_tile<64, 64>(c, a, b, lda, M, N, K, s);
}

hipError_t launch_swap_trsm(double *a, int lda, int pb, int w, int c0, int c1,
                            const int *swap_list, int nb, hipStream_t s) {
  if (c1 <= c0 || w <= 0) return hipSuccess;
  constexpr int NS = 16;
  const dim3 grid((c1 - c0 + NS - 1) / NS), block(256);
  switch (nb) {
    case 64: hipLaunchKernelGGL((swap_trsm_kernel<64, NS>), grid, block, 0, s, a, lda, pb, w, c0, c1, swap_list); break;
    case 128: hipLaunchKernelGGL((swap_trsm_kernel<128, NS>), grid, block, 0, s, a, lda, pb, w, c0, c1, swap_list); break;
    case 256: hipLaunchKernelGGL((swap_trsm_kernel<256, NS>), grid, block, 0, s, a, lda, pb, w, c0, c1, swap_list); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace dlu
