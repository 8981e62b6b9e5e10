// Level-3 kernels of the delayed-pivoting LU on gfx950: the MFMA trailing update and the
// triangular solve that forms the U block row, in FP64 and FP32.
//
// Reference behaviour restated here (delayed_lu/delayed_dgetrf.cpp):
//   finish_panel: U12 = L11^{-1} A12 (unit lower), A22 -= L21 U12        :61-85
//   materialize_partial_panel: the same two steps for an accepted prefix  :101-125
//   row interchanges span the columns right of the panel                  :202-205
#include "dlu_internal.hpp"

namespace dlu {
namespace {

// ---------------------------------------------------------------------------------------------
// C(M x N) -= A(M x K) * B(K x N), column-major, leading dimensions ldc, lda, ldb.
//
// Workgroup tile BM x BN, 256 threads = 4 waves in a 2 x 2 grid, each wave owns a
// (BM/2) x (BN/2) block made of 16 x 16 MFMA tiles (Mfma<T>).  K advances in chunks of BK staged
// through double-buffered LDS; the next chunk's global loads are in flight while the current
// chunk feeds the MFMA pipe.  Global loads are lane-contiguous (consecutive lanes read
// consecutive rows of A, consecutive k of B), 16 bytes per lane when the operands allow it.
// ---------------------------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK, bool VEC>
__global__ __launch_bounds__(256, 2) void gemm_sub_kernel(T *__restrict__ C, int ldc,
                                                         const T *__restrict__ A, int lda,
                                                         const T *__restrict__ B, int ldb,
                                                         int M, int N, int K, int tiles_m,
                                                         int tiles_n) {
  typedef Mfma<T> MF;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int AS = BM + 16;  // A tile k-row stride: consecutive k-rows a fragment read touches
                               // land on opposite LDS bank halves
  constexpr int VW = VEC ? 16 / static_cast<int>(sizeof(T)) : 1;  // elements per global load
  constexpr int BS = BK + 16 / static_cast<int>(sizeof(T));       // B tile column stride
  constexpr int A_LD = BM * BK / (256 * VW);  // loads per thread per chunk
  constexpr int B_LD = BN * BK / (256 * VW);
  constexpr int KS = BK / 4;                  // MFMA k-steps per chunk
  static_assert(A_LD >= 1 && B_LD >= 1, "tile too small for the load width");
  __shared__ T As[2][BK][AS];
  __shared__ T Bs[2][BN][BS];

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

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // Global loads are branch-free: an out-of-range element is read from a clamped in-range
  // address, and replaced by zero only when it is stashed in LDS, after the current chunk's
  // MFMAs, so the loads of the next chunk stay in flight while this one computes.  Row / column
  // clamps are fixed per thread; the k clamp is needed only in a last, partial chunk.  On the
  // vector path M and K are multiples of the vector width, so a vector is either wholly inside
  // or wholly outside.
  typedef T dv_t __attribute__((ext_vector_type(VW)));
  constexpr int AR = BM / VW, BR = BK / VW;  // threads per A column / per B column
  static_assert(256 % AR == 0 && 256 % BR == 0, "load layout");
  dv_t ra[A_LD], rb[B_LD];
  // A: every load of a thread reads the same row, k rows apart by 256 / AR.
  const int a_row = row0 + (t % AR) * VW;
  const bool a_ok = a_row < M;
  const int a_k = t / AR;
  const T *pa = A + static_cast<size_t>(a_ok ? a_row : 0);
  // B: every load of a thread reads the same k, columns apart by 256 / BR.
  const int b_k = (t % BR) * VW;
  const int b_col = col0 + t / BR;
  const T *pb = B + static_cast<size_t>(b_k);
  auto load = [&](int k0) {
    const bool full = k0 + BK <= K;  // wave-uniform
#pragma unroll
    for (int i = 0; i < A_LD; ++i) {
      const int kk = k0 + a_k + i * (256 / AR);
      const int kc = (full || kk < K) ? kk : 0;
      ra[i] = *reinterpret_cast<const dv_t *>(pa + static_cast<size_t>(kc) * lda);
    }
#pragma unroll
    for (int i = 0; i < B_LD; ++i) {
      const int col = b_col + i * (256 / BR);
      const int kc = (full || k0 + b_k < K) ? k0 : -b_k;  // a k beyond K reads row 0
      rb[i] = *reinterpret_cast<const dv_t *>(
          pb + kc + static_cast<size_t>(col < N ? col : 0) * ldb);
    }
  };
  auto stash = [&](int buf, int k0) {
#pragma unroll
    for (int i = 0; i < A_LD; ++i) {
      const int kr = a_k + i * (256 / AR);
      const bool ok = a_ok && k0 + kr < K;
      *reinterpret_cast<dv_t *>(&As[buf][kr][(t % AR) * VW]) =
          ok ? -ra[i] : dv_t(T(0));  // A is stored negated: acc = C + (-A) B
    }
#pragma unroll
    for (int i = 0; i < B_LD; ++i) {
      const int cr = t / BR + i * (256 / BR);
      const bool ok = b_col + i * (256 / BR) < N && k0 + b_k < K;
      *reinterpret_cast<dv_t *>(&Bs[buf][cr][b_k]) = ok ? rb[i] : dv_t(T(0));
    }
  };

  // The accumulators start from C (loaded together with the first chunk, so its latency is not
  // paid again at the end) and the MFMAs add (-A) * B: the epilogue only stores.
  typename MF::acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = col0 + wn * WN + 16 * j + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wm * WM + 16 * i + MF::acc_row(lane, r);
        const bool ok = row < M && col < N;
        const T v = C[static_cast<size_t>(ok ? row : 0) + static_cast<size_t>(ok ? col : 0) * ldc];
        acc[i][j][r] = ok ? v : T(0);
      }
    }

  T af[2][FM], bf[2][FN];
  auto frags = [&](int buf, int kk, int slot) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
      af[slot][i] = As[buf][kk + (lane >> 4)][wm * WM + 16 * i + (lane & 15)];
#pragma unroll
    for (int j = 0; j < FN; ++j)
      bf[slot][j] = Bs[buf][wn * WN + 16 * j + (lane & 15)][kk + (lane >> 4)];
  };

  const int nk = (K + BK - 1) / BK;
  load(0);
  stash(0, 0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
    frags(cur, 0, 0);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      // LDS reads for the next k-step are issued before this step's MFMAs.
      if (s + 1 < KS) frags(cur, 4 * (s + 1), (s + 1) & 1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(af[s & 1][i], bf[s & 1][j], acc[i][j]);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (kt + 1 < nk) stash(cur ^ 1, (kt + 1) * BK);
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
        const int row = row0 + wm * WM + 16 * i + MF::acc_row(lane, r);
        if (row < M) C[static_cast<size_t>(row) + static_cast<size_t>(col) * ldc] = acc[i][j][r];
      }
    }
}

// ---------------------------------------------------------------------------------------------
// Row interchange + unit-lower triangular solve on a strip of 16 columns.
//
// swap_list holds the panel's net row movement as (dst, src) pairs: entries [0, w) are rows
// pb .. pb+w-1 in order, later entries are rows >= pb+w that received a row from the panel.
// The strip's touched rows are gathered (all reads before any write), displaced rows are written
// back in place, and rows [pb, pb+w) are solved in LDS:
//   X := L11^{-1} X  with L11 = A(pb:pb+w, pb:pb+w) unit lower,
// by 16-row blocks on MFMA: X_b := inv(L_bb) X_b (the 16 x 16 diagonal inverses are formed by
// every workgroup for itself), then X_rest -= L(rest, b) X_b with L's column slice staged in
// LDS.  Rows at and beyond w are zero-padded and never written back.
// ---------------------------------------------------------------------------------------------
template <typename T, int NB>
__global__ __launch_bounds__(256) void swap_trsm_kernel(T *__restrict__ a, int lda, int pb,
                                                       int w, int c0, int c1,
                                                       const int *__restrict__ swap_list) {
  typedef Mfma<T> MF;
  constexpr int NS = 16;
  constexpr int XS = NS + 2;  // row stride of X
  __shared__ T X[NB][XS];
  __shared__ T E[NB][NS + 1];
  __shared__ T invL[NB / 16][16][17];
  __shared__ T Ls[NB][17];  // rows [r0+16, w) of L's current 16-column slice
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  // No list: rows [pb, pb+w) in place (triangular solve only).
  const int cnt = swap_list ? swap_list[4 * NB] : w;
  const int nx = cnt - w;  // displaced rows below the panel
  const int *dst = swap_list;
  const int *src = swap_list ? swap_list + 2 * NB : nullptr;
  const int cs = c0 + blockIdx.x * NS;
  const int ncol = (c1 - cs) < NS ? (c1 - cs) : NS;
  const int nblk = (w + 15) / 16;
  const T *L = a + static_cast<size_t>(pb) + static_cast<size_t>(pb) * lda;

  // Gather every touched row of the strip before anything is written (rows are contiguous in
  // lane order for a fixed column).
  for (int idx = t; idx < cnt * NS; idx += 256) {
    const int i = idx % cnt, col = idx / cnt;
    const int srow = src ? src[i] : pb + i;
    const T val =
        col < ncol ? a[static_cast<size_t>(srow) + static_cast<size_t>(cs + col) * lda] : T(0);
    if (i < w)
      X[i][col] = val;
    else
      E[i - w][col] = val;
  }
  for (int idx = t; idx < (nblk * 16 - w) * NS; idx += 256) X[w + idx / NS][idx % NS] = T(0);

  // Inverses of the diagonal blocks: thread (blk, col) solves L_bb x = e_col.
  if (t < 16 * nblk) {
    const int blk = t >> 4, col = t & 15;
    const int r0 = blk * 16;
    const int nr = (w - r0) < 16 ? (w - r0) : 16;
    T x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      T sum = (r == col) ? T(1) : T(0);
      if (r > col && r < nr) {
#pragma unroll
        for (int q = 0; q < 16; ++q)
          if (q >= col && q < r)
            sum -= L[static_cast<size_t>(r0 + r) + static_cast<size_t>(r0 + q) * lda] * x[q];
      }
      x[r] = (r < nr) ? sum : T(0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) invL[blk][r][col] = x[r];
  }
  __syncthreads();
  for (int idx = t; idx < nx * ncol; idx += 256) {
    const int i = idx % nx, col = idx / nx;
    a[static_cast<size_t>(dst[w + i]) + static_cast<size_t>(cs + col) * lda] = E[i][col];
  }

  const int lr = lane & 15, lk = lane >> 4;
  for (int blk = 0; blk < nblk; ++blk) {
    const int r0 = blk * 16;
    const int below = nblk * 16 - (r0 + 16);  // padded rows under the block
    // Stage L(r0+16 .. w, r0 .. r0+16) (zero beyond w) while wave 0 solves the diagonal block.
    for (int idx = t; idx < below * 16; idx += 256) {
      const int rr = idx % below, q = idx / below;
      const int row = r0 + 16 + rr, colq = r0 + q;
      Ls[rr][q] = (row < w && colq < w)
                      ? L[static_cast<size_t>(row) + static_cast<size_t>(colq) * lda]
                      : T(0);
    }
    if (wave == 0) {
      typename MF::acc_t d = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int k4 = 0; k4 < 16; k4 += 4) {
        const T av = invL[blk][lr][k4 + lk];
        const T bv = X[r0 + k4 + lk][lr];
        d = MF::mma(av, bv, d);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) X[r0 + MF::acc_row(lane, r)][lr] = d[r];
    }
    __syncthreads();
    // X(rows below) -= L(rows below, block) * X_b, one 16 x 16 tile per wave at a time.
    for (int tb = wave; tb < below / 16; tb += 4) {
      const int rt = r0 + 16 + 16 * tb;
      typename MF::acc_t d;
#pragma unroll
      for (int r = 0; r < 4; ++r) d[r] = X[rt + MF::acc_row(lane, r)][lr];
#pragma unroll
      for (int k4 = 0; k4 < 16; k4 += 4) {
        const T av = -Ls[16 * tb + lr][k4 + lk];
        const T bv = X[r0 + k4 + lk][lr];
        d = MF::mma(av, bv, d);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) X[rt + MF::acc_row(lane, r)][lr] = d[r];
    }
    __syncthreads();
  }

  for (int idx = t; idx < w * ncol; idx += 256) {
    const int i = idx % w, col = idx / w;
    a[static_cast<size_t>(pb + i) + static_cast<size_t>(cs + col) * lda] = X[i][col];
  }
}

template <typename T, int BM, int BN, int BK>
hipError_t launch_gemm_tile(T *c, int ldc, const T *a, int lda, const T *b, int ldb, int M, int N,
                            int K, hipStream_t s) {
  const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  // 16-byte loads: 16-byte aligned origins, leading dimensions, M and K multiples of the width.
  constexpr int vw = 16 / static_cast<int>(sizeof(T));
  const bool vec = (lda % vw == 0) && (ldb % vw == 0) && (M % vw == 0) && (K % vw == 0) &&
                   (reinterpret_cast<uintptr_t>(a) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(b) % 16 == 0);
  if (vec)
    hipLaunchKernelGGL((gemm_sub_kernel<T, BM, BN, BK, true>), dim3(tm * tn), dim3(256), 0, s, c,
                       ldc, a, lda, b, ldb, M, N, K, tm, tn);
  else
    hipLaunchKernelGGL((gemm_sub_kernel<T, BM, BN, BK, false>), dim3(tm * tn), dim3(256), 0, s,
                       c, ldc, a, lda, b, ldb, M, N, K, tm, tn);
  return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_gemm_sub(T *c, int ldc, const T *a, int lda, const T *b, int ldb, int M, int N,
                           int K, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  if (gemm_uses_large_tiles(M, N))
    return launch_gemm_tile<T, 128, 128, 16>(c, ldc, a, lda, b, ldb, M, N, K, s);
  return launch_gemm_tile<T, 64, 64, 16>(c, ldc, a, lda, b, ldb, M, N, K, s);
}

template <typename T>
hipError_t launch_swap_trsm(T *a, int lda, int pb, int w, int c0, int c1, const int *swap_list,
                            int nb, hipStream_t s) {
  if (c1 <= c0 || w <= 0) return hipSuccess;
  const dim3 grid((c1 - c0 + 15) / 16), block(256);
  switch (nb) {
    case 64: hipLaunchKernelGGL((swap_trsm_kernel<T, 64>), grid, block, 0, s, a, lda, pb, w, c0, c1, swap_list); break;
    case 128: hipLaunchKernelGGL((swap_trsm_kernel<T, 128>), grid, block, 0, s, a, lda, pb, w, c0, c1, swap_list); break;
    case 256: hipLaunchKernelGGL((swap_trsm_kernel<T, 256>), grid, block, 0, s, a, lda, pb, w, c0, c1, swap_list); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define DLU_BLAS3_INSTANTIATE(T)                                                                 \
  template hipError_t launch_gemm_sub<T>(T *, int, const T *, int, const T *, int, int, int, int, \
                                         hipStream_t);                                           \
  template hipError_t launch_swap_trsm<T>(T *, int, int, int, int, int, const int *, int,        \
                                          hipStream_t);
DLU_BLAS3_INSTANTIATE(float)
DLU_BLAS3_INSTANTIATE(double)
#undef DLU_BLAS3_INSTANTIATE

}  // namespace dlu
