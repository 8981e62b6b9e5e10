// Level-3 kernels of the delayed-pivoting LU on gfx950: the MFMA trailing update and the
// triangular solve that forms the U block row, in FP64 and FP32.
//
// Reference behaviour restated here (delayed_lu/delayed_dgetrf.cpp):
//   finish_panel: U12 = L11^{-1} A12 (unit lower), A22 -= L21 U12        :61-85
//   materialize_partial_panel: the same two steps for an accepted prefix  :101-125
//   row interchanges span the columns right of the panel                  :202-205
#include "dlu_internal.hpp"
#include "dlu_tuning.inc"

namespace dlu {
namespace {

// ---------------------------------------------------------------------------------------------
// C(M x N) -= A(M x K) * B(K x N), column-major, leading dimensions ldc, lda, ldb.
//
// Workgroup tile BM x BN, 256 threads = 4 waves in a 2 x 2 grid, each wave owns a
// (BM/2) x (BN/2) block made of 16 x 16 MFMA tiles (Mfma<T>).  K advances in chunks of BK staged
// through double-buffered LDS with one barrier per chunk:
//   k-step 0         the next chunk's global loads are issued;
//   k-steps 0..KS-2  each step reads the fragments of the following step from LDS, then
//                    issues its own MFMAs;
//   k-step KS-1      the next chunk is stored to the other LDS buffer, the barrier is passed,
//                    the next chunk's first fragments are read, and only then are this step's
//                    MFMAs issued, so they cover the LDS latency after the barrier.
// The last chunk is peeled, which keeps every k-step free of branches.
//
// The MFMA operands are swapped (the B fragment is the first operand): a lane's accumulator
// entries then lie in one row of C and 16 consecutive lanes hold 16 consecutive rows, so the
// column-major C is loaded and stored in whole 16-row runs (one 128-byte line per instruction
// and 16 lanes in FP64) instead of 16 scattered columns.
//
// Global loads are lane-contiguous (consecutive lanes read consecutive rows of A, consecutive k
// of B), 16 bytes per lane when the operands allow it.
// ---------------------------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK, bool VEC, int OCC, bool AMAX>
__global__ __launch_bounds__(256, OCC) void gemm_sub_kernel(T *__restrict__ C, int ldc,
                                                         const T *__restrict__ A, int lda,
                                                         const T *__restrict__ B, int ldb,
                                                         int M, int N, int K, int tiles_m,
                                                         int tiles_n, Pred pred, AmaxOut amax) {
  if (skipped(pred)) return;
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
  static_assert(KS % 2 == 0, "fragment slots alternate: an even number of k-steps per chunk");
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
  // MFMAs were issued, so the loads stay in flight meanwhile.  Row / column clamps are fixed per
  // thread; the k clamp is needed only in a last, partial chunk.  On the vector path M and K are
  // multiples of the vector width, so a vector is either wholly inside or wholly outside.
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

  // Accumulator (i, j) entry r of this lane is C(crow(i), ccol(j, r)).
  const int crow_base = row0 + wm * WM + (lane & 15);
  const int ccol_base = col0 + wn * WN;
  auto ccol = [&](int j, int r) { return ccol_base + 16 * j + MF::acc_row(lane, r); };

  // The accumulators start from C (loaded together with the first chunk, so its latency is not
  // paid again at the end) and the MFMAs add (-A) * B: the epilogue only stores.
  typename MF::acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = crow_base + 16 * i;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = ccol(j, r);
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
  auto mmas = [&](int slot) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(bf[slot][j], af[slot][i], acc[i][j]);
  };

  const int nk = (K + BK - 1) / BK;
  load(0);
  stash(0, 0);
  __syncthreads();
  frags(0, 0, 0);
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      __builtin_amdgcn_sched_barrier(0);
      if (s == 0) load((kt + 1) * BK);
      if (s + 1 < KS) {
        frags(cur, 4 * (s + 1), (s + 1) & 1);
      } else {
        stash(cur ^ 1, (kt + 1) * BK);
        __syncthreads();
        frags(cur ^ 1, 0, 0);
      }
      mmas(s & 1);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  {
    const int cur = (nk - 1) & 1;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      __builtin_amdgcn_sched_barrier(0);
      if (s + 1 < KS) frags(cur, 4 * (s + 1), (s + 1) & 1);
      mmas(s & 1);
      __builtin_amdgcn_sched_barrier(0);
    }
  }

#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = crow_base + 16 * i;
    if (row >= M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = ccol(j, r);
        if (col < N) C[static_cast<size_t>(row) + static_cast<size_t>(col) * ldc] = acc[i][j][r];
      }
  }
  if (AMAX) {  // the full-pivot argmax of this tile of the result
    double bv = 0.0;  // a zero entry never displaces the initial candidate
    unsigned long long bo = 0ull;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int row = crow_base + 16 * i;
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = ccol(j, r);
          const double v = (row < M && col < N) ? fabs(static_cast<double>(acc[i][j][r])) : 0.0;
          const unsigned long long o =
              static_cast<unsigned long long>(col + amax.col_shift) *
                  static_cast<unsigned long long>(amax.rows) +
              static_cast<unsigned long long>(row);
          if (better(v, o, bv, bo)) {
            bv = v;
            bo = o;
          }
        }
    }
    block_best(bv, bo);
    if (threadIdx.x == 0) {
      amax.v[blockIdx.x] = bv;
      amax.o[blockIdx.x] = bo;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Row interchange + unit-lower triangular solve on a strip of 16 columns, one wave per strip.
//
// swap_list holds the panel's net row movement as (dst, src) pairs: entries [0, w) are rows
// pb .. pb+w-1 in order, later entries are rows >= pb+w that received a row from the panel (no
// list: rows [pb, pb+w) in place, solve only).  The wave owns its 16 columns outright, so no
// other wave touches them: it reads every touched row of the strip, writes the displaced rows
// back, and solves
//   X := L11^{-1} X  with L11 = A(pb:pb+w, pb:pb+w) unit lower,
// with X held in registers as NB/16 MFMA result tiles: per 16-row block b, X_b := inv(L_bb) X_b,
// then X_t -= L_tb X_b for every tile t below.  A result tile is the B operand of the next
// product as it stands when k runs in the result's row order (k-step s of lane l is row
// acc_row(l, s)); the A operand is read in that same k order.  The 16 x 16 diagonal inverses
// are formed by the wave itself in LDS.  Rows at and beyond w are zero-padded and never
// written back.
// ---------------------------------------------------------------------------------------------
template <typename T, int NB, bool LIST, bool FULL, bool TWO = false>
__global__ __launch_bounds__(64) void swap_trsm_kernel(T *__restrict__ a, int lda, int pb, int w,
                                                      int c0, int c1,
                                                      const int *__restrict__ swap_list, int ls,
                                                      Pred pred,
                                                      const int *__restrict__ list2) {
  if (skipped(pred)) return;
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int NT = NB / 16;  // 16-row tiles of the panel
  constexpr int NJ = NB / 64;  // rows per lane in a 64-row sweep
  __shared__ T X[NB][17];
  __shared__ T D[NT][16][17];  // diagonal blocks of L11, then their inverses
  const int l = threadIdx.x;
  const int lr = l & 15;
  // The list was written for panels of width ls >= NB (its layout: ls is the stride).
  const int cnt = LIST ? swap_list[4 * ls] : w;
  const int *src = swap_list + 2 * ls;  // read only with a list
  const int cs = c0 + blockIdx.x * 16;
  const int ncol = (c1 - cs) < 16 ? (c1 - cs) : 16;
  const T *L = a + static_cast<size_t>(pb) + static_cast<size_t>(pb) * lda;
  T *S = a + static_cast<size_t>(cs) * lda;  // column cs: the strip

  // 1. Every read of the strip before any write: rows [0, w) after the interchange into LDS,
  // displaced rows into registers.  Lanes run down a column, 64 consecutive rows per load.
  // Columns past the strip's end read column cmax instead and are never stored (no column of X
  // feeds another).  FULL (w == NB) needs no row mask, so no loaded value goes through a select
  // and the loads issue back to back; a partial panel zeroes rows and columns at and beyond w.
  const int cmax = ncol - 1;
  int xsrc[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = 64 * j + l;
    const int from = LIST ? src[i] : pb + i;  // i < NB: inside the list
    xsrc[j] = i < w ? from : -1;
  }
  T e[NJ][16];
  int edst[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int idx = w + 64 * j + l;  // < 2 NB: inside the list; displaced rows number <= w
    const bool ok = LIST && idx < cnt;
    const int to = LIST ? swap_list[idx] : -1, from = LIST ? src[idx] : pb;
    edst[j] = ok ? to : -1;
    const int es = ok ? from : pb;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      e[j][c] = S[static_cast<size_t>(es) + static_cast<size_t>(c < cmax ? c : cmax) * lda];
  }
#pragma unroll
  for (int c = 0; c < 16; ++c)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int row = (FULL || xsrc[j] >= 0) ? xsrc[j] : pb;
      const T v = S[static_cast<size_t>(row) + static_cast<size_t>(c < cmax ? c : cmax) * lda];
      X[64 * j + l][c] = (FULL || xsrc[j] >= 0) ? v : T(0);
    }
  // Diagonal blocks of L11 (zero outside [0, w)), 16 consecutive rows per column segment.
#pragma unroll
  for (int j = 0; j < 4 * NT; ++j) {
    const int idx = 64 * j + l;
    const int b = idx >> 8, q = (idx >> 4) & 15, r = idx & 15;
    const int row = 16 * b + r, col = 16 * b + q;
    const bool ok = FULL || (row < w && col < w);
    const T v = L[ok ? static_cast<size_t>(row) + static_cast<size_t>(col) * lda : 0];
    D[b][r][q] = ok ? v : T(0);
  }
  __syncthreads();  // the strip's reads have all returned (their values sit in LDS)

  // 2. Displaced rows go back in place.
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (edst[j] >= 0)
#pragma unroll
      for (int c = 0; c < 16; ++c)
        if (c < ncol) S[static_cast<size_t>(edst[j]) + static_cast<size_t>(c) * lda] = e[j][c];

  // 3. inv(L_bb), column lr of block b per lane: forward substitution on e_lr.  Blocks past w
  // are zero, so their inverse is zero and the solve below needs no branch on w.
  T x[NT / 4][16];
#pragma unroll
  for (int h = 0; h < NT / 4; ++h) {
    const int b = (l >> 4) + 4 * h;
#pragma unroll
    for (int r = 0; r < 16; ++r) x[h][r] = (r == lr) ? T(1) : T(0);
#pragma unroll
    for (int q = 0; q < 15; ++q)
#pragma unroll
      for (int r = q + 1; r < 16; ++r) x[h][r] -= D[b][r][q] * x[h][q];
    if (!FULL) {
      const int nr = w - 16 * b;
#pragma unroll
      for (int r = 0; r < 16; ++r) x[h][r] = (r < nr) ? x[h][r] : T(0);
    }
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < NT / 4; ++h)
#pragma unroll
    for (int r = 0; r < 16; ++r) D[(l >> 4) + 4 * h][r][lr] = x[h][r];
  __syncthreads();

  // 4. The solve on MFMA with X in registers.  The A operands of block b + 1 (its column slice
  // of L, negated) are loaded while block b's products run.
  acc_t xr[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) xr[t][r] = X[16 * t + MF::acc_row(l, r)][lr];
  auto lslice = [&](int b, int t, int s) -> T {
    const int row = 16 * t + lr, k = 16 * b + MF::acc_row(l, s);
    const bool ok = FULL || (row < w && k < w);
    const T v = L[ok ? static_cast<size_t>(row) + static_cast<size_t>(k) * lda : 0];
    return ok ? -v : T(0);
  };
  T lcur[NT][4];
#pragma unroll
  for (int t = 1; t < NT; ++t)
#pragma unroll
    for (int s = 0; s < 4; ++s) lcur[t][s] = lslice(0, t, s);
#pragma unroll
  for (int b = 0; b < NT; ++b) {
    T lnext[NT][4];
#pragma unroll
    for (int t = b + 2; t < NT; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) lnext[t][s] = lslice(b + 1, t, s);
    acc_t d = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int s = 0; s < 4; ++s) d = MF::mma(D[b][lr][MF::acc_row(l, s)], xr[b][s], d);
    xr[b] = d;
#pragma unroll
    for (int t = b + 1; t < NT; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) xr[t] = MF::mma(lcur[t][s], xr[b][s], xr[t]);
#pragma unroll
    for (int t = b + 2; t < NT; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) lcur[t][s] = lnext[t][s];
  }

  // 5. Rows [pb, pb+w) of the strip, through LDS so the stores run down columns.
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) X[16 * t + MF::acc_row(l, r)][lr] = xr[t][r];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 16; ++c)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int i = 64 * j + l;
      if (i < w && c < ncol) S[static_cast<size_t>(pb + i) + static_cast<size_t>(c) * lda] = X[i][c];
    }
  if (!TWO) return;

  // 6. TWO (the first panel of a pair, full): the second panel's movement (list2, rows from
  // pb + NB on) on the same strip, which reads rows this wave's displaced-row stores may just
  // have written (its own stores: a workgroup-scope fence orders them).
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __syncthreads();
  const int pb2 = pb + NB;
  const int cnt2 = list2[4 * ls];
  const int *src2 = list2 + 2 * ls;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int idx = NB + 64 * j + l;
    const bool ok = idx < cnt2;
    edst[j] = ok ? list2[idx] : -1;
    const int es = ok ? src2[idx] : pb2;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      e[j][c] = S[static_cast<size_t>(es) + static_cast<size_t>(c < cmax ? c : cmax) * lda];
  }
#pragma unroll
  for (int c = 0; c < 16; ++c)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int row = src2[64 * j + l];
      X[64 * j + l][c] = S[static_cast<size_t>(row) + static_cast<size_t>(c < cmax ? c : cmax) * lda];
    }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (edst[j] >= 0)
#pragma unroll
      for (int c = 0; c < 16; ++c)
        if (c < ncol) S[static_cast<size_t>(edst[j]) + static_cast<size_t>(c) * lda] = e[j][c];
#pragma unroll
  for (int c = 0; c < 16; ++c)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
      if (c < ncol)
        S[static_cast<size_t>(pb2 + 64 * j + l) + static_cast<size_t>(c) * lda] = X[64 * j + l][c];
}

template <typename T, int BM, int BN, int BK, int OCC>
hipError_t launch_gemm_tile(T *c, int ldc, const T *a, int lda, const T *b, int ldb, int M, int N,
                            int K, hipStream_t s, Pred pred, AmaxOut amax) {
  const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  // 16-byte loads: 16-byte aligned origins, leading dimensions, M and K multiples of the width.
  constexpr int vw = 16 / static_cast<int>(sizeof(T));
  const bool vec = (lda % vw == 0) && (ldb % vw == 0) && (M % vw == 0) && (K % vw == 0) &&
                   (reinterpret_cast<uintptr_t>(a) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(b) % 16 == 0);
#define DLU_GEMM_LAUNCH(V, A)                                                                    \
  hipLaunchKernelGGL((gemm_sub_kernel<T, BM, BN, BK, V, OCC, A>), dim3(tm * tn), dim3(256), 0, s, \
                     c, ldc, a, lda, b, ldb, M, N, K, tm, tn, pred, amax)
  if (amax.v != nullptr) {
    if (vec)
      DLU_GEMM_LAUNCH(true, true);
    else
      DLU_GEMM_LAUNCH(false, true);
  } else if (vec) {
    DLU_GEMM_LAUNCH(true, false);
  } else {
    DLU_GEMM_LAUNCH(false, false);
  }
#undef DLU_GEMM_LAUNCH
  return hipGetLastError();
}

}  // namespace

namespace {

// The tile configurations of the trailing-update GEMM (workgroup tile, workgroups per CU).
template <typename T>
hipError_t launch_gemm_config(int config, T *c, int ldc, const T *a, int lda, const T *b,
                              int ldb, int M, int N, int K, hipStream_t s, Pred pred,
                              AmaxOut amax) {
  switch (config) {
    case 0: return launch_gemm_tile<T, 64, 64, 16, 4>(c, ldc, a, lda, b, ldb, M, N, K, s, pred, amax);
    case 1:
      return launch_gemm_tile<T, 128, 128, 16, 2>(c, ldc, a, lda, b, ldb, M, N, K, s, pred, amax);
    case 2: return launch_gemm_tile<T, 128, 64, 16, 3>(c, ldc, a, lda, b, ldb, M, N, K, s, pred, amax);
    case 3: return launch_gemm_tile<T, 64, 128, 16, 3>(c, ldc, a, lda, b, ldb, M, N, K, s, pred, amax);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int gemm_workgroups(int config, int M, int N) {
  static const int bm[kGemmConfigs] = {64, 128, 128, 64}, bn[kGemmConfigs] = {64, 128, 64, 128};
  if (config < 0 || config >= kGemmConfigs) return 0;
  return ((M + bm[config] - 1) / bm[config]) * ((N + bn[config] - 1) / bn[config]);
}

int tuned_gemm_config(int M, int N, int K) {
  const long long tiles = static_cast<long long>((M + 127) / 128) * ((N + 127) / 128);
  const int size_class = tiles < 64 ? 0 : tiles < 256 ? 1 : tiles < 512 ? 2 : tiles < 2048 ? 3 : 4;
  return kTunedGemmTiles[K > 160 ? 1 : 0][size_class];
}


template <typename T>
hipError_t launch_gemm_sub(T *c, int ldc, const T *a, int lda, const T *b, int ldb, int M, int N,
                           int K, hipStream_t s, int config, Pred pred, AmaxOut amax) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  // The tile configuration comes from the table the autotuner wrote (by depth and launch size)
  // unless the caller names one.
  return launch_gemm_config<T>(config >= 0 ? config : tuned_gemm_config(M, N, K), c, ldc, a, lda,
                               b, ldb, M, N, K, s, pred, amax);
}

template <typename T>
hipError_t launch_swap_trsm(T *a, int lda, int pb, int w, int c0, int c1, const int *swap_list,
                            int nb, hipStream_t s, Pred pred) {
  if (c1 <= c0 || w <= 0) return hipSuccess;
  const dim3 grid((c1 - c0 + 15) / 16), block(64);
  // The solve's tile height follows w (a restarted panel's accepted prefix is often a few rows):
  // the 64-row kernel serves any list while w <= 64.
  const int tile = w <= 64 ? 64 : nb;
#define DLU_TRSM_CASE(NBV)                                                                      \
  case NBV:                                                                                      \
    if (swap_list && w == NBV)                                                                   \
      hipLaunchKernelGGL((swap_trsm_kernel<T, NBV, true, true>), grid, block, 0, s, a, lda, pb,  \
                         w, c0, c1, swap_list, nb, pred, nullptr);                                        \
    else if (swap_list)                                                                          \
      hipLaunchKernelGGL((swap_trsm_kernel<T, NBV, true, false>), grid, block, 0, s, a, lda, pb, \
                         w, c0, c1, swap_list, nb, pred, nullptr);                                        \
    else if (w == NBV)                                                                           \
      hipLaunchKernelGGL((swap_trsm_kernel<T, NBV, false, true>), grid, block, 0, s, a, lda, pb, \
                         w, c0, c1, swap_list, nb, pred, nullptr);                                        \
    else                                                                                         \
      hipLaunchKernelGGL((swap_trsm_kernel<T, NBV, false, false>), grid, block, 0, s, a, lda,    \
                         pb, w, c0, c1, swap_list, nb, pred, nullptr);                                    \
    break;
  if (nb != 64 && nb != 128 && nb != 256) return hipErrorInvalidValue;
  switch (tile) {
    DLU_TRSM_CASE(64)
    DLU_TRSM_CASE(128)
    DLU_TRSM_CASE(256)
    default: return hipErrorInvalidValue;
  }
#undef DLU_TRSM_CASE
  return hipGetLastError();
}


template <typename T>
hipError_t launch_swap_trsm_pair(T *a, int lda, int b0, int c0, int c1, const int *l0,
                                 const int *l1, int nb, hipStream_t s, Pred pred) {
  if (c1 <= c0) return hipSuccess;
  const dim3 grid((c1 - c0 + 15) / 16), block(64);
  switch (nb) {
    case 64:
      hipLaunchKernelGGL((swap_trsm_kernel<T, 64, true, true, true>), grid, block, 0, s, a, lda,
                         b0, 64, c0, c1, l0, nb, pred, l1);
      break;
    case 128:
      hipLaunchKernelGGL((swap_trsm_kernel<T, 128, true, true, true>), grid, block, 0, s, a, lda,
                         b0, 128, c0, c1, l0, nb, pred, l1);
      break;
    case 256:
      hipLaunchKernelGGL((swap_trsm_kernel<T, 256, true, true, true>), grid, block, 0, s, a, lda,
                         b0, 256, c0, c1, l0, nb, pred, l1);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define DLU_BLAS3_INSTANTIATE(T)                                                                 \
  template hipError_t launch_gemm_sub<T>(T *, int, const T *, int, const T *, int, int, int, int, \
                                         hipStream_t, int, Pred, AmaxOut);                       \
  template hipError_t launch_swap_trsm<T>(T *, int, int, int, int, int, const int *, int,        \
                                          hipStream_t, Pred);                                    \
  template hipError_t launch_swap_trsm_pair<T>(T *, int, int, int, int, const int *, const int *, \
                                               int, hipStream_t, Pred);
DLU_BLAS3_INSTANTIATE(float)
DLU_BLAS3_INSTANTIATE(double)
#undef DLU_BLAS3_INSTANTIATE

}  // namespace dlu
