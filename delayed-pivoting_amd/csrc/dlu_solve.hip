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

__device__ __forceinline__ void wave_sync_lds_solve() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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
// A whole triangular solve in one launch: X := T^{-1} X with T the unit lower (L) or non-unit
// upper (U) triangle of an n x n factor and X n x nrhs (nrhs <= 16), rows padded in x to a
// multiple of 64.  Rows go in blocks of 64, one workgroup each.  Workgroups take blocks in
// dependency order (L: top first, U: bottom first) from a ticket counter, so every block a
// workgroup waits on is held by a workgroup that already runs.
//
// A solved block travels through xw as tagged exchange words (an LL protocol: 32 payload bits
// and the launch's 32-bit epoch in each 64-bit word, two words per FP64 value), so a reader
// spins on the data itself with relaxed loads and no flag, fence or cache invalidation sits on
// the chain.  Per block i, before the wait that closes it: the four waves form the inverse of
// the 64 x 64 diagonal block D (its 16 x 16 diagonal sub-blocks inverted in LDS, then a block
// substitution on the identity, 16 columns per wave) and M = D^{-1} (-T_{i,jp}), the coupling to
// the block jp solved just before; they subtract the contribution of every earlier block as
// its words arrive (MFMA into four accumulators, the A operands of the next block loaded
// ahead), and form y = D^{-1} (what is left of the right-hand side).  The chain's critical path
// per block is then the arrival of block jp's words, x_i = y + M x_jp (16 MFMAs per wave, no
// exchange between the waves) and the publication of the rows.  Rows at and beyond n are zero
// in y and M, and so in x_i (the inverse is the identity there).  Waits are bounded: after
// spin_limit ticks of the 100 MHz clock a wave sets *err and goes on.
//
// The launch's ticket base and epoch come from the arguments, or, when clock is not null, from
// device memory (written by chain_clock_kernel earlier on the stream): a launch recorded in a
// HIP graph then takes fresh values on every replay.
// ---------------------------------------------------------------------------------------------
template <typename T>
struct ChainWords;
template <>
struct ChainWords<double> {
  static constexpr int n = 2;
  static __device__ __forceinline__ unsigned part(double x, int i) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    return i ? static_cast<unsigned>(b >> 32) : static_cast<unsigned>(b);
  }
  static __device__ __forceinline__ double join(unsigned lo, unsigned hi) {
    return __longlong_as_double(
        static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
  }
};
template <>
struct ChainWords<float> {
  static constexpr int n = 1;
  static __device__ __forceinline__ unsigned part(float x, int) { return __float_as_uint(x); }
  static __device__ __forceinline__ float join(unsigned lo, unsigned) { return __uint_as_float(lo); }
};

template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void trsv_chain_kernel(
    const T *__restrict__ lu, int lda, int n, T *__restrict__ x, int ldx, int nrhs,
    unsigned long long *__restrict__ xw, unsigned long long *__restrict__ ticket,
    unsigned long long ticket_base_arg, unsigned epoch_arg, const ChainClock *__restrict__ clock,
    long long spin_limit, int *__restrict__ err) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  typedef ChainWords<T> CW;
  constexpr int WPV = CW::n;
  __shared__ int s_blk;
  __shared__ T s_d[4][16][17];  // diagonal sub-blocks, then their inverses
  __shared__ T s_m[64][65];     // the block's inverse
  __shared__ T s_c[64][65];     // M, the coupling to the block solved just before
  __shared__ T s_acc[64][17];   // the block's right-hand sides after all contributions
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nblk = (n + 63) / 64;
  const unsigned epoch = clock ? clock->epoch : epoch_arg;
  if (t == 0) {
    const unsigned long long ticket_base = clock ? clock->base : ticket_base_arg;
    const unsigned long long tk = atomicAdd(ticket, 1ull) - ticket_base;
    s_blk = tk < static_cast<unsigned long long>(nblk) ? static_cast<int>(tk) : -1;
    if (s_blk < 0) atomicExch(err, 2);  // the counter and the host's base disagree
  }
  __syncthreads();
  if (s_blk < 0) return;
  const int i = UPPER ? nblk - 1 - s_blk : s_blk;
  const int r0 = 64 * i;
  const int col = lr < nrhs ? lr : nrhs - 1;  // columns past nrhs repeat the last, never stored
  const int last = n - 1;
  const long long t0 = static_cast<long long>(wall_clock64());

  // 1. The diagonal block's inverse, before any wait.  Its four 16 x 16 diagonal sub-blocks go
  // to LDS (entries outside [0, n) become those of the identity) and are inverted in place by
  // one thread per sub-block column; then wave w runs the block substitution over the four
  // sub-blocks on identity columns [16w, 16w + 16), the off-diagonal sub-blocks as negated A
  // operands in the k order of an MFMA result tile (k-step s of lane l is row acc_row(l, s) of
  // the B tile), zero outside [0, n).
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int idx = 256 * j + t;
    const int sb = idx >> 8, q = (idx >> 4) & 15, r = idx & 15;
    const int row = r0 + 16 * sb + r, cl = r0 + 16 * sb + q;
    const T v = lu[static_cast<size_t>(row < n ? row : last) +
                   static_cast<size_t>(cl < n ? cl : last) * lda];
    s_d[sb][r][q] = (row < n && cl < n) ? v : (r == q ? T(1) : T(0));
  }
  __syncthreads();
  if (wave == 0) {
    const int sb = lane >> 4, c = lr;
    T inv[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) inv[r] = (r == c) ? T(1) : T(0);
    if (!UPPER) {  // unit lower: forward substitution on e_c
#pragma unroll
      for (int q = 0; q < 15; ++q)
#pragma unroll
        for (int r = q + 1; r < 16; ++r) inv[r] -= s_d[sb][r][q] * inv[q];
    } else {  // non-unit upper: back substitution on e_c
#pragma unroll
      for (int q = 15; q >= 0; --q) {
        inv[q] = inv[q] / s_d[sb][q][q];
#pragma unroll
        for (int r = 0; r < q; ++r) inv[r] -= s_d[sb][r][q] * inv[q];
      }
    }
    wave_sync_lds_solve();
#pragma unroll
    for (int r = 0; r < 16; ++r) s_d[sb][r][c] = inv[r];
  }
  {
    T dt[6][4];
    int k6 = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (UPPER ? a < b : a > b) {
          const int row = r0 + 16 * a + lr;
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            const int cl = r0 + 16 * b + MF::acc_row(lane, s4);
            const T v = lu[static_cast<size_t>(row < n ? row : last) +
                           static_cast<size_t>(cl < n ? cl : last) * lda];
            dt[k6][s4] = (row < n && cl < n) ? -v : T(0);
          }
          ++k6;
        }
    __syncthreads();  // the inverted sub-blocks
    acc_t xr[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        xr[tt][r] = (16 * tt + MF::acc_row(lane, r) == 16 * wave + lr) ? T(1) : T(0);
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const int b = UPPER ? 3 - st : st;
      acc_t dd = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4)
        dd = MF::mma(s_d[b][lr][MF::acc_row(lane, s4)], xr[b][s4], dd);
      xr[b] = dd;
      k6 = 0;
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (UPPER ? a < c : a > c) {
            if (c == b) {
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4) xr[a] = MF::mma(dt[k6][s4], xr[b][s4], xr[a]);
            }
            ++k6;
          }
    }
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r) s_m[16 * tt + MF::acc_row(lane, r)][16 * wave + lr] = xr[tt][r];
  }
  __syncthreads();  // the inverse, every column group
  const int nj = UPPER ? nblk - 1 - i : i;  // blocks solved before this one
  const int jp = UPPER ? i + 1 : i - 1;     // the one solved just before

  // 2. M = D^{-1} (-T_{i,jp}), the coupling to the block solved just before: this wave's 16
  // rows of the inverse (A operand k-step s: column 4s + lk) times T_{i,jp} (B operand k-step
  // s: row 4s + lk of this block, 16 columns of block jp per tile; entries outside [0, n)
  // zero), into this wave's rows of s_c.
  if (nj > 0) {
#pragma unroll 1
    for (int ct = 0; ct < 4; ++ct) {
      acc_t mres = {T(0), T(0), T(0), T(0)};
      const int cl = 64 * jp + 16 * ct + lr;
#pragma unroll
      for (int s4 = 0; s4 < 16; ++s4) {
        const int row = r0 + 4 * s4 + lk;
        const T v = lu[static_cast<size_t>(row < n ? row : last) +
                       static_cast<size_t>(cl < n ? cl : last) * lda];
        mres = MF::mma(s_m[16 * wave + lr][4 * s4 + lk], (row < n && cl < n) ? -v : T(0), mres);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) s_c[16 * wave + MF::acc_row(lane, r)][16 * ct + lr] = mres[r];
    }
  }

  // 3. This wave's 16 rows of the right-hand side, as an MFMA result tile; contributions go
  // to four accumulators by k-step (four independent MFMA chains), summed at the end.
  const int trow = r0 + 16 * wave;
  acc_t acc[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    acc[0][r] =
        x[static_cast<size_t>(trow + MF::acc_row(lane, r)) + static_cast<size_t>(col) * ldx];
    acc[1][r] = acc[2][r] = acc[3][r] = T(0);
  }
  bool late = false;
  // Waits until block j's words carry this launch's epoch; its B operand (k-step s4: row
  // 4 s4 + lk of block j, column lr) lands in lo/hi.  A block's words lie part by part (all low
  // halves, then all high halves), rows of 16 columns within a part, so one load or store
  // instruction covers 512 contiguous bytes.  Off the critical path (critical = false) the wait
  // first spins on the one word that block j's last store instruction writes (row 63, column
  // 15, last part): a single transaction per wave, so the waiting workgroups do not load every
  // word through the coherence point that the stores of the block being solved go through.
  auto wait_block = [&](int j, bool critical, unsigned(&lo)[16], unsigned(&hi)[16]) {
    const unsigned long long *wj = xw + static_cast<size_t>(j) * (64 * 16 * WPV) + lk * 16 + lr;
    if (!critical) {
      const unsigned long long *wlast =
          xw + static_cast<size_t>(j) * (64 * 16 * WPV) + ((WPV - 1) * 64 + 63) * 16 + 15;
      while (!late) {
        const unsigned long long w =
            __hip_atomic_load(wlast, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (static_cast<unsigned>(w >> 32) == epoch) break;
        if (static_cast<long long>(wall_clock64()) - t0 > spin_limit) break;  // reported below
        __builtin_amdgcn_s_sleep(1);
      }
    }
    for (;;) {
      // Tag mismatches OR-ed together (no short circuit, so every load stays unconditional).
      unsigned stale = 0u;
#pragma unroll
      for (int s4 = 0; s4 < 16; ++s4) {
        const unsigned long long w0 =
            __hip_atomic_load(wj + 4 * 16 * s4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        lo[s4] = static_cast<unsigned>(w0);
        stale |= static_cast<unsigned>(w0 >> 32) ^ epoch;
        if (WPV == 2) {
          const unsigned long long w1 = __hip_atomic_load(
              wj + 64 * 16 + 4 * 16 * s4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          hi[s4] = static_cast<unsigned>(w1);
          stale |= static_cast<unsigned>(w1 >> 32) ^ epoch;
        } else {
          hi[s4] = 0u;
        }
      }
      if (__all(stale == 0u)) break;
      if (late || static_cast<long long>(wall_clock64()) - t0 > spin_limit) {
        if (!late && lane == 0) atomicExch(err, 1);
        late = true;
        break;
      }
    }
  };

  // Contributions of the blocks solved before jp, as their words arrive (the A operands of the
  // next block loaded ahead).
  const int arow = trow + lr < n ? trow + lr : last;
  const T *arow_p = lu + static_cast<size_t>(arow);
  T a_next[16];
  auto load_a = [&](int j, T(&dst)[16]) {
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) {
      const int kc = 64 * j + 4 * s4 + lk;
      dst[s4] = -arow_p[static_cast<size_t>(kc < n ? kc : last) * lda];
    }
  };
  if (nj > 1) load_a(UPPER ? nblk - 1 : 0, a_next);
  for (int d = 0; d + 1 < nj; ++d) {
    const int j = UPPER ? nblk - 1 - d : d;
    T a_cur[16];
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) a_cur[s4] = a_next[s4];
    if (d + 2 < nj) load_a(UPPER ? j - 1 : j + 1, a_next);
    unsigned lo[16], hi[16];
    wait_block(j, false, lo, hi);
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4)
      acc[s4 & 3] = MF::mma(a_cur[s4], CW::join(lo[s4], hi[s4]), acc[s4 & 3]);
  }

  // 4. y = D^{-1} (right-hand side after those contributions): the 64 rows (past n zeroed) meet
  // in LDS and each wave multiplies its rows of the inverse by them.
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int rr = 16 * wave + MF::acc_row(lane, r);
    const T v = (acc[0][r] + acc[1][r]) + (acc[2][r] + acc[3][r]);
    s_acc[rr][lr] = (r0 + rr < n) ? v : T(0);
  }
  __syncthreads();
  acc_t xp[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) xp[q] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int s4 = 0; s4 < 16; ++s4)
    xp[s4 & 3] =
        MF::mma(s_m[16 * wave + lr][4 * s4 + lk], s_acc[4 * s4 + lk][lr], xp[s4 & 3]);

  // 5. The chain step: x_i = y + M x_jp as soon as block jp's words arrive.
  if (nj > 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) xp[0][r] = (xp[0][r] + xp[1][r]) + (xp[2][r] + xp[3][r]);
    xp[1] = xp[2] = xp[3] = acc_t{T(0), T(0), T(0), T(0)};
    T am[16];  // this wave's rows of M (its own rows of s_c), read before the wait
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) am[s4] = s_c[16 * wave + lr][4 * s4 + lk];
    unsigned lo[16], hi[16];
    wait_block(jp, true, lo, hi);
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4)
      xp[s4 & 3] = MF::mma(am[s4], CW::join(lo[s4], hi[s4]), xp[s4 & 3]);
  }

  // 6. Publish: tagged words for the blocks that wait on this one, plain values for the next
  // kernel.
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = trow + MF::acc_row(lane, r);
    const T v = (xp[0][r] + xp[1][r]) + (xp[2][r] + xp[3][r]);
    unsigned long long *w =
        xw + static_cast<size_t>(i) * (64 * 16 * WPV) + (row - r0) * 16 + lr;
#pragma unroll
    for (int p = 0; p < WPV; ++p)
      __hip_atomic_store(w + p * 64 * 16,
                         (static_cast<unsigned long long>(epoch) << 32) | CW::part(v, p),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lr < nrhs) x[static_cast<size_t>(row) + static_cast<size_t>(lr) * ldx] = v;
  }
}

// ---------------------------------------------------------------------------------------------
// R = B - A X, FP64, for 64 rows x 16 columns of R per workgroup.  The product reads A once
// (it is bound by HBM), so the workgroup takes 64 rows, whose four 16-row MFMA tiles cover 512
// contiguous bytes of each column of A, and each X operand feeds the four tiles.  Sixteen
// waves take interleaved k-steps of v_mfma_f64_16x16x4_f64 (measured at n = 16384: 4 waves
// 2.5 TB/s, 8 waves 4.5, 16 waves 5.7) and their partial tiles are summed in LDS in a fixed
// order, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------------------------
constexpr int kResRows = 64;
constexpr int kResWaves = 16;
__global__ __launch_bounds__(64 * kResWaves) void residual_kernel(
    const double *__restrict__ a, int lda, const double *__restrict__ xm, int ldx,
    const double *__restrict__ b, int ldb, double *__restrict__ r, int ldr, int n, int nrhs) {
  typedef Mfma<double> MF;
  __shared__ double part[kResWaves][4][64][4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r0 = blockIdx.x * kResRows, c0 = blockIdx.y * 16;
  const int lr = lane & 15, lk = lane >> 4;
  const int col = c0 + lr;
  const bool col_ok = col < nrhs;
  const double *xp = xm + static_cast<size_t>(col_ok ? col : 0) * ldx;
  const double *ap[4];
  bool rok[4];
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) {
    const int row = r0 + 16 * tt + lr;
    rok[tt] = row < n;
    ap[tt] = a + static_cast<size_t>(rok[tt] ? row : 0);
  }
  MF::acc_t acc[4];
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) acc[tt] = MF::acc_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
  for (int k = 4 * wave; k < n; k += 4 * kResWaves) {
    const int kk = k + lk;
    const bool ok = kk < n;
    const double xv = (col_ok && ok) ? xp[kk] : 0.0;
    double av[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt)
      av[tt] = (rok[tt] && ok) ? ap[tt][static_cast<size_t>(kk) * lda] : 0.0;
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) acc[tt] = MF::mma(av[tt], xv, acc[tt]);
  }
#pragma unroll
  for (int tt = 0; tt < 4; ++tt)
#pragma unroll
    for (int q = 0; q < 4; ++q) part[wave][tt][lane][q] = acc[tt][q];
  __syncthreads();
  if (t < 256) {  // one output tile per wave of the first four, the partials in order
    const int tt = t >> 6, oc = c0 + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < kResWaves; ++w) s += part[w][tt][lane][q];
      const int orow = r0 + 16 * tt + MF::acc_row(lane, q);
      if (orow < n && oc < nrhs)
        r[static_cast<size_t>(orow) + static_cast<size_t>(oc) * ldr] =
            b[static_cast<size_t>(orow) + static_cast<size_t>(oc) * ldb] - s;
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

// Hands out the ticket bases and epochs of the next nslots chain launches (each of `blocks`
// tickets) and zeroes zero[0, nzero).  Every earlier chain launch on the stream has finished,
// so the counter holds its final value and no reader is left on the words.  When the epochs
// would wrap, the words are cleared and the epochs start again from 1.
__global__ __launch_bounds__(256) void chain_clock_kernel(ChainClocks *st, int nslots, int blocks,
                                                          unsigned long long *words,
                                                          size_t nwords, double *zero,
                                                          int nzero) {
  const unsigned e0 = st->epoch;
  const unsigned long long t0 = st->ticket;
  const bool wrap = e0 > 0xFFFFFFFFu - static_cast<unsigned>(nslots);
  if (wrap)
    for (size_t i = threadIdx.x; i < nwords; i += 256) words[i] = 0ull;
  for (int i = threadIdx.x; i < nzero; i += 256) zero[i] = 0.0;
  __syncthreads();  // every thread has read st->epoch before thread 0 moves it on
  if (threadIdx.x == 0) {
    const unsigned e = wrap ? 0u : e0;
    for (int j = 0; j < nslots; ++j) {
      st->slot[j].base = t0 + static_cast<unsigned long long>(j) * blocks;
      st->slot[j].epoch = e + 1u + static_cast<unsigned>(j);
    }
    st->epoch = e + static_cast<unsigned>(nslots);
  }
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
hipError_t launch_trsv_chain(const T *lu, int lda, int n, T *x, int ldx, int nrhs, bool upper,
                             const ChainState &cs, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs > 16 || ldx < (n + 63) / 64 * 64) return hipErrorInvalidValue;
  const int nblk = (n + 63) / 64;
  if (upper)
    hipLaunchKernelGGL((trsv_chain_kernel<T, true>), dim3(nblk), dim3(256), 0, s, lu, lda, n, x,
                       ldx, nrhs, cs.words, cs.ticket, cs.ticket_base, cs.epoch, cs.clock,
                       cs.spin_limit, cs.err);
  else
    hipLaunchKernelGGL((trsv_chain_kernel<T, false>), dim3(nblk), dim3(256), 0, s, lu, lda, n, x,
                       ldx, nrhs, cs.words, cs.ticket, cs.ticket_base, cs.epoch, cs.clock,
                       cs.spin_limit, cs.err);
  return hipGetLastError();
}

hipError_t launch_chain_clock(ChainClocks *st, int nslots, int blocks, unsigned long long *words,
                              size_t nwords, double *zero, int nzero, hipStream_t s) {
  if (nslots < 0 || nslots > kChainSlots || blocks < 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(chain_clock_kernel, dim3(1), dim3(256), 0, s, st, nslots, blocks, words,
                     nwords, zero, nzero);
  return hipGetLastError();
}

hipError_t launch_residual(const double *a, int lda, const double *x, int ldx, const double *b,
                           int ldb, double *r, int ldr, int n, int nrhs, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  hipLaunchKernelGGL(residual_kernel, dim3((n + kResRows - 1) / kResRows, (nrhs + 15) / 16),
                     dim3(64 * kResWaves), 0, s, a, lda, x, ldx, b, ldb, r, ldr, n, nrhs);
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
template hipError_t launch_trsv_chain<float>(const float *, int, int, float *, int, int, bool,
                                             const ChainState &, hipStream_t);
template hipError_t launch_trsv_chain<double>(const double *, int, int, double *, int, int, bool,
                                              const ChainState &, hipStream_t);

}  // namespace dlu
