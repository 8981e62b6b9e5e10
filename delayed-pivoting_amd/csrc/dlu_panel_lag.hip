// Panel factorization of the delayed-pivoting LU on gfx950 with one exchange per pivot step.
//
// Reference behaviour restated here (delayed_lu/delayed_dgetrf.cpp):
//   column pivot by first-index argmax of |A(k:m-1,k)|          :170-172
//   acceptance test !(alpha > 0) (zero or NaN column -> stop)   :177-178, 182
//   whole-row interchange                                       :202-205
//   multipliers = column * (1 / pivot)                          :221-228
//   rank-1 update restricted to the panel's columns             :230-239
//   the pre-placed pivot of a restarted panel                   :165-167
//
// Layout.  As in the exchange-per-step panel kernel: workgroup g holds the R = 64*RPL rows
// [pb + g*R, pb + (g+1)*R) of the panel in registers, wave w the 16 columns [16w, 16w + 16), lane
// l the rows pb + g*R + 64*r + l (r < RPL).
//
// Rows never move during the panel.  Each held row carries its logical position (the row it
// occupies after the interchanges so far); an interchange of rows k and p only swaps two
// positions.  A row chosen as pivot is final (its L part, the pivot and its U part) and is
// written to row k straight away; the other rows go to their final positions when the panel
// ends, which also yields the panel's net row movement.  So no row data crosses workgroups for an
// interchange, and the only data a step needs from elsewhere is the pivot row.
//
// Roles.  Wave 0 leads: it keeps a private window of every held row -- W0 = column t and W1 =
// column t+1 up to date through step t-1, W2 = column t+2 through step t-2, and Lp = the row's
// multiplier of step t-1 -- searches W0, and publishes one header: the local candidate's W0,
// W1, W2, Lp, logical position and (for the reported margins) the workgroup's second-largest
// magnitude.  It then reads every workgroup's header, picks the winner (largest magnitude,
// lowest logical position on ties: BLAS idamax semantics) and advances its window.  The
// winner's W2 is one step behind; the missing term, l(t-1) * U(t-1, t+2), comes from the pivot
// row of step t-1, which the owner waves fetched while the header travelled.  So a step costs
// one exchange round trip on the critical path.
//
// Every wave (wave 0 included) owns 16 columns and lags the leader by one step: after the
// leader has decided step t, the owners fetch the winner's row (its U part), apply step t's
// rank-1 update to their columns with the leader's multipliers, hand the leader the window
// column it needs next, and publish their part of the next candidate row.  Each step has one
// workgroup barrier; the leader's multipliers and candidate reach the owners through LDS flags.
//
// Every exchange word carries the tag of the step that wrote it (LL protocol: header words hold
// 48 payload bits under a 16-bit tag, row words 32 under a 32-bit tag), so readers spin on
// relaxed device-scope loads and no cache maintenance sits on the per-step path.  All copies of
// a value (the leader's window, the owners' registers, the published row) come from the same
// fma sequence, so they agree bit for bit.  Every wait is bounded (PanelParams::spin_limit).
#include <climits>

#include "dlu_internal.hpp"
#include "dlu_wave.hpp"

namespace dlu {
namespace {

// Header stride per workgroup and step parity (in words).
constexpr int kHdrWords = 8;

__device__ __forceinline__ uint64_t ld_word(const uint64_t *p) {
  return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_word(uint64_t *p, uint64_t w) {
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t *hdr_slot(uint64_t *slots, int par, int g) {
  return slots + (static_cast<size_t>(par) * kLagMaxGroups + g) * kHdrWords;
}
template <int NB>
__device__ __forceinline__ uint64_t *row_slot(uint64_t *slots, int q, int g) {
  return slots + 2 * kLagMaxGroups * kHdrWords +
         (static_cast<size_t>(q) * kLagMaxGroups + g) * (2 * NB);
}

// A 16-byte device-coherent load (sc1, the cache policy of an agent-scope atomic load) from the
// header area: one request for two tagged words.  Each 8-byte half carries its own tag, so a
// value that arrives in pieces is seen as stale, not as wrong.
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void ld_pair(__amdgpu_buffer_rsrc_t rsrc, int byte_off, uint64_t &w0,
                                        uint64_t &w1) {
  const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, byte_off, 0, 16 /* sc1 */);
  w0 = (static_cast<uint64_t>(v[1]) << 32) | v[0];
  w1 = (static_cast<uint64_t>(v[3]) << 32) | v[2];
}

// A T as 32-bit halves and back.
__device__ __forceinline__ void put_halves(double x, unsigned *h) {
  h[0] = lo32(x);
  h[1] = hi32(x);
}
__device__ __forceinline__ void put_halves(float x, unsigned *h) { h[0] = __float_as_uint(x); }
template <typename T>
__device__ __forceinline__ T get_halves(const unsigned *h);
template <>
__device__ __forceinline__ double get_halves<double>(const unsigned *h) {
  return mk(h[0], h[1]);
}
template <>
__device__ __forceinline__ float get_halves<float>(const unsigned *h) {
  return __uint_as_float(h[0]);
}

// Header payload: 32-bit halves packed three to a pair of 48-bit words.
template <int NHALF>
__device__ __forceinline__ uint64_t pack_word(const unsigned *h, int j, unsigned tag) {
  const int p = j >> 1;
  const uint64_t a = h[3 * p], b = h[3 * p + 1], c = h[3 * p + 2];
  const uint64_t w = (j & 1) ? ((b >> 16) | (c << 16)) : (a | ((b & 0xFFFFull) << 32));
  return (static_cast<uint64_t>(tag & 0xFFFFu) << 48) | w;
}
template <int NHALF>
__device__ __forceinline__ void unpack_words(const uint64_t *w, unsigned *h) {
#pragma unroll
  for (int p = 0; p < NHALF / 3; ++p) {
    const uint64_t x = w[2 * p], y = w[2 * p + 1];
    h[3 * p] = static_cast<unsigned>(x);
    h[3 * p + 1] = static_cast<unsigned>((x >> 32) & 0xFFFFull) |
                   (static_cast<unsigned>(y & 0xFFFFull) << 16);
    h[3 * p + 2] = static_cast<unsigned>(y >> 16);
  }
}

// Element `row` of a column whose base address is wave-uniform: a 32-bit byte offset on a scalar
// base, so a column's loads or stores need one address register per row, not one per element.
template <typename T>
__device__ __forceinline__ T &col_at(T *colp, int row) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(colp) +
                                static_cast<unsigned>(row) * static_cast<unsigned>(sizeof(T)));
}

// Rows per lane an owner wave keeps in registers; the rest (FP64 beyond two) live in LDS, which
// keeps the owners within the register share of a nine-wave workgroup.
template <typename T, int RPL>
struct OwnerRows {
  static constexpr int regs = (sizeof(T) == 8 && RPL > 2) ? 2 : RPL;
  static constexpr int lds = RPL - regs;
};

template <typename T, int NB, int RPL>
struct LagShared {
  static constexpr int RL = OwnerRows<T, RPL>::lds;
  // Owner rows kept in LDS: [wave][row - regs][column][lane], one 64-lane run per column.
  T vl[RL > 0 ? NB / 16 : 1][RL > 0 ? RL : 1][RL > 0 ? 16 : 1][RL > 0 ? 64 : 1];
  T l[2][64 * RPL];      // multipliers of step t (leader -> owners), by step parity
  T h[2][64 * RPL];      // window column handed to the leader (owners -> leader)
  T u[2][NB];            // pivot row of step t (the owners' fetch; the leader reads one value)
  T stage[NB / 16][16];  // a candidate row slice, staged for the publishing lanes
  int pos[64 * RPL];     // final logical position of each held row (+ 2^30 once eliminated)
  int cand[2];           // local index of this workgroup's candidate (-1: none)
  int win[2][4];         // the step's winner: workgroup, logical position, local index, stop
  int flag_l, flag_c;    // last step whose multipliers / candidate are in LDS
  int fail;
  int hist[NB];
  double g1[NB], g2[NB];
  unsigned long long ph[9];
};

// True once another workgroup aborted or this wait ran past the bound (which then aborts all).
__device__ __forceinline__ bool expired(Control *ctl, long long spin_limit, long long t0) {
  if (__hip_atomic_load(&ctl->abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return true;
  if (static_cast<long long>(wall_clock64()) - t0 > spin_limit) {
    __hip_atomic_store(&ctl->abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicMax(&ctl->error, 2);
    return true;
  }
  return false;
}

// Bounded wait until an LDS step counter reaches `target` (false: expired, or a failure).
__device__ __forceinline__ bool lds_wait(int *flag, int target, int *fail, Control *ctl,
                                         long long spin_limit, long long t0) {
  for (int it = 1;; ++it) {
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      return true;
    }
    if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return false;
    if ((it & 63) == 0 && expired(ctl, spin_limit, t0)) {
      __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      return false;
    }
  }
}

__device__ __forceinline__ void lds_signal(int *flag, int value) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __hip_atomic_store(flag, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T, int NB, int RPL>
struct Lag {
  // Sixteen columns of one row as a register vector (a select over its elements by a uniform
  // column index stays in registers).
  typedef T row16_t __attribute__((ext_vector_type(16)));
  static constexpr int WPV = Words<T>::n;
  static constexpr int R = 64 * RPL;
  static constexpr int RR = OwnerRows<T, RPL>::regs;  // rows in registers (v)
  static constexpr int RL = OwnerRows<T, RPL>::lds;   // rows in LDS (vl)

  PanelParams<T> P;  // a copy: a reference would keep the kernel arguments in scratch memory
  LagShared<T, NB, RPL> &S;
  int lane, wave, g, base, width, col0;
  long long t0;

  __device__ unsigned seq(int t) const { return P.seq_base + static_cast<unsigned>(t) + 1u; }
  // Row RR + rr of this lane, column j, in LDS.
  __device__ __forceinline__ T &vl(int rr, int j) { return S.vl[wave][rr][j][lane]; }

  // The owners' part of step t on this wave's columns: fetch the pivot row's U part, write the
  // pivot row out when this workgroup holds it, apply the rank-1 update with the leader's
  // multipliers, and hand the leader column t+4 (through step t).
  __device__ __forceinline__ void apply(int t, row16_t (&v)[RR]) {
    const int gwin = S.win[t & 1][0];
    const bool has_u = col0 + 15 > t;  // a column of this wave lies right of t
    uint64_t got = 0;
    if (has_u) {  // loads first: the stores below would hold up the wait for them otherwise
      const bool mine = lane < 16 * WPV;
      const uint64_t *src = row_slot<NB>(P.slots, t & 3, gwin) + WPV * col0 + lane;
      const unsigned sq = seq(t);
      for (int it = 1;; ++it) {
        if (mine) got = ld_word(src);
        if (__all(!mine || static_cast<unsigned>(got >> 32) == sq)) break;
        if ((it & 31) == 0 && expired(P.ctl, P.spin_limit, t0)) {
          __hip_atomic_store(&S.fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          return;
        }
      }
    }
    // The pivot row (through step t-1) is this workgroup's candidate, still staged in LDS.
    if (gwin == g && col0 + 15 >= t && lane < 16 && col0 + lane < width)
      P.a[static_cast<size_t>(P.pb + t) + static_cast<size_t>(P.pb + col0 + lane) * P.lda] =
          S.stage[wave][lane];
    if (has_u) {
      if (lane < 16 * WPV)
        reinterpret_cast<unsigned *>(&S.u[t & 1][col0])[lane] = static_cast<unsigned>(got);
      wave_sync_lds();
    }
    const T *u = &S.u[t & 1][col0];  // read column by column below (stale unless has_u: unused)
    const int jp = t - col0;
    if (jp >= 16) return;
    // The update waits until the leader's next header is out, so that its fmas share the SIMD
    // with the leader's wait rather than with its search.
    if (t + 1 < width && !lds_wait(&S.flag_c, t + 1, &S.fail, P.ctl, P.spin_limit, t0)) return;
    if (!lds_wait(&S.flag_l, t, &S.fail, P.ctl, P.spin_limit, t0)) return;
    T l[RR];  // the LDS rows read theirs where they use it
#pragma unroll
    for (int r = 0; r < RR; ++r) l[r] = S.l[t & 1][64 * r + lane];
    const T *lx = &S.l[t & 1][64 * RR + lane];
    if (jp < 0) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const T uj = u[j];
#pragma unroll
        for (int r = 0; r < RR; ++r) v[r][j] = fma(-l[r], uj, v[r][j]);
#pragma unroll
        for (int rr = 0; rr < RL; ++rr) vl(rr, j) = fma(-lx[64 * rr], uj, vl(rr, j));
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const T uj = u[j];
#pragma unroll
        for (int r = 0; r < RR; ++r) {
          const T f = fma(-l[r], uj, v[r][j]);
          v[r][j] = (j == jp) ? l[r] : ((j > jp) ? f : v[r][j]);
        }
#pragma unroll
        for (int rr = 0; rr < RL; ++rr) {
          if (j > jp) vl(rr, j) = fma(-lx[64 * rr], uj, vl(rr, j));
          if (j == jp) vl(rr, j) = lx[64 * rr];
        }
      }
    }
    const int th = t + 4;
    if (th < width && (th >> 4) == wave) {
      const int jh = th & 15;
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        T x = v[r][0];
#pragma unroll
        for (int j = 1; j < 16; ++j) x = (j == jh) ? v[r][j] : x;
        S.h[(t + 1) & 1][64 * r + lane] = x;
      }
#pragma unroll
      for (int rr = 0; rr < RL; ++rr) S.h[(t + 1) & 1][64 * (RR + rr) + lane] = vl(rr, jh);
    }
  }

  // Stage this wave's slice of the candidate row of step t (through step t-1) in LDS and publish
  // it when it holds columns right of t.
  __device__ __forceinline__ void publish(int t, const row16_t (&v)[RR]) {
    if (t >= width || col0 + 15 < t) return;
    if (!lds_wait(&S.flag_c, t, &S.fail, P.ctl, P.spin_limit, t0)) return;
    const int cl = S.cand[t & 1];
    if (cl < 0) return;
    // (a per-lane test of the held row: a test of the uniform index alone lets the compiler
    // index the register array dynamically, through scratch memory)
#pragma unroll
    for (int r = 0; r < RR; ++r)
      if (64 * r + lane == cl) {
#pragma unroll
        for (int j = 0; j < 16; ++j) S.stage[wave][j] = v[r][j];
      }
#pragma unroll
    for (int rr = 0; rr < RL; ++rr)
      if (64 * (RR + rr) + lane == cl) {
#pragma unroll
        for (int j = 0; j < 16; ++j) S.stage[wave][j] = vl(rr, j);
      }
    wave_sync_lds();
    if (col0 + 15 > t && lane < 16 * WPV) {
      const int j = lane / WPV, part = lane % WPV;
      const T x = S.stage[wave][j];
      st_word(row_slot<NB>(P.slots, t & 3, g) + WPV * col0 + lane,
              (static_cast<uint64_t>(seq(t)) << 32) | Words<T>::part(x, part));
    }
  }
};

// INSTR: the launch may track margins or time its phases (P.margins, P.phase_ticks); the plain
// variant compiles neither, which keeps the leader's step short.
template <typename T, int NB, int RPL, int GQ, bool INSTR>
__global__ __launch_bounds__(NB * 4 + 64) void panel_lag_kernel(PanelParams<T> P) {
  typedef typename Lag<T, NB, RPL>::row16_t row16_t;
  constexpr int WPV = Words<T>::n;
  constexpr int R = 64 * RPL;
  // Header halves, three per pair of words: W0 and the logical position first (the first 16
  // bytes are all the search needs), then W1, W2, Lp, and (reported margins only) the
  // second-largest key at HM, HM+1.
  constexpr int HP = WPV;
  constexpr int H1 = WPV + 1, H2 = 2 * WPV + 1, HL = 3 * WPV + 1;
  constexpr int HM = WPV == 2 ? 9 : 6;
  constexpr int NHALF = HM + 3;
  constexpr int NCORE = 2 * HM / 3;  // words without the margin key
  constexpr int NHW = NCORE + 2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = blockIdx.x;
  const int base = P.pb + g * R;
  const int width = P.pe - P.pb;
  const int col0 = wave * 16;
  const int m = P.m;
  const bool margins = INSTR && P.margins != nullptr;

  __shared__ LagShared<T, NB, RPL> S;

  constexpr int NW = NB / 16;  // owner waves; wave NW leads
  constexpr int RR = OwnerRows<T, RPL>::regs;
  row16_t v[RR];
  if (wave < NW) {
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int row = base + 64 * r + lane;
      T x[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        T *colp = P.a + static_cast<size_t>(P.pb + col0 + j) * P.lda;  // uniform
        x[j] = (row < m && col0 + j < width) ? col_at(colp, row) : T(0);
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        if (r < RR) v[r < RR ? r : 0][j] = x[j];
        else S.vl[wave][r - RR][j][lane] = x[j];
      }
      if (wave == 0) {  // the leader's first window: columns 0-2, and column 3 handed over
        S.l[0][64 * r + lane] = x[0];
        S.l[1][64 * r + lane] = x[1];
        S.h[1][64 * r + lane] = x[2];
        S.h[0][64 * r + lane] = x[3];
      }
    }
  }
  if (threadIdx.x == 0) {
    S.flag_l = -1;
    S.flag_c = -1;
    S.fail = 0;
  }
  // Nothing to factor: a speculative restart after a deferral, or a launch queued behind a panel
  // that did not eliminate all its columns.
  const bool deferred = P.ctl->deferred != 0 || skipped(P.pred);
  __syncthreads();
  if (deferred) {
    if (g == 0 && threadIdx.x == 0) {
      P.ctl->stop = P.pb;
      P.ctl->zero_column = 0;
      P.swap_list[4 * NB] = 0;
      publish_done(P.ctl, P.done, P.done_seq);
    }
    return;
  }

  const long long t0 = static_cast<long long>(wall_clock64());
  Lag<T, NB, RPL> own{P, S, lane, wave, g, base, width, col0, t0};
  int stop = P.pe;
  int zero_col = 0;
  bool failed = false;

  if (wave < NW) {
    // ---- owners -------------------------------------------------------------------------
    own.publish(0, v);
    __syncthreads();
    for (int t = 0;; ++t) {
      // The leader's verdict on step t (1: zero column, 2: a wait failed somewhere): every wave
      // reads the same word, which nothing rewrites before the next barrier.
      const int verdict = S.win[t & 1][3];
      if (verdict == 2) {
        failed = true;
        break;
      }
      if (verdict == 1) {
        stop = P.pb + t;
        zero_col = 1;
        break;
      }
      own.apply(t, v);
      own.publish(t + 1, v);
      if (t + 1 == width) break;
      __syncthreads();
    }
    __syncthreads();  // the leader's final positions are in LDS
    if (!failed && !S.fail) {
      // Rows not yet written: every row still in the search, and the L-only slices of the pivot
      // rows (a slice holding the pivot or U columns went out when its row was chosen).
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int e = S.pos[64 * r + lane];
        const bool done = e >= (1 << 30);
        const int row = done ? e - (1 << 30) : e;
        const bool out = row < m && (!done || col0 + 15 < row - P.pb);
        if (out) {
#pragma unroll
          for (int j = 0; j < 16; ++j)
            if (col0 + j < width)
              col_at(P.a + static_cast<size_t>(P.pb + col0 + j) * P.lda, row) =
                  r < RR ? v[r < RR ? r : 0][j] : S.vl[wave][r < RR ? 0 : r - RR][j][lane];
        }
      }
    }
  } else {
    // ---- leader ----------------------------------------------------------------------------
    // Step t: search column t (after the winner of step t-1 is known), barrier (the owners have
    // finished step t-2, so the candidate's lagged window values are in LDS), header out, then the
    // rest of the window and the owners' inputs while the header travels, then every header in
    // and the winner of step t.
    __builtin_amdgcn_s_setprio(2);
    // Per-lane values of the RPL held rows as register vectors: the candidate's element is read
    // by a uniform index, which a vector keeps in registers (an array would go to scratch).
    typedef T tvec __attribute__((ext_vector_type(RPL)));
    typedef int ivec __attribute__((ext_vector_type(RPL)));
    tvec W0, W1, W2, Lp;  // the window of the step whose header goes out next
    ivec pos;
    unsigned elig = 0;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      W0[r] = S.l[0][64 * r + lane];
      W1[r] = S.l[1][64 * r + lane];
      W2[r] = S.h[1][64 * r + lane];
      Lp[r] = T(0);
      pos[r] = base + 64 * r + lane;
      if (pos[r] < m) elig |= 1u << r;
    }
    const bool timed = INSTR && P.phase_ticks != nullptr && g == 0 && lane == 0;
    if (timed) {
      for (int i = 0; i < 8; ++i) S.ph[i] = 0;
      S.ph[8] = wall_clock64();
    }
#define DLU_PHASE(i)                              \
  if (timed) {                                    \
    const unsigned long long tn = wall_clock64(); \
    S.ph[i] += tn - S.ph[8];                      \
    S.ph[8] = tn;                                 \
  }
    const int G = P.groups;
    // The winner of the last reduced step: 1 / pivot, its W1 (column t), W2 and Lp, position.
    T rcp = T(0), u1 = T(0), w2 = T(0), lw = T(0);
    int gwin = 0, p = 0;
    bool zero_now = false;
    bool owner_failed = false;  // an owner's failure as read while the last header travelled
    for (int t = 0;; ++t) {
      const int k = P.pb + t;
      tvec l, n0;  // multipliers of step t-1; column t through step t-1
      if (t > 0) {
        if (owner_failed) failed = true;
        if (lane == 0) {
          S.win[(t - 1) & 1][0] = gwin;
          S.win[(t - 1) & 1][1] = p;
          S.win[(t - 1) & 1][3] = failed ? 2 : (zero_now ? 1 : 0);
          if (failed) S.fail = 1;
        }
        if (failed || zero_now) {
          __syncthreads();
          break;
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
          l[r] = W0[r] * rcp;
          n0[r] = fma(-l[r], u1, W1[r]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
          l[r] = T(0);
          n0[r] = W0[r];
        }
      }

      // Local candidate: the largest magnitude of column t among the rows not yet eliminated,
      // the lowest logical position among equal magnitudes (NaN never wins).
      uint64_t key[RPL];
      uint64_t lk = 0;
      int cpos = INT_MAX, cr = 0, cl = 0;
      if (t < width) {
        const bool forced = P.positioned && t == 0;
        uint64_t m1 = 0;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
          const bool ok = ((elig >> r) & 1u) && (!forced || pos[r] == k);
          key[r] = mag_key(n0[r], ok);
          m1 = max(m1, key[r]);
        }
        lk = wave_max_key_fast(m1);
        // A single holder of the largest key is the candidate; only ties need the lowest position.
        int holders = 0;
#pragma unroll
        for (int r = 0; r < RPL; ++r) holders += __popcll(__ballot(key[r] == lk));
        if (holders == 1) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) {
            const uint64_t b = __ballot(key[r] == lk);
            if (b != 0) {
              cr = r;
              cl = __ffsll(static_cast<long long>(b)) - 1;
            }
          }
          cpos = __builtin_amdgcn_readlane(pos[cr], cl);
        } else {
          int mp = INT_MAX;
#pragma unroll
          for (int r = 0; r < RPL; ++r) mp = (key[r] == lk) ? min(mp, pos[r]) : mp;
          cpos = wave_min(mp);
          bool found = false;
#pragma unroll
          for (int r = 0; r < RPL; ++r) {
            const uint64_t b = __ballot(pos[r] == cpos);
            if (!found && b != 0) {
              found = true;
              cr = r;
              cl = __ffsll(static_cast<long long>(b)) - 1;
            }
          }
        }
      }
      DLU_PHASE(0)
      if (t > 0) {
        __syncthreads();  // the owners have finished step t-2
        DLU_PHASE(1)
        if (t == width) {  // the last step's multipliers, for the owners
#pragma unroll
          for (int r = 0; r < RPL; ++r) S.l[(t - 1) & 1][64 * r + lane] = l[r];
          lds_signal(&S.flag_l, t - 1);
          break;
        }
      }
      const bool has = lk != 0;
      // Step t-1's update of column t+1 needs U(t-2, t+1) for the lagged W2 (the owners' copy);
      // column t+2 through step t-2 is the owners' hand-over.  Both LDS reads go out together.
      const bool lag = t >= 2 && t + 1 < width;
      const T up = lag ? S.u[(t - 2) & 1][t + 1] : T(0);
      tvec n2;
#pragma unroll
      for (int r = 0; r < RPL; ++r)
        n2[r] = (t > 0 && t + 2 < width) ? S.h[(t - 1) & 1][64 * r + lane] : T(0);
      const T u2 = lag ? fma(-lw, up, w2) : w2;  // U(t-1, t+1)

      // The header: the candidate's window values, computed for it alone first.
      {
        const unsigned sq = P.seq_base + static_cast<unsigned>(t) + 1u;
        const T c0 = lane_value(n0[cr], cl);
        T c1, c2, c3;
        if (t > 0) {
          const T cw2 = lane_value(W2[cr], cl), clp = lane_value(Lp[cr], cl);
          c3 = lane_value(l[cr], cl);
          c1 = fma(-c3, u2, lag ? fma(-clp, up, cw2) : cw2);
          c2 = lane_value(n2[cr], cl);
        } else {
          c1 = lane_value(W1[cr], cl);
          c2 = lane_value(W2[cr], cl);
          c3 = T(0);
        }
        unsigned h[NHALF];
#pragma unroll
        for (int i = 0; i < NHALF; ++i) h[i] = 0;
        put_halves(has ? c0 : T(0), h);
        h[HP] = has ? static_cast<unsigned>(cpos) : static_cast<unsigned>(INT_MAX);
        put_halves(c1, h + H1);
        put_halves(c2, h + H2);
        put_halves(c3, h + HL);
        if (margins) {  // second-largest magnitude of this workgroup's rows
          uint64_t m2 = 0;
#pragma unroll
          for (int r = 0; r < RPL; ++r) m2 = max(m2, pos[r] != cpos ? key[r] : 0ull);
          m2 = wave_max_key(m2);
          h[HM] = static_cast<unsigned>(m2);
          h[HM + 1] = static_cast<unsigned>(m2 >> 32);
        }
        uint64_t mine = 0;
#pragma unroll
        for (int j = 0; j < NHW; ++j) {
          const uint64_t wj = pack_word<NHALF>(h, j, sq);
          mine = (lane == j) ? wj : mine;
        }
        if (lane < (margins ? NHW : NCORE))
          st_word(hdr_slot(P.slots, t & 1, g) + lane, mine);
      }
      DLU_PHASE(2)

      // While the header travels: the rest of the window, then the owners' inputs.
      if (t > 0) {
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
          const T w2c = lag ? fma(-Lp[r], up, W2[r]) : W2[r];
          const T n1 = fma(-l[r], u2, w2c);
          S.l[(t - 1) & 1][64 * r + lane] = l[r];
          W0[r] = n0[r];
          W1[r] = n1;
          W2[r] = n2[r];
          Lp[r] = l[r];
        }
        lds_signal(&S.flag_l, t - 1);
      }
      S.cand[t & 1] = has ? 64 * cr + cl : -1;
      lds_signal(&S.flag_c, t);  // the candidate, and the header is out
      owner_failed =
          __hip_atomic_load(&S.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
      DLU_PHASE(3)

      // Every workgroup's header: lane i reads workgroups i and i + 64.  Each poll loads the
      // whole header (its words arrive together, one store of the writer) and tests the first
      // 16 bytes (the pivot candidate and its position); the winner's other words are then
      // already in registers, not one more device-scope round trip away.
      if (!failed) {
        const bool q0 = lane < G, q1 = GQ > 1 && lane + 64 < G;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            P.slots, 0, 2 * kLagMaxGroups * kHdrWords * 8, 0x00020000);
        const int o0 = ((t & 1) * kLagMaxGroups + lane) * kHdrWords * 8;
        const int o1 = ((t & 1) * kLagMaxGroups + lane + 64) * kHdrWords * 8;
        const unsigned tg = (P.seq_base + static_cast<unsigned>(t) + 1u) & 0xFFFFu;
        const int nw = margins ? NHW : NCORE;
        uint64_t hw0[NHW], hw1[NHW];
#pragma unroll
        for (int j = 0; j < NHW; ++j) hw0[j] = hw1[j] = 0ull;
        auto fresh2 = [&](const uint64_t *w, int j) {
          return static_cast<unsigned>(w[j] >> 48) == tg &&
                 static_cast<unsigned>(w[j + 1] >> 48) == tg;
        };
        for (int it = 1;; ++it) {
#pragma unroll
          for (int j = 0; j < NHW; j += 2)
            if (j < nw) {
              if (q0) ld_pair(rs, o0 + 8 * j, hw0[j], hw0[j + 1]);
              if (q1) ld_pair(rs, o1 + 8 * j, hw1[j], hw1[j + 1]);
            }
          if (__all((!q0 || fresh2(hw0, 0)) && (!q1 || fresh2(hw1, 0)))) break;
          if ((it & 31) == 0 && expired(P.ctl, P.spin_limit, t0)) {
            failed = true;
            break;
          }
        }
        DLU_PHASE(4)
        if (!failed) {
          unsigned a0[3], a1[3];  // pair 0: W0 and the position
          unpack_words<3>(hw0, a0);
          unpack_words<3>(hw1, a1);
          const uint64_t k0 = q0 ? mag_key(get_halves<T>(a0), true) : 0ull;
          const uint64_t k1 = q1 ? mag_key(get_halves<T>(a1), true) : 0ull;
          const int p0 = q0 ? static_cast<int>(a0[HP]) : INT_MAX;
          const int p1 = q1 ? static_cast<int>(a1[HP]) : INT_MAX;
          const bool use1 = k1 > k0 || (k1 == k0 && p1 < p0);
          const uint64_t bk = use1 ? k1 : k0;
          const int bp = use1 ? p1 : p0;
          // 1 / pivot of this lane's better header, formed on every lane while the reduction
          // runs; the winner's lane holds the step's value.
          const T lrcp = T(1) / (use1 ? get_halves<T>(a1) : get_halves<T>(a0));
          const uint64_t gmax = wave_max_key_fast(bk);
          if (gmax == 0) {  // zero (or NaN-only) column: stop the panel at k
            zero_now = true;
            stop = k;
            zero_col = 1;
          } else {
            uint64_t wb = __ballot(bk == gmax);
            if (__popcll(wb) == 1) {  // one workgroup holds the largest magnitude
              p = __builtin_amdgcn_readlane(bp, __ffsll(static_cast<long long>(wb)) - 1);
            } else {
              p = wave_min(bk == gmax ? bp : INT_MAX);
              wb = __ballot(bk == gmax && bp == p);
            }
            const int wlane = __ffsll(static_cast<long long>(wb)) - 1;
            const bool w1 = __builtin_amdgcn_readlane(use1 ? 1 : 0, wlane) != 0;
            gwin = wlane + (w1 ? 64 : 0);
            // The winner's other words (re-read in the rare case they were not yet all fresh).
            for (int it = 1;; ++it) {
              bool ok = true;
#pragma unroll
              for (int j = 2; j < NHW; j += 2)
                if (j < nw) ok = ok && (!q0 || fresh2(hw0, j)) && (!q1 || fresh2(hw1, j));
              if (__all(ok)) break;
#pragma unroll
              for (int j = 2; j < NHW; j += 2)
                if (j < nw) {
                  if (q0) ld_pair(rs, o0 + 8 * j, hw0[j], hw0[j + 1]);
                  if (q1) ld_pair(rs, o1 + 8 * j, hw1[j], hw1[j + 1]);
                }
              if ((it & 31) == 0 && expired(P.ctl, P.spin_limit, t0)) {
                failed = true;
                break;
              }
            }
            uint64_t ww[NHW];
#pragma unroll
            for (int j = 0; j < NHW; ++j) {
              const uint64_t x = w1 ? hw1[j] : hw0[j];
              ww[j] = (static_cast<uint64_t>(static_cast<unsigned>(
                           __builtin_amdgcn_readlane(static_cast<int>(x >> 32), wlane)))
                       << 32) |
                      static_cast<unsigned>(
                          __builtin_amdgcn_readlane(static_cast<int>(x), wlane));
            }
            unsigned x[NHALF];
            unpack_words<NHALF>(ww, x);
            rcp = lane_value(lrcp, wlane);
            u1 = get_halves<T>(x + H1);
            w2 = get_halves<T>(x + H2);
            lw = get_halves<T>(x + HL);
            if (g == 0 && margins) {  // margin: the largest candidate elsewhere
              const uint64_t o0k = (q0 && lane != gwin) ? k0 : 0ull;
              const uint64_t o1k = (q1 && lane + 64 != gwin) ? k1 : 0ull;
              const uint64_t o = wave_max_key(max(o0k, o1k));
              const uint64_t m2w = (static_cast<uint64_t>(x[HM + 1]) << 32) | x[HM];
              if (lane == 0) {
                S.g1[t] = key_value(gmax);
                S.g2[t] = key_value(max(o, m2w));
              }
            }
            // Interchange of logical positions k and p; the winner's row leaves the search.
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
              const bool winner = ((elig >> r) & 1u) && pos[r] == p;
              if (winner) {
                pos[r] = k;
              } else if (pos[r] == k) {
                pos[r] = p;
              }
              elig &= winner ? ~(1u << r) : ~0u;
            }
            if (g == 0 && lane == 0) S.hist[t] = p;
          }
        }
      }
      DLU_PHASE(5)
    }
    if (!failed) {
#pragma unroll
      for (int r = 0; r < RPL; ++r)
        S.pos[64 * r + lane] = pos[r] + (((elig >> r) & 1u) ? 0 : (1 << 30));
    }
    if (timed) {
      DLU_PHASE(7)
      for (int i = 0; i < 8; ++i) P.phase_ticks[i] += S.ph[i];
    }
#undef DLU_PHASE
    __syncthreads();
  }
  if (failed || S.fail) goto leave;
  {
    const int steps = stop - P.pb;
    // Net row movement as (dst, src) pairs: the pivot rows of steps [0, steps) at their step's
    // index, then every other row whose position changed (appended).
    int *dst = P.swap_list;
    int *src = P.swap_list + 2 * NB;
    if (wave == 0) {
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int e = S.pos[64 * r + lane];
        const bool done = e >= (1 << 30);
        const int row = done ? e - (1 << 30) : e;
        const int from = base + 64 * r + lane;
        if (row < m) {
          if (done) {
            dst[row - P.pb] = row;
            src[row - P.pb] = from;
          } else if (row != from) {
            const int at = steps + atomicAdd(&P.swap_list[4 * NB + 1], 1);
            dst[at] = row;
            src[at] = from;
          }
        }
      }
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
      if (atomicAdd(&P.swap_list[4 * NB + 2], 1) == P.groups - 1) {
        __threadfence();
        const int appended = __hip_atomic_load(&P.swap_list[4 * NB + 1], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
        P.swap_list[4 * NB] = steps + appended;
        P.swap_list[4 * NB + 1] = 0;
        P.swap_list[4 * NB + 2] = 0;
      }
    }
    if (g == 0) {
      for (int s = threadIdx.x; s < steps; s += blockDim.x) {
        if (!(P.positioned && s == 0)) P.ipiv[P.pb + s] = S.hist[s];
        if (margins)
          P.margins[P.pb + s] =
              (P.positioned && s == 0) ? -1.0 : (S.g1[s] - fmax(S.g2[s], 0.0)) / S.g1[s];
      }
      if (wave == 0) {  // tie statistics of the column steps (when margins are tracked)
        double lo = 2.0;
        int near = 0, exact = 0;
        for (int s0 = 0; margins && s0 < steps; s0 += 64) {
          const int s = s0 + lane;
          const bool valid = s < steps && !(P.positioned && s == 0);
          const double mg = valid ? (S.g1[s] - fmax(S.g2[s], 0.0)) / S.g1[s] : 2.0;
          lo = fmin(lo, mg);
          near += __popcll(__ballot(valid && mg < 1e-10));
          exact += __popcll(__ballot(valid && mg == 0.0));
        }
        const double min_margin = -wave_max(-lo);
        if (lane == 0) {
          P.ctl->stop = stop;
          P.ctl->zero_column = zero_col;
          if (stop == P.pe && zero_col == 0) P.ctl->go = P.done_seq;
          P.ctl->near_ties += near;
          P.ctl->exact_ties += exact;
          if (min_margin < P.ctl->min_margin) P.ctl->min_margin = min_margin;
          publish_done(P.ctl, P.done, P.done_seq);
        }
      }
    }
  }
  return;

leave:
  if (g == 0 && threadIdx.x == 0) {
    P.ctl->stop = P.pb;
    publish_done(P.ctl, P.done, P.done_seq);
  }
}

template <typename T, int NB, int GQ, bool INSTR>
hipError_t launch_lag_gq(const PanelParams<T> &p, int rows_per_lane, hipStream_t s) {
  const dim3 grid(p.groups), block(NB * 4 + 64);
  switch (rows_per_lane) {
    case 1: hipLaunchKernelGGL((panel_lag_kernel<T, NB, 1, GQ, INSTR>), grid, block, 0, s, p); break;
    case 2: hipLaunchKernelGGL((panel_lag_kernel<T, NB, 2, GQ, INSTR>), grid, block, 0, s, p); break;
    case 3: hipLaunchKernelGGL((panel_lag_kernel<T, NB, 3, GQ, INSTR>), grid, block, 0, s, p); break;
    case 4: hipLaunchKernelGGL((panel_lag_kernel<T, NB, 4, GQ, INSTR>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The leader reads one header per lane up to 64 workgroups, two beyond.
template <typename T, int NB>
hipError_t launch_lag_nb(const PanelParams<T> &p, int rows_per_lane, hipStream_t s) {
  const bool instr = p.margins != nullptr || p.phase_ticks != nullptr;
  if (p.groups <= 64)
    return instr ? launch_lag_gq<T, NB, 1, true>(p, rows_per_lane, s)
                 : launch_lag_gq<T, NB, 1, false>(p, rows_per_lane, s);
  return instr ? launch_lag_gq<T, NB, 2, true>(p, rows_per_lane, s)
               : launch_lag_gq<T, NB, 2, false>(p, rows_per_lane, s);
}

}  // namespace

template <typename T>
hipError_t launch_panel_lag(const PanelParams<T> &p, int nb, int rows_per_lane, hipStream_t s) {
  if (p.groups < 1 || p.groups > kLagMaxGroups) return hipErrorInvalidValue;
  switch (nb) {
    case 64: return launch_lag_nb<T, 64>(p, rows_per_lane, s);
    case 128: return launch_lag_nb<T, 128>(p, rows_per_lane, s);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_panel_lag<float>(const PanelParams<float> &, int, int, hipStream_t);
template hipError_t launch_panel_lag<double>(const PanelParams<double> &, int, int, hipStream_t);

}  // namespace dlu
