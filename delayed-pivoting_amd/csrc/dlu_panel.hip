// Panel factorization of the delayed-pivoting LU on gfx950.
//
// Reference behaviour restated here (delayed_lu/delayed_dgetrf.cpp):
//   column pivot by first-index argmax of |A(k:m-1,k)|          :170-172
//   acceptance test !(alpha > 0) (zero or NaN column -> stop)   :177-178, 182
//   whole-row interchange (inside the panel's columns here)     :202-205
//   multipliers = column * (1 / pivot)                          :221-228
//   rank-1 update restricted to the panel's columns             :230-239
//   the pre-placed pivot of a restarted panel                   :165-167
//
// Layout.  The panel is rows [pb, m) x columns [pb, pe) of the column-major matrix, at most NB
// columns.  It is held in registers for the whole factorization of the panel: workgroup g owns
// the R = 64*RPL rows [pb + g*R, pb + (g+1)*R); wave w of a workgroup owns the 16 columns
// [pb + 16w, pb + 16w + 16); lane l holds rows pb + g*R + 64*r + l (r < RPL) of those columns.
// Besides its own columns, every wave keeps private copies of the next two pivot columns (the
// "window"), updated with the same fma as the owner's copy, so the pivot search of the next step
// needs no hand-over between waves and a step has a single workgroup barrier.
//
// One pivot step costs one all-to-all exchange between the workgroups.  Each workgroup publishes
// its local candidate (|value|, row, second-largest |value|) together with the candidate row's
// NB values; the owner of row k also publishes row k.  Every workgroup then reduces all candidates
// (largest magnitude, lowest row on ties -- BLAS idamax semantics), reads the winner's row and
// row k, performs the interchange on its own rows, scales its part of column k by the reciprocal
// of the pivot and applies the rank-1 update to its part of the panel.  Every exchange word
// carries the 16-bit tag of the step that wrote it next to its payload (an LL-style protocol:
// header words hold 48 payload bits, data words 32), so readers spin on relaxed device-scope
// loads and no cache writeback/invalidate fence sits on the per-step path.  Every wait is
// bounded (PanelParams::spin_limit); a workgroup whose wait expires raises ctl->abort_flag and
// all workgroups leave.
//
// Net row movement.  Each workgroup also keeps, per row it holds, the original index of the row
// whose content sits there (in LDS, read and written only by its last wave); the candidate's and
// row k's origins travel in the same exchange, so at the end every workgroup lists its own moved
// rows in parallel (an atomic append) instead of one lane replaying the pivot history.
#include "dlu_internal.hpp"
#include "dlu_wave.hpp"

#include <climits>

namespace dlu {
namespace {

__device__ __forceinline__ void ll_put(uint64_t *p, unsigned seq, unsigned payload) {
  __hip_atomic_store(p, (static_cast<uint64_t>(seq) << 32) | payload, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void ll_put_double(uint64_t *p, unsigned seq, double v) {
  const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
  ll_put(p, seq, static_cast<unsigned>(bits));
  ll_put(p + 1, seq, static_cast<unsigned>(bits >> 32));
}

__device__ __forceinline__ uint64_t ll_load(const uint64_t *p) {
  return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait for `count` consecutive exchange words tagged `seq`; payloads land in out[].
// Returns false when the wait expired or another workgroup aborted.
template <int COUNT, typename Params>
__device__ bool ll_get(const uint64_t *p, unsigned seq, unsigned *out, const Params &P,
                       long long t0) {
  for (int it = 1;; ++it) {
    bool fresh = true;
#pragma unroll
    for (int i = 0; i < COUNT; ++i) {
      const uint64_t w = ll_load(p + i);
      out[i] = static_cast<unsigned>(w);
      fresh = fresh && (static_cast<unsigned>(w >> 32) == seq);
    }
    if (fresh) return true;
    if ((it & 31) == 0) {
      if (__hip_atomic_load(&P.ctl->abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return false;
      if (static_cast<long long>(wall_clock64()) - t0 > P.spin_limit) {
        __hip_atomic_store(&P.ctl->abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicMax(&P.ctl->error, 2);
        return false;
      }
    }
  }
}

// Header words: 48-bit payload, 16-bit step tag on top.  The candidate's |value| (64 bits) and
// row (32 bits) fit in two words: word 0 = low 48 bits of the value, word 1 = the value's top 16
// bits above the row.
__device__ __forceinline__ void hdr_put(uint64_t *p, unsigned seq, double v, int row) {
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
  const uint64_t tag = static_cast<uint64_t>(seq) << 48;
  __hip_atomic_store(p, tag | (b & 0xFFFFFFFFFFFFull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, tag | ((b >> 48) << 32) | static_cast<unsigned>(row), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait for a fresh header; false when the wait expired or another workgroup aborted.
template <typename Params>
__device__ bool hdr_get(const uint64_t *p, unsigned seq, double &v, int &row, const Params &P,
                        long long t0) {
  for (int it = 1;; ++it) {
    const uint64_t w0 = ll_load(p), w1 = ll_load(p + 1);
    if (static_cast<unsigned>(w0 >> 48) == seq && static_cast<unsigned>(w1 >> 48) == seq) {
      v = __longlong_as_double(
          static_cast<long long>((w0 & 0xFFFFFFFFFFFFull) | (((w1 >> 32) & 0xFFFFull) << 48)));
      row = static_cast<int>(static_cast<unsigned>(w1));
      return true;
    }
    if ((it & 31) == 0) {
      if (__hip_atomic_load(&P.ctl->abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return false;
      if (static_cast<long long>(wall_clock64()) - t0 > P.spin_limit) {
        __hip_atomic_store(&P.ctl->abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicMax(&P.ctl->error, 2);
        return false;
      }
    }
  }
}

__device__ __forceinline__ double join_double(unsigned lo, unsigned hi) {
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

template <typename T>
__device__ __forceinline__ void ll_put_value(uint64_t *p, unsigned seq, T v) {
#pragma unroll
  for (int i = 0; i < Words<T>::n; ++i) ll_put(p + i, seq, Words<T>::part(v, i));
}

// An integer carried bit for bit through the exchange as a T.
__device__ __forceinline__ void as_word(double &x, int i) { x = mk(static_cast<unsigned>(i), 0u); }
__device__ __forceinline__ void as_word(float &x, int i) { x = __int_as_float(i); }
__device__ __forceinline__ int from_word(double x) { return static_cast<int>(lo32(x)); }
__device__ __forceinline__ int from_word(float x) { return __float_as_int(x); }

template <typename T, int NB, int RPL>
__global__ __launch_bounds__(NB * 4) void panel_kernel(PanelParams<T> P) {
  // Sixteen columns of one row, kept as a register vector so that a wave-uniform runtime column
  // index lowers to indirect register moves rather than scratch memory.
  typedef T row16_t __attribute__((ext_vector_type(16)));
  constexpr int WPV = Words<T>::n;  // exchange words per element
  constexpr int R = 64 * RPL;
  constexpr int SW = slot_words(NB);
  constexpr int NW = NB / 16;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = blockIdx.x;
  const int base = P.pb + g * R;
  const int width = P.pe - P.pb;
  const int col0 = wave * 16;
  const int m = P.m;

  // s_win[0..1]: column c+2, handed over by its owner wave during step c (alternating by step);
  // s_win[2..3]: columns 0 and 1 at the start.
  __shared__ T s_win[4][R];
  __shared__ T s_u[NW][16];        // per-wave slice of the pivot row
  __shared__ T s_lp[NW][R];        // per-wave copy of the pending step's multipliers
  __shared__ T s_k[NW][16];        // per-wave slice of the old row k (owner of row p)
  __shared__ T s_pub[NW][2][16];   // per-wave slices of the candidate row and of row k
  __shared__ uint64_t s_ck[2][kMaxPanelGroups];  // candidate keys, alternating by step
  __shared__ int s_crow[2][kMaxPanelGroups];
  __shared__ int s_fail;
  __shared__ int s_hist[NB];
  __shared__ int s_orig[NB];    // per step: original index of the row that became row k
  __shared__ int s_from[R];     // per held row: original index of the content it carries
  __shared__ double s_g1[NB];   // per step: largest and second-largest |value| (workgroup 0)
  __shared__ double s_g2[NB];

  row16_t v[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int row = base + 64 * r + lane;
    if (wave == NW - 1) s_from[64 * r + lane] = row;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int col = col0 + j;
      v[r][j] = (row < m && col < width)
                    ? P.a[static_cast<size_t>(row) + static_cast<size_t>(P.pb + col) * P.lda]
                    : T(0);
    }
    if (wave == 0) {
      s_win[2][64 * r + lane] = v[r][0];
      s_win[3][64 * r + lane] = v[r][1];
    }
  }
  if (threadIdx.x == 0) s_fail = 0;
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
  T wa[RPL], wb[RPL];  // private copies of columns c and c+1
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    wa[r] = s_win[2][64 * r + lane];
    wb[r] = s_win[3][64 * r + lane];
  }

  const long long t0 = static_cast<long long>(wall_clock64());
  int stop = P.pe;
  int zero_col = 0;
  double min_margin = 2.0;
  int near_ties = 0, exact_ties = 0;
  // The rank-1 update of a step reaches the bulk of the panel (all but the window) in the next
  // step, after that step's header is out: the pending multipliers s_lp[wave] and the pivot row
  // slice s_u[wave] (both private to the wave) are applied while the header travels, before the
  // candidate row is staged.
  int jp = 0;         // this wave's column index of the pending step's column (wave-uniform)
  bool pend = false;  // a step's bulk update is pending
#define DLU_APPLY_PENDING()                                                          \
  if (pend) {                                                                        \
    pend = false;                                                                    \
    T u[16];                                                                         \
    _Pragma("unroll") for (int j = 0; j < 16; ++j) u[j] = s_u[wave][j];              \
    _Pragma("unroll") for (int r = 0; r < RPL; ++r) {                                \
      const T l = s_lp[wave][64 * r + lane];                                         \
      if (jp < 0) {                                                                  \
        _Pragma("unroll") for (int j = 0; j < 16; ++j) v[r][j] = fma(-l, u[j], v[r][j]); \
      } else if (jp < 16) {                                                          \
        _Pragma("unroll") for (int j = 0; j < 16; ++j) {                             \
          if (j == jp) v[r][j] = l;                                                  \
          if (j > jp) v[r][j] = fma(-l, u[j], v[r][j]);                              \
        }                                                                            \
      }                                                                              \
    }                                                                                \
  }
  // Margin bookkeeping of the previous step, finished by wave 1 of workgroup 0 while wave 0
  // polls the next step's headers (off the per-step path).
  int prev_gwin = -1;
  double prev_w2 = 0.0;
  // Optional phase timing (workgroup 0, thread 0; 100 MHz ticks).
  const bool timed = P.phase_ticks != nullptr && g == 0 && threadIdx.x == 0;
  // Totals live in LDS (s_ph[0..7], s_ph[8] = last stamp) so they cost no registers.
  __shared__ unsigned long long s_ph[9];
  if (timed) {
    for (int i = 0; i < 8; ++i) s_ph[i] = 0;
    s_ph[8] = wall_clock64();
  }
#define DLU_PHASE(i)                              \
  if (timed) {                                    \
    const unsigned long long tn = wall_clock64(); \
    s_ph[i] += tn - s_ph[8];                      \
    s_ph[8] = tn;                                 \
  }

  for (int c = 0; c < width; ++c) {
    const int k = P.pb + c;
    const unsigned seq = P.seq_base + static_cast<unsigned>(c) + 1u;
    const int par = static_cast<int>(seq & 1u);
    const int cb = c & 1;
    const bool forced = P.positioned && c == 0;
    const bool k_here = k >= base && k < base + R;

    // Local candidate from the private copy of column c: the largest magnitude and the lowest
    // row among equal magnitudes (NaN never wins).
    uint64_t key[RPL];
    bool ok[RPL];
    uint64_t m1 = 0;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int row = base + 64 * r + lane;
      ok[r] = row >= k && row < m && (!forced || row == k);
      key[r] = mag_key(wa[r], ok[r]);
      m1 = max(m1, key[r]);
    }
    const uint64_t lk = wave_max_key(m1);
    int myrow = INT_MAX;
#pragma unroll
    for (int r = RPL - 1; r >= 0; --r)
      myrow = (ok[r] && key[r] == lk) ? base + 64 * r + lane : myrow;
    const int lrow = wave_min(myrow);
    DLU_PHASE(0)

    // The header first: it is what every workgroup polls on.  The candidate row's data words
    // are read only after the reduction, so they follow once the pending update is applied.
    uint64_t *slot = P.slots + (static_cast<size_t>(par) * P.groups + g) * SW;
    uint64_t *rk = P.rowk + static_cast<size_t>(par) * rowk_words(NB);
    if (wave == NW - 1 && lane == 0) hdr_put(slot, seq, key_value(lk), lrow);
    DLU_APPLY_PENDING()

    // Stage this wave's 16 columns of the candidate row and of row k in LDS (so 32 lanes each
    // store one exchange word), and hand column c+2 (current through step c-1) to every wave.
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int row = base + 64 * r + lane;
      if (row == lrow) {
#pragma unroll
        for (int j = 0; j < 16; ++j) s_pub[wave][0][j] = v[r][j];
      }
      if (row == k) {
#pragma unroll
        for (int j = 0; j < 16; ++j) s_pub[wave][1][j] = v[r][j];
      }
    }
    if (c + 2 < width && ((c + 2) >> 4) == wave) {
      const int j2 = (c + 2) & 15;
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j == j2) {
#pragma unroll
          for (int r = 0; r < RPL; ++r) s_win[cb][64 * r + lane] = v[r][j];
        }
    }
    wave_sync_lds();

    // Publish: lanes 0-31 store this wave's slice of the candidate row, lanes 32-63 that of
    // row k, one exchange word each.
    {
      const int j = (lane & 31) / WPV, part = lane % WPV;
      const bool cand_here = lrow >= base && lrow < base + R;
      if (j < 16 && col0 + j < width && (lane < 32 ? cand_here : k_here)) {
        const T x = s_pub[wave][lane >> 5][j];
        uint64_t *dstw = (lane < 32 ? slot + kSlotHeaderWords : rk) + WPV * (col0 + j) + part;
        ll_put(dstw, seq, Words<T>::part(x, part));
      }
    }
    if (wave == NW - 1) {
      // The candidate's origin in the slot's spare words, row k's next to its values.
      T ow;
      if (lane == 1 && lrow >= base && lrow < base + R) {
        as_word(ow, s_from[lrow - base]);
        ll_put_value(slot + 4, seq, ow);
      }
      if (lane == 2 && k_here) {
        as_word(ow, s_from[k - base]);
        ll_put_value(rk + 2 * NB, seq, ow);
      }
      // The second-largest |value| of this workgroup's rows, for the reported margins only
      // (when the caller asked for them): published after the header.
      if (P.margins != nullptr) {
        uint64_t m2 = 0;
#pragma unroll
        for (int r = 0; r < RPL; ++r) m2 = max(m2, (base + 64 * r + lane != lrow) ? key[r] : 0ull);
        m2 = wave_max_key(m2);
        if (lane == 0) ll_put_value(slot + 2, seq, static_cast<T>(key_value(m2)));
      }
    }
    DLU_PHASE(1)

    // Margin of the previous step: the largest candidate outside the winner's workgroup (only
    // when margins were asked for: workgroup 0 is the one most often last to publish its header,
    // and this work delays its gather).
    if (g == 0 && wave == 1 && prev_gwin >= 0 && P.margins != nullptr) {
      constexpr int GQ = kMaxPanelGroups / 64;
      uint64_t o = 0;
#pragma unroll
      for (int q = 0; q < GQ; ++q) {
        const int gi = lane + 64 * q;
        const uint64_t kq = s_ck[cb ^ 1][gi];
        o = max(o, (gi < P.groups && gi != prev_gwin) ? kq : 0ull);
      }
      const double gs = key_value(wave_max_key(o));
      if (lane == 0) s_g2[c - 1] = fmax(gs, prev_w2);
    }

    // Gather every workgroup's header.
    if (threadIdx.x < P.groups) {
      const uint64_t *hs = P.slots + (static_cast<size_t>(par) * P.groups + threadIdx.x) * SW;
      double hv = 0.0;
      int hrow = 0;
      if (!hdr_get(hs, seq, hv, hrow, P, t0)) s_fail = 1;
      s_ck[cb][threadIdx.x] = static_cast<uint64_t>(__double_as_longlong(hv));
      s_crow[cb][threadIdx.x] = hrow;
    }
    __syncthreads();
    DLU_PHASE(2)
    if (s_fail) goto leave;

    {
      T wc[RPL];  // private copy of column c+2
#pragma unroll
      for (int r = 0; r < RPL; ++r) wc[r] = (c + 2 < width) ? s_win[cb][64 * r + lane] : T(0);

      // Every wave reduces all candidates itself, with the same rule.
      constexpr int GQ = kMaxPanelGroups / 64;
      uint64_t gk[GQ];
      int gq[GQ];
#pragma unroll
      for (int q = 0; q < GQ; ++q) {  // unconditional LDS reads, then selects
        gk[q] = s_ck[cb][lane + 64 * q];
        gq[q] = s_crow[cb][lane + 64 * q];
      }
      uint64_t gm = 0;
#pragma unroll
      for (int q = 0; q < GQ; ++q) {
        gk[q] = lane + 64 * q < P.groups ? gk[q] : 0ull;
        gm = max(gm, gk[q]);
      }
      const uint64_t g1k = wave_max_key(gm);
      if (g1k == 0) {  // zero (or NaN-only) column: stop the panel at k
        stop = k;
        zero_col = 1;
        break;
      }
      int gr = INT_MAX;
#pragma unroll
      for (int q = 0; q < GQ; ++q) gr = (gk[q] == g1k) ? min(gr, gq[q]) : gr;
      const int p = wave_min(gr);
      const int gwin = (p - P.pb) / R;  // rows are dealt to workgroups in blocks of R
      DLU_PHASE(3)
      const bool p_here = p != k && p >= base && p < base + R;

      // One exchange round trip for everything this wave needs: its 16 columns of the pivot row
      // (lanes 0-15), the pivot row at columns c, c+1, c+2 (lanes 16-18), and, when row p lives
      // here, the old row k at columns c, c+1, c+2 (lanes 19-21) and its 16 columns (32-47);
      // lane 22 reads the winner's second-largest |value| (for the reported margin), lane 23 the
      // winner's origin and lane 24 row k's origin.
      const uint64_t *wslot = P.slots + (static_cast<size_t>(par) * P.groups + gwin) * SW;
      const uint64_t *wrow = wslot + kSlotHeaderWords;
      const uint64_t *src = nullptr;
      if (lane < 16 && col0 + lane < width) src = wrow + WPV * (col0 + lane);
      if (lane >= 16 && lane < 19 && c + lane - 16 < width) src = wrow + WPV * (c + lane - 16);
      if (p_here && lane >= 19 && lane < 22 && c + lane - 19 < width)
        src = rk + WPV * (c + lane - 19);
      if (lane == 22 && P.margins != nullptr) src = wslot + 2;
      if (lane == 23) src = wslot + 4;
      if (p_here && lane == 24) src = rk + 2 * NB;
      if (p_here && lane >= 32 && lane < 48 && col0 + lane - 32 < width)
        src = rk + WPV * (col0 + lane - 32);
      T got = T(0);
      if (src != nullptr) {
        unsigned w[WPV];
        if (!ll_get<WPV>(src, seq, w, P, t0)) s_fail = 1;
        got = Words<T>::join(w);
      }
      if (lane < 16) s_u[wave][lane] = got;
      if (lane >= 32 && lane < 48) s_k[wave][lane - 32] = got;
      const T piv = lane_value(got, 16);
      const T u1 = lane_value(got, 17), u2 = lane_value(got, 18);
      const T kc = lane_value(got, 19);
      const T k1 = lane_value(got, 20), k2 = lane_value(got, 21);
      const double w2 = static_cast<double>(lane_value(got, 22));
      const int o_p = from_word(lane_value(got, 23)), o_k = from_word(lane_value(got, 24));
      wave_sync_lds();
      DLU_PHASE(4)

      // Row k is final now: it is the pivot row (the L part, the pivot and the U part).  The
      // workgroup that published it still holds it staged in LDS and writes it out; the register
      // copy of row k (like those of the rows above it) is never read again.
      if (g == gwin && lane < 16 && col0 + lane < width)
        P.a[static_cast<size_t>(k) + static_cast<size_t>(P.pb + col0 + lane) * P.lda] =
            s_pub[wave][0][lane];

      // Move the old row k into row p (only the lane holding row p), form the multipliers
      // (column c times the reciprocal of the pivot) and update the window; the rank-1 update
      // of the panel columns right of c becomes pending.  Every row is updated: those at or above
      // k hold values that are never read again, and padding rows beyond m are never written back.
      const T recip = T(1) / piv;
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int row = base + 64 * r + lane;
        const bool is_p = p_here && row == p;
        if (is_p) {  // divergent: only the lane holding row p
#pragma unroll
          for (int j = 0; j < 16; ++j) v[r][j] = s_k[wave][j];
          wb[r] = k1;
          wc[r] = k2;
          if (wave == NW - 1) s_from[row - base] = o_k;
        }
        const T l = (is_p ? kc : wa[r]) * recip;
        s_lp[wave][64 * r + lane] = l;
        wb[r] = fma(-l, u1, wb[r]);
        wc[r] = fma(-l, u2, wc[r]);
      }
      jp = c - col0;  // may be out of 0..15
      pend = true;
      if (g == 0 && threadIdx.x == 0) {
        s_hist[c] = p;
        s_orig[c] = o_p;
        s_g1[c] = key_value(g1k);
      }
      prev_gwin = gwin;
      prev_w2 = w2;
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        wa[r] = wb[r];
        wb[r] = wc[r];
      }
      DLU_PHASE(5)
    }
  }
  DLU_APPLY_PENDING()
#undef DLU_APPLY_PENDING
  __syncthreads();  // exchange reads of the last step are checked here
  if (s_fail) goto leave;
  {
    const int steps = stop - P.pb;
    // Margin of the last eliminated step (its successor never ran).
    if (g == 0 && wave == 1 && prev_gwin >= 0 && P.margins != nullptr) {
      constexpr int GQ = kMaxPanelGroups / 64;
      const int cl = steps - 1;
      uint64_t o = 0;
#pragma unroll
      for (int q = 0; q < GQ; ++q) {
        const int gi = lane + 64 * q;
        const uint64_t kq = s_ck[cl & 1][gi];
        o = max(o, (gi < P.groups && gi != prev_gwin) ? kq : 0ull);
      }
      const double gs = key_value(wave_max_key(o));
      if (lane == 0) s_g2[cl] = fmax(gs, prev_w2);
    }

    // Write the rows not yet written back: L multipliers and the Schur columns of rows [stop, m)
    // (rows [pb, stop) went out as they were finished).
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int row = base + 64 * r + lane;
      if (row >= stop && row < m) {
#pragma unroll
        for (int j = 0; j < 16; ++j)
          if (col0 + j < width)
            P.a[static_cast<size_t>(row) + static_cast<size_t>(P.pb + col0 + j) * P.lda] = v[r][j];
      }
    }

    // Net row movement as (dst, src) pairs: rows [pb, stop) first, in order (workgroup 0), then
    // every row at or below stop whose content changed (each workgroup its own, appended).
    int *dst = P.swap_list;
    int *src = P.swap_list + 2 * NB;
    if (wave == 0) {
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int row = base + 64 * r + lane;
        const int from = s_from[64 * r + lane];
        if (row >= stop && row < m && from != row) {
          const int at = steps + atomicAdd(&P.swap_list[4 * NB + 1], 1);
          dst[at] = row;
          src[at] = from;
        }
      }
    }
    if (g == 0) {
      for (int i = threadIdx.x; i < steps; i += blockDim.x) {
        dst[i] = P.pb + i;
        src[i] = s_orig[i];
      }
    }
    __threadfence();
    __syncthreads();  // s_g2 complete; this workgroup's list entries are out
    if (threadIdx.x == 0) {
      // The last workgroup to finish closes the list and clears the counters for the next panel
      // that uses this list.
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
      // Pivot history and margins of the eliminated steps (a forced step keeps the row the
      // fallback recorded and has no margin).
      for (int s = threadIdx.x; s < steps; s += blockDim.x) {
        if (!(P.positioned && s == 0)) P.ipiv[P.pb + s] = s_hist[s];
        if (P.margins)
          P.margins[P.pb + s] =
              (P.positioned && s == 0) ? -1.0 : (s_g1[s] - fmax(s_g2[s], 0.0)) / s_g1[s];
      }
      if (wave == 0) {  // tie statistics of the column steps (when margins are tracked)
        double lo = 2.0;
        int near = 0, exact = 0;
        for (int s0 = 0; P.margins != nullptr && s0 < steps; s0 += 64) {
          const int s = s0 + lane;
          const bool valid = s < steps && !(P.positioned && s == 0);
          const double mg = valid ? (s_g1[s] - fmax(s_g2[s], 0.0)) / s_g1[s] : 2.0;
          lo = fmin(lo, mg);
          near += __popcll(__ballot(valid && mg < 1e-10));
          exact += __popcll(__ballot(valid && mg == 0.0));
        }
        min_margin = -wave_max(-lo);
        near_ties = near;
        exact_ties = exact;
        if (lane == 0) {
          if (timed) {
            DLU_PHASE(7)
            for (int i = 0; i < 8; ++i) P.phase_ticks[i] += s_ph[i];
          }
          P.ctl->stop = stop;
          P.ctl->zero_column = zero_col;
          if (stop == P.pe && zero_col == 0) P.ctl->go = P.done_seq;
          P.ctl->near_ties += near_ties;
          P.ctl->exact_ties += exact_ties;
          if (min_margin < P.ctl->min_margin) P.ctl->min_margin = min_margin;
          publish_done(P.ctl, P.done, P.done_seq);
        }
      }
    }
  }
  return;
#undef DLU_PHASE

leave:
  if (g == 0 && threadIdx.x == 0) {
    P.ctl->stop = P.pb;
    publish_done(P.ctl, P.done, P.done_seq);
  }
}

template <typename T, int NB>
hipError_t launch_nb(const PanelParams<T> &p, int rows_per_lane, hipStream_t s) {
  const dim3 grid(p.groups), block(NB * 4);
  switch (rows_per_lane) {
    case 1: hipLaunchKernelGGL((panel_kernel<T, NB, 1>), grid, block, 0, s, p); break;
    case 2: hipLaunchKernelGGL((panel_kernel<T, NB, 2>), grid, block, 0, s, p); break;
    case 4: hipLaunchKernelGGL((panel_kernel<T, NB, 4>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_panel(const PanelParams<T> &p, int nb, int rows_per_lane, hipStream_t s) {
  if (p.groups < 1 || p.groups > kMaxPanelGroups) return hipErrorInvalidValue;
  switch (nb) {
    case 64: return launch_nb<T, 64>(p, rows_per_lane, s);
    case 128: return launch_nb<T, 128>(p, rows_per_lane, s);
    case 256: return launch_nb<T, 256>(p, rows_per_lane, s);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_panel<float>(const PanelParams<float> &, int, int, hipStream_t);
template hipError_t launch_panel<double>(const PanelParams<double> &, int, int, hipStream_t);

}  // namespace dlu
