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
//
// One pivot step costs one all-to-all exchange between the workgroups.  Each workgroup publishes
// its local candidate (|value|, row, second-largest |value|) together with the candidate row's
// NB values; the owner of row k also publishes row k.  Every workgroup then reduces all candidates
// (largest magnitude, lowest row on ties -- BLAS idamax semantics), reads the winner's row and
// row k, performs the interchange on its own rows, scales its part of column k by the reciprocal
// of the pivot and applies the rank-1 update to its part of the panel.  Exchange words carry a
// 32-bit payload and the 32-bit sequence number of the step that wrote them (an LL-style
// protocol), so readers spin on relaxed device-scope loads and no cache writeback/invalidate
// fence sits on the per-step path.  Every wait is bounded (PanelParams::spin_limit); a
// workgroup whose wait expires raises ctl->abort_flag and all workgroups leave.
#include "dlu_internal.hpp"

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
template <int COUNT>
__device__ bool ll_get(const uint64_t *p, unsigned seq, unsigned *out, const PanelParams &P,
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

__device__ __forceinline__ double join_double(unsigned lo, unsigned hi) {
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Candidate order: larger magnitude first, then lower row.  NaN never compares larger.
__device__ __forceinline__ bool beats(double v, int row, double bv, int brow) {
  return v > bv || (v == bv && row < brow);
}

// Wave-wide (value, row, second value, tag) reduction keeping the two largest magnitudes.
__device__ __forceinline__ void wave_top2(double &v1, int &row, double &v2, int &tag) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov1 = __shfl_xor(v1, off);
    const int orow = __shfl_xor(row, off);
    const double ov2 = __shfl_xor(v2, off);
    const int otag = __shfl_xor(tag, off);
    if (beats(ov1, orow, v1, row)) {
      v2 = fmax(v1, ov2);
      v1 = ov1;
      row = orow;
      tag = otag;
    } else {
      v2 = fmax(v2, ov1);
    }
  }
}

// Sixteen columns of one row, kept as a register vector so that a wave-uniform runtime column
// index lowers to indirect register moves rather than scratch memory.
typedef double row16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NB, int RPL>
__global__ __launch_bounds__(NB * 4) void panel_kernel(PanelParams P) {
  constexpr int R = 64 * RPL;
  constexpr int SW = slot_words(NB);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = blockIdx.x;
  const int base = P.pb + g * R;
  const int width = P.pe - P.pb;
  const int col0 = wave * 16;
  const int m = P.m;

  __shared__ double s_lbuf[R];
  __shared__ double s_urow[NB];
  __shared__ double s_krow[NB];
  __shared__ double s_cv[kMaxPanelGroups];
  __shared__ double s_cv2[kMaxPanelGroups];
  __shared__ int s_crow[kMaxPanelGroups];
  __shared__ double s_loc_v, s_loc_v2;
  __shared__ int s_loc_row;
  __shared__ int s_fail;
  __shared__ int s_hist[NB];
  __shared__ int s_cur[NB];
  __shared__ int s_out_row[NB];
  __shared__ int s_out_cur[NB];

  row16_t v[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int row = base + 64 * r + lane;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int col = col0 + j;
      v[r][j] = (row < m && col < width)
                    ? P.a[static_cast<size_t>(row) + static_cast<size_t>(P.pb + col) * P.lda]
                    : 0.0;
    }
  }
  if (threadIdx.x == 0) s_fail = 0;
  const bool deferred = P.ctl->deferred != 0;  // a speculative restart after a deferral
  __syncthreads();
  if (deferred) {
    if (g == 0 && threadIdx.x == 0) {
      P.ctl->stop = P.pb;
      P.ctl->zero_column = 0;
      P.swap_list[4 * NB] = 0;
    }
    return;
  }

  const long long t0 = static_cast<long long>(wall_clock64());
  int stop = P.pe;
  int zero_col = 0;
  double min_margin = 2.0;
  int near_ties = 0, exact_ties = 0;
  // Optional phase timing (workgroup 0, thread 0; 100 MHz ticks).
  const bool timed = P.phase_ticks != nullptr && g == 0 && threadIdx.x == 0;
  unsigned long long ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long tp = timed ? wall_clock64() : 0;
#define DLU_PHASE(i)                         \
  if (timed) {                               \
    const unsigned long long tn = wall_clock64(); \
    ph[i] += tn - tp;                        \
    tp = tn;                                 \
  }

  for (int c = 0; c < width; ++c) {
    const int k = P.pb + c;
    const unsigned seq = P.seq_base + static_cast<unsigned>(c) + 1u;
    const int par = static_cast<int>(seq & 1u);
    const bool forced = P.positioned && c == 0;
    const int owner_wave = c >> 4;
    const int cc = c & 15;

    // Local candidate of this workgroup (only the wave holding column c has it).
    if (wave == owner_wave) {
      double b1 = -1.0, b2 = -1.0;
      int brow = INT_MAX, tag = 0;
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int row = base + 64 * r + lane;
        const bool ok = row >= k && row < m && (!forced || row == k);
        if (ok) {
          const double x = fabs(v[r][cc]);
          if (beats(x, row, b1, brow)) {
            b2 = fmax(b1, b2);
            b1 = x;
            brow = row;
          } else {
            b2 = fmax(b2, x);
          }
        }
      }
      wave_top2(b1, brow, b2, tag);
      if (lane == 0) {
        s_loc_v = b1;
        s_loc_v2 = b2;
        s_loc_row = brow;
      }
    }
    __syncthreads();
    DLU_PHASE(0)
    const int lrow = s_loc_row;

    // Publish: header, candidate row, and row k when this workgroup owns it.
    uint64_t *slot = P.slots + (static_cast<size_t>(par) * P.groups + g) * SW;
    uint64_t *rk = P.rowk + static_cast<size_t>(par) * 2 * NB;
    if (threadIdx.x == 0) {
      ll_put_double(slot + 0, seq, s_loc_v);
      ll_put_double(slot + 2, seq, s_loc_v2);
      ll_put(slot + 4, seq, static_cast<unsigned>(lrow));
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int row = base + 64 * r + lane;
      uint64_t *dst = row == lrow ? slot + kSlotHeaderWords : (row == k ? rk : nullptr);
      if (dst != nullptr) {
        // One lane per wave publishes; keep the loop rolled (runtime j indexes the vector).
#pragma unroll 1
        for (int j = 0; j < 16; ++j)
          if (col0 + j < width) ll_put_double(dst + 2 * (col0 + j), seq, v[r][j]);
        if (row == lrow && row == k) {
#pragma unroll 1
          for (int j = 0; j < 16; ++j)
            if (col0 + j < width) ll_put_double(rk + 2 * (col0 + j), seq, v[r][j]);
        }
      }
    }

    DLU_PHASE(1)
    // Gather every workgroup's header.
    if (threadIdx.x < P.groups) {
      const uint64_t *hs = P.slots + (static_cast<size_t>(par) * P.groups + threadIdx.x) * SW;
      unsigned w[5];
      if (!ll_get<5>(hs, seq, w, P, t0)) s_fail = 1;
      s_cv[threadIdx.x] = join_double(w[0], w[1]);
      s_cv2[threadIdx.x] = join_double(w[2], w[3]);
      s_crow[threadIdx.x] = static_cast<int>(w[4]);
    }
    __syncthreads();
    DLU_PHASE(2)
    if (s_fail) goto leave;

    {
      // Every wave reduces all candidates itself (no further barrier needed for the winner).
      double g1 = -1.0, g2 = -1.0;
      int grow = INT_MAX, gwin = 0;
      for (int i = lane; i < P.groups; i += 64) {
        const double cv = s_cv[i];
        const int crow = s_crow[i];
        if (beats(cv, crow, g1, grow)) {
          g2 = fmax(fmax(g1, g2), s_cv2[i]);
          g1 = cv;
          grow = crow;
          gwin = i;
        } else {
          g2 = fmax(fmax(g2, cv), s_cv2[i]);
        }
      }
      wave_top2(g1, grow, g2, gwin);
      DLU_PHASE(3)

      if (!(g1 > 0.0)) {  // zero (or NaN-only) column: stop the panel at k
        stop = k;
        zero_col = 1;
        break;
      }
      const int p = grow;

      // This wave's 16 columns of the pivot row and of row k.
      if (lane < 16 && col0 + lane < width) {
        const uint64_t *ws = P.slots + (static_cast<size_t>(par) * P.groups + gwin) * SW +
                             kSlotHeaderWords + 2 * (col0 + lane);
        unsigned w[2];
        if (!ll_get<2>(ws, seq, w, P, t0)) s_fail = 1;
        s_urow[col0 + lane] = join_double(w[0], w[1]);
        if (p != k) {
          if (!ll_get<2>(rk + 2 * (col0 + lane), seq, w, P, t0)) s_fail = 1;
          s_krow[col0 + lane] = join_double(w[0], w[1]);
        }
      }
      wave_sync_lds();
      DLU_PHASE(4)

      // Interchange rows k and p on this workgroup's rows.
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int row = base + 64 * r + lane;
        if (row == k) {
#pragma unroll 1
          for (int j = 0; j < 16; ++j) v[r][j] = s_urow[col0 + j];
        } else if (row == p) {
#pragma unroll 1
          for (int j = 0; j < 16; ++j) v[r][j] = s_krow[col0 + j];
        }
      }

      // Multipliers: column c times the reciprocal of the pivot.
      if (wave == owner_wave) {
        const double recip = 1.0 / s_urow[c];
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
          const int row = base + 64 * r + lane;
          double l = 0.0;
          if (row > k && row < m) {
            l = v[r][cc] * recip;
            v[r][cc] = l;
          }
          s_lbuf[64 * r + lane] = l;
        }
      }
      if (g == 0 && threadIdx.x == 0) {
        if (!forced) P.ipiv[k] = p;  // a forced step keeps the fallback's recorded row
        s_hist[c] = p;
        double margin = -1.0;
        if (!forced) {
          margin = (g1 - fmax(g2, 0.0)) / g1;
          min_margin = fmin(min_margin, margin);
          if (margin == 0.0) ++exact_ties;
          if (margin < 1e-10) ++near_ties;
        }
        if (P.margins) P.margins[k] = margin;
      }
      __syncthreads();
      DLU_PHASE(5)
      if (s_fail) goto leave;

      // Rank-1 update of this workgroup's rows, panel columns right of c.
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int row = base + 64 * r + lane;
        if (row > k && row < m) {
          const double l = s_lbuf[64 * r + lane];
#pragma unroll
          for (int j = 0; j < 16; ++j)
            if (col0 + j > c && col0 + j < width) v[r][j] = fma(-l, s_urow[col0 + j], v[r][j]);
        }
      }
      DLU_PHASE(6)
    }
  }

finish:
  // Write the panel back (L below the diagonal, U above, untouched Schur columns after a stop).
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int row = base + 64 * r + lane;
    if (row < m) {
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (col0 + j < width)
          P.a[static_cast<size_t>(row) + static_cast<size_t>(P.pb + col0 + j) * P.lda] = v[r][j];
    }
  }
  if (g == 0) {
    // Net row movement of steps [pb, stop) as (dst, src) pairs: rows [pb, stop) first, in order,
    // then every other row whose content changed.  Built by wave 0 from the pivot history.
    const int steps = stop - P.pb;
    if (wave == 0) {
      for (int i = lane; i < NB; i += 64) s_cur[i] = P.pb + i;
      int out_cnt = 0;
      wave_sync_lds();
      for (int s = 0; s < steps; ++s) {
        const int p = s_hist[s];
        if (p < P.pb + NB) {
          if (lane == 0) {
            const int t = s_cur[s];
            s_cur[s] = s_cur[p - P.pb];
            s_cur[p - P.pb] = t;
          }
        } else {
          int found = -1;
          for (int j0 = 0; j0 < out_cnt; j0 += 64) {
            const int j = j0 + lane;
            const unsigned long long hit = __ballot(j < out_cnt && s_out_row[j] == p);
            if (hit) {
              found = j0 + __ffsll(static_cast<long long>(hit)) - 1;
              break;
            }
          }
          if (found < 0) {
            found = out_cnt++;
            if (lane == 0) {
              s_out_row[found] = p;
              s_out_cur[found] = p;
            }
            wave_sync_lds();
          }
          if (lane == 0) {
            const int t = s_cur[s];
            s_cur[s] = s_out_cur[found];
            s_out_cur[found] = t;
          }
        }
        wave_sync_lds();
      }
      if (lane == 0) {
        int *dst = P.swap_list;
        int *src = P.swap_list + 2 * NB;
        int cnt = 0;
        for (int i = 0; i < steps; ++i) {
          dst[cnt] = P.pb + i;
          src[cnt] = s_cur[i];
          ++cnt;
        }
        for (int i = steps; i < NB; ++i)
          if (s_cur[i] != P.pb + i) {
            dst[cnt] = P.pb + i;
            src[cnt] = s_cur[i];
            ++cnt;
          }
        for (int j = 0; j < out_cnt; ++j)
          if (s_out_cur[j] != s_out_row[j]) {
            dst[cnt] = s_out_row[j];
            src[cnt] = s_out_cur[j];
            ++cnt;
          }
        P.swap_list[4 * NB] = cnt;
        if (timed) {
          DLU_PHASE(7)
          for (int i = 0; i < 8; ++i) P.phase_ticks[i] += ph[i];
        }
        P.ctl->stop = stop;
        P.ctl->zero_column = zero_col;
        P.ctl->near_ties += near_ties;
        P.ctl->exact_ties += exact_ties;
        if (min_margin < P.ctl->min_margin) P.ctl->min_margin = min_margin;
      }
    }
  }
  return;
#undef DLU_PHASE

leave:
  if (g == 0 && threadIdx.x == 0) P.ctl->stop = P.pb;
}

template <int NB>
hipError_t launch_nb(const PanelParams &p, int rows_per_lane, hipStream_t s) {
  const dim3 grid(p.groups), block(NB * 4);
  switch (rows_per_lane) {
    case 1: hipLaunchKernelGGL((panel_kernel<NB, 1>), grid, block, 0, s, p); break;
    case 2: hipLaunchKernelGGL((panel_kernel<NB, 2>), grid, block, 0, s, p); break;
    case 4: hipLaunchKernelGGL((panel_kernel<NB, 4>), grid, block, 0, s, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_panel(const PanelParams &p, int nb, int rows_per_lane, hipStream_t s) {
  if (p.groups < 1 || p.groups > kMaxPanelGroups) return hipErrorInvalidValue;
  switch (nb) {
    case 64: return launch_nb<64>(p, rows_per_lane, s);
    case 128: return launch_nb<128>(p, rows_per_lane, s);
    case 256: return launch_nb<256>(p, rows_per_lane, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace dlu
