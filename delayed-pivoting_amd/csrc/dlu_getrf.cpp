// Host driver and C ABI of libdlu: the delayed-pivoting LU on one MI355X.
//
// Reference interface replaced: delayed_lu_cuda/delayed_dgetrf_cuda.hpp:7-63 and
// delayed_lu_cuda/delayed_dgetrf_cuda.cu:370-517 (argument checks, status mapping, device/host
// entry points, aliases, status strings).  Algorithm control flow restated from
// delayed_lu/delayed_dgetrf.cpp:129-251 (panel loop, full-pivot fallback, panel restart,
// deferral) with decisions taken on the device:
//   * the panel kernel chooses every column pivot and stops by itself at a zero column;
//   * the full-pivot search, the interchanges it implies and the deferral flag are kernels too;
//   * the host reads a small control block once per panel (not once per column) to learn where
//     the panel stopped, and uses it only to size the next launches.
// Look-ahead: after panel j, the columns of panel j+1 are updated and factored on the caller's
// stream while the update of the far columns runs on a second stream of the same device.
#include "dlu.h"
#include "dlu_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <vector>

namespace dlu {
namespace {

// delayed_lu_cuda/delayed_dgetrf_cuda.cu:370-383: lda*n*8 must not overflow size_t.
bool matrix_bytes(int lda, int n, size_t &bytes, size_t elem = sizeof(double)) {
  const size_t cols = static_cast<size_t>(n), lead = static_cast<size_t>(lda);
  if (cols != 0 && lead > std::numeric_limits<size_t>::max() / cols) return false;
  const size_t elems = lead * cols;
  if (elems > std::numeric_limits<size_t>::max() / elem) return false;
  bytes = elems * elem;
  return true;
}

// delayed_lu_cuda/delayed_dgetrf_cuda.cu:385-397 and delayed_lu/delayed_dgetrf.cpp:132-142.
bool valid_arguments(int m, int n, const void *a, int lda, const int *ipiv, const int *jpiv) {
  if (m < 0 || n < 0 || lda < std::max(1, m)) return false;
  const int steps = std::min(m, n);
  if ((m > 0 && n > 0 && a == nullptr) || (steps > 0 && (ipiv == nullptr || jpiv == nullptr)))
    return false;
  size_t ignored = 0;
  return matrix_bytes(lda, n, ignored);
}

// Per-device state reused across calls (one factorization at a time per device).
struct Workspace {
  int device = -1;
  int steps_cap = 0;
  int cus = 0;
  Control *d_ctl = nullptr;
  Control *h_ctl = nullptr;
  // Pinned, coherent: panel launch `seq` publishes into slot seq % kDoneSlots (a panel queued on
  // speculation may finish before the host has read the one ahead of it).
  PanelDone *h_done = nullptr;
  PanelDone *d_done = nullptr;   // the same memory as the device sees it
  unsigned long long done_seq = 0;
  int *d_ipiv = nullptr;
  int *d_jpiv = nullptr;
  double *d_margins = nullptr;
  uint64_t *d_slots = nullptr;
  uint64_t *d_rowk = nullptr;
  int *d_lists = nullptr;
  double *d_pv = nullptr;
  unsigned long long *d_po = nullptr;
  double *d_gpv = nullptr;               // an update's argmax partials (kAmaxPartials each)
  unsigned long long *d_gpo = nullptr;
  unsigned long long *d_phase = nullptr;
  unsigned seq = 0;
  // Deferred left interchanges: every panel's swap list kept (archive), the composed
  // permutation's copies per closed block, and the per-column plan (pinned host staging and its
  // device copy).
  int *d_archive = nullptr;
  int archive_cap = 0;  // lists
  int *d_snaps = nullptr;
  size_t snaps_cap = 0;  // ints
  int *d_plan = nullptr;
  int *h_plan = nullptr;
  int plan_cap = 0;      // ints
  // Internal streams: panels run on `panel`, everything else on `update`.  When the device
  // allows it the two streams own disjoint sets of CUs, so the panel kernel (which needs all of
  // its workgroups resident at once) never waits for trailing-update workgroups to drain.
  struct Streams {
    int panel_cus = -1;  // CUs of the panel stream (0: both streams use the whole device)
    hipStream_t panel = nullptr, update = nullptr;
  };
  Streams streams[4];
  int nstreams = 0;
  int next_replace = 0;  // slot replaced when all four hold other partitions
  hipEvent_t ev_fork = nullptr, ev_next = nullptr, ev_join = nullptr;
  hipEvent_t ev_solved = nullptr, ev_tail = nullptr;  // far U block row done / tail GEMM done
  hipEvent_t ev_pan[4] = {};                          // panel launch seq ends: ev_pan[seq % 4]
  hipEvent_t ev_near = nullptr, ev_far = nullptr;     // panel-stream near work / update-stream work
  hipEvent_t ev_pair = nullptr;  // a pair's second interchange on the first panel's columns
  bool ready = false;
};

constexpr int kMaxDevices = 64;
constexpr int kCusPerXcd = 32;  // MI355X (gfx950): 8 XCDs of 32 CUs
constexpr int kSearchPartials = 2048;
constexpr int kListStride = 4 * 256 + 4;  // swap_list_ints(256), enough for every nb
constexpr int kLists = 4;  // panels j and j+4 share a list (a pair's far update reads two)
constexpr int kDoneSlots = 4;  // host-visible control block copies, by panel launch

// Exchange buffers, sized for the largest panel width (256) and workgroup count.
constexpr size_t slot_bytes() {
  return 2ull * kMaxPanelGroups * static_cast<size_t>(slot_words(256)) * sizeof(uint64_t);
}
constexpr size_t rowk_bytes() { return 2ull * rowk_words(256) * sizeof(uint64_t); }

// Factorization workspaces: a small pool per device.  A call leases one for its whole run, so
// calls from several host threads (the serving workers, or a caller's own threads) overlap on
// the device.  Admission keeps panels from blocking one another: a panel kernel needs all of its
// workgroups resident at once (one per CU), so the panel workgroups of the calls running on a
// device never exceed its CU count, and at most one call at a time uses a CU-partitioned panel
// stream (every partition covers the same CUs).  Every other kernel of a factorization runs to
// completion on its own or waits only on workgroups that already run.
constexpr int kPoolSize = 4;
struct DevicePool {
  std::mutex mu;
  std::condition_variable cv;
  Workspace ws[kPoolSize];
  bool busy[kPoolSize] = {};
  int groups = 0;       // panel workgroups of the running calls
  bool masked = false;  // a running call uses a CU-partitioned panel stream
};
DevicePool g_pool[kMaxDevices];

class Lease {
 public:
  Lease(int device, int groups, bool masked, int cus)
      : p_(g_pool[device]), groups_(groups), masked_(masked) {
    std::unique_lock<std::mutex> lk(p_.mu);
    p_.cv.wait(lk, [&] {
      bool free = false;
      for (bool b : p_.busy) free = free || !b;
      return free && (p_.groups == 0 || p_.groups + groups <= cus) && !(masked && p_.masked);
    });
    for (idx_ = 0; p_.busy[idx_]; ++idx_) {
    }
    p_.busy[idx_] = true;
    p_.groups += groups;
    p_.masked = p_.masked || masked;
  }
  ~Lease() {
    {
      std::lock_guard<std::mutex> lk(p_.mu);
      p_.busy[idx_] = false;
      p_.groups -= groups_;
      if (masked_) p_.masked = false;
    }
    p_.cv.notify_all();
  }
  Lease(const Lease &) = delete;
  Lease &operator=(const Lease &) = delete;
  Workspace &workspace() { return p_.ws[idx_]; }

 private:
  DevicePool &p_;
  int groups_;
  bool masked_;
  int idx_ = 0;
};

// Rows per lane and workgroups of the panel kernel for `rows` rows: the panel is held in
// registers, so as few workgroups as possible (the most rows per lane the panel width allows).
inline void panel_shape_for(int rows, int nb, int forced_rpl, int max_groups, int &groups,
                            int &rpl) {
  if (forced_rpl > 0) {
    rpl = forced_rpl == 3 ? 4 : forced_rpl;  // 3 rows per lane: the one-exchange kernel only
  } else {
    rpl = nb == 256 ? 2 : 4;
    if (nb == 256 && (rows + 63) / 64 <= max_groups) rpl = 1;  // avoid the spilling variant
  }
  groups = (rows + 64 * rpl - 1) / (64 * rpl);
}

// Shape of the one-exchange-per-step panel kernel: FP64 panels hold 2 rows per lane (a
// workgroup of nine waves keeps 16 x 2 doubles per lane without spilling), FP32 panels 4.  False
// where that kernel does not take the panel (nb 256, or more than kLagMaxGroups workgroups).
inline bool lag_shape_for(int rows, int nb, size_t elem, int forced_rpl, int max_groups,
                          int &groups, int &rpl) {
  if (nb != 64 && nb != 128) return false;
  // Rows per lane: fewer make a shorter step, more leave more CUs to the trailing GEMM that runs
  // beside the panel while the update is the longer chain.  FP64: 2 in registers, 3 (one in
  // LDS) above 10000 rows (n = 16384: 100.7 -> 99.9 ms), 4 (two in LDS) above 12288 rows, where
  // the panel has most slack beside the GEMM (n = 16384: 95.7-96.0 -> 95.4 ms; 4 from 10000
  // rows slows n = 12288), 1 up to 4096 rows (at most 64 workgroups: n = 2048 5.94 -> 5.45 ms,
  // n = 4096 12.37 -> 11.66 ms).  FP32: 4, and 2 up to 4096 rows (n = 2048: 5.89 -> 5.18 ms; at
  // n = 16384 4 is best: 68.1 against 71.3 ms).
  if (forced_rpl > 0)
    rpl = forced_rpl;
  else if (elem == 8)
    rpl = rows > 12288 ? 4 : (rows > 10000 ? 3 : (rows > 4096 ? 2 : 1));
  else
    rpl = rows > 4096 ? 4 : 2;
  groups = (rows + 64 * rpl - 1) / (64 * rpl);
  return groups <= std::min(kLagMaxGroups, max_groups);
}

// Buffers of the deferred left interchanges for m rows, n columns and at most `lists` panels.
hipError_t prepare_defer(Workspace &w, int m, int n, int lists) {
  hipError_t e;
  if (lists > w.archive_cap) {
    hipFree(w.d_archive);
    w.d_archive = nullptr;
    w.archive_cap = 0;
    if ((e = hipMalloc(&w.d_archive, static_cast<size_t>(lists) * kListStride * sizeof(int))) !=
        hipSuccess)
      return e;
    w.archive_cap = lists;
  }
  const size_t snaps = static_cast<size_t>(lists) * m;
  if (snaps > w.snaps_cap) {
    hipFree(w.d_snaps);
    w.d_snaps = nullptr;
    w.snaps_cap = 0;
    if ((e = hipMalloc(&w.d_snaps, snaps * sizeof(int))) != hipSuccess) return e;
    w.snaps_cap = snaps;
  }
  const int plan = lists + 1 + 2 * n;
  if (plan > w.plan_cap) {
    hipFree(w.d_plan);
    hipHostFree(w.h_plan);
    w.d_plan = w.h_plan = nullptr;
    w.plan_cap = 0;
    if ((e = hipMalloc(&w.d_plan, plan * sizeof(int))) != hipSuccess) return e;
    if ((e = hipHostMalloc(&w.h_plan, plan * sizeof(int))) != hipSuccess) return e;
    w.plan_cap = plan;
  }
  return hipSuccess;
}

hipError_t prepare(Workspace &w, int device, int steps) {
  if (!w.ready) {
    w.device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return e;
    w.cus = prop.multiProcessorCount;
    if ((e = hipMalloc(&w.d_ctl, sizeof(Control))) != hipSuccess) return e;
    if ((e = hipHostMalloc(&w.h_ctl, sizeof(Control))) != hipSuccess) return e;
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&w.h_done), kDoneSlots * sizeof(PanelDone),
                           hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess)
      return e;
    std::memset(static_cast<void *>(w.h_done), 0, kDoneSlots * sizeof(PanelDone));
    if ((e = hipHostGetDevicePointer(reinterpret_cast<void **>(&w.d_done), w.h_done, 0)) !=
        hipSuccess)
      return e;
    if ((e = hipMalloc(&w.d_slots, slot_bytes())) != hipSuccess) return e;
    if ((e = hipMemset(w.d_slots, 0, slot_bytes())) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_rowk, rowk_bytes())) != hipSuccess) return e;
    if ((e = hipMemset(w.d_rowk, 0, rowk_bytes())) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_lists, kLists * kListStride * sizeof(int))) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_pv, kSearchPartials * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_po, kSearchPartials * sizeof(unsigned long long))) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_gpv, kAmaxPartials * sizeof(double))) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_gpo, kAmaxPartials * sizeof(unsigned long long))) != hipSuccess)
      return e;
    if ((e = hipMalloc(&w.d_phase, 8 * sizeof(unsigned long long))) != hipSuccess) return e;
    for (hipEvent_t *ev : {&w.ev_fork, &w.ev_next, &w.ev_join, &w.ev_solved, &w.ev_tail,
                           &w.ev_pan[0], &w.ev_pan[1], &w.ev_pan[2], &w.ev_pan[3], &w.ev_near,
                           &w.ev_far, &w.ev_pair})
      if ((e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipDeviceSynchronize()) != hipSuccess) return e;
    w.ready = true;
  }
  if (steps > w.steps_cap) {
    hipFree(w.d_ipiv);
    hipFree(w.d_jpiv);
    hipFree(w.d_margins);
    w.d_ipiv = w.d_jpiv = nullptr;
    w.d_margins = nullptr;
    w.steps_cap = 0;
    hipError_t e;
    if ((e = hipMalloc(&w.d_ipiv, steps * sizeof(int))) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_jpiv, steps * sizeof(int))) != hipSuccess) return e;
    if ((e = hipMalloc(&w.d_margins, steps * sizeof(double))) != hipSuccess) return e;
    w.steps_cap = steps;
  }
  return hipSuccess;
}

// The stream pair whose panel stream owns `panel_cus` CUs (0: no partition).  CUs
// [0, panel_cus) of the device mask go to the panel stream, the rest to the update stream.
hipError_t stream_pair(Workspace &w, int panel_cus, Workspace::Streams *&out) {
  for (int i = 0; i < w.nstreams; ++i)
    if (w.streams[i].panel_cus == panel_cus) {
      out = &w.streams[i];
      return hipSuccess;
    }
  // All slots taken by other partitions: replace one (round robin).  The workspace is leased by
  // this call alone and every earlier call joined its streams before returning, so the pair
  // replaced holds no work.
  int slot = w.nstreams;
  if (slot == 4) {
    slot = w.next_replace;
    w.next_replace = (w.next_replace + 1) % 4;
    Workspace::Streams &old = w.streams[slot];
    if (old.panel) hipStreamDestroy(old.panel);
    if (old.update) hipStreamDestroy(old.update);
    old = Workspace::Streams{};
  }
  Workspace::Streams &st = w.streams[slot];
  hipError_t e = hipSuccess;
  bool masked = false;
  if (panel_cus > 0 && panel_cus < w.cus) {
    const int words = (w.cus + 31) / 32;
    std::vector<uint32_t> pm(words, 0u), um(words, 0u);
    for (int c = 0; c < w.cus; ++c) (c < panel_cus ? pm : um)[c / 32] |= 1u << (c % 32);
    masked = hipExtStreamCreateWithCUMask(&st.panel, words, pm.data()) == hipSuccess &&
             hipExtStreamCreateWithCUMask(&st.update, words, um.data()) == hipSuccess;
    if (!masked) {
      if (st.panel) hipStreamDestroy(st.panel);
      if (st.update) hipStreamDestroy(st.update);
      st.panel = st.update = nullptr;
      (void)hipGetLastError();
    }
  }
  if (!masked) {
    if ((e = hipStreamCreateWithFlags(&st.panel, hipStreamNonBlocking)) != hipSuccess) return e;
    if ((e = hipStreamCreateWithFlags(&st.update, hipStreamNonBlocking)) != hipSuccess) return e;
  }
  st.panel_cus = masked ? panel_cus : 0;
  if (slot == w.nstreams) ++w.nstreams;
  out = &st;
  return hipSuccess;
}

template <typename T>
inline T *at(T *a, int lda, int i, int j) {
  return a + static_cast<size_t>(i) + static_cast<size_t>(j) * static_cast<size_t>(lda);
}

// One factorization of a column-major T matrix (T = double, or float for the FP32 factor of the
// mixed-precision solver).
template <typename T>
struct Run {
  Workspace &w;
  int m, n, lda, steps, nb;
  T *a;
  hipStream_t s;  // the caller's stream
  long long spin;
  bool profile = false;
  // Per-step top-2 margins and the tie statistics: tracked only when the caller asks for the
  // margins (the reference reports none, and the bookkeeping sits on the panel step).
  bool track_margins = false;
  // Everything in order on the caller's stream, with no events and no pinned host memory (the
  // allocation-free entry point); look-ahead, pairs and the panel-stream tail are off then.
  bool single = false;
  int panels = 0;
  int max_groups = kMaxPanelGroups;  // workgroups the panel stream can hold at once
  int forced_rpl = 0;                // rows per lane requested by the caller (0: automatic)
  int merge_rows = 3000;             // pair panels while at least this many rows remain
  bool split = true;                 // tail of the far GEMM on the panel stream's CUs
  int forced_split = 0;              // > 0: that many far columns always go to the panel stream
  int panel_cus = 0;                 // CUs of the panel stream (0: no partition)
  int variant = DLU_FALLBACK_FULL;   // fallback search: full pivot or rightmost nonzero column
  int panel_kernel = DLU_PANEL_AUTO; // which register-resident panel kernel factors a panel
  // Deferred left interchanges: while on, each panel writes its swap list into its own archive
  // slot (panel e: event e), and the rows of the columns left of a panel are not interchanged
  // when it ends; each closed column block records the first event it has not yet taken, and
  // one pass at the end (or before a full-pivot fallback, which interchanges whole rows) applies
  // the rest.  Off for single-stream runs and tall or very tall matrices.
  bool defer = false;
  int events = 0;
  struct Block {
    int c0, c1, first;  // columns [c0, c1) have taken the row movements of events < first
  };
  std::vector<Block> blocks;
  // A panel launch: its block, swap list (and archive event, deferred mode), done token (its
  // done_seq, which also picks its host-visible slot and end event) and whether it was queued on
  // speculation (before the panel ahead of it finished).
  struct Launch {
    bool valid = false;
    int pb = 0, pe = 0;
    const int *list = nullptr;
    int event = -1;
    unsigned long long seq = 0;
    bool spec = false;
  };
  Launch launched;      // the last panel launch
  int launches = 0;     // panel launches so far (they take the swap lists in turn)
  // Predicate of the launches being queued: set while a speculative step is queued.
  Pred pred{nullptr, 0};
  hipStream_t ps = nullptr, us = nullptr;  // internal panel / update streams

  // Opt-in launch timing: an event pair around each launch, on the launch's own stream.
  struct Timed {
    hipEvent_t begin, end;
    double flops, bytes;  // algorithmic flops and memory bytes of the launch
  };
  std::vector<Timed> t_update, t_update_big, t_panel, t_solve, t_search;
  hipError_t mark(std::vector<Timed> &v, hipStream_t st, double flops, double bytes = 0.0) {
    Timed t{nullptr, nullptr, flops, bytes};
    hipError_t e = hipEventCreate(&t.begin);
    if (e == hipSuccess) e = hipEventCreate(&t.end);
    if (e == hipSuccess) e = hipEventRecord(t.begin, st);
    v.push_back(t);
    return e;
  }
  hipError_t close(std::vector<Timed> &v, hipStream_t st) { return hipEventRecord(v.back().end, st); }
  static double total(const std::vector<Timed> &v, double *flops, double *bytes = nullptr) {
    double sec = 0.0, fl = 0.0, by = 0.0;
    for (const Timed &t : v) {
      float ms = 0.0f;
      if (t.begin && t.end && hipEventElapsedTime(&ms, t.begin, t.end) == hipSuccess) sec += ms * 1e-3;
      fl += t.flops;
      by += t.bytes;
    }
    if (flops) *flops = fl;
    if (bytes) *bytes = by;
    return sec;
  }
  void release() {
    for (auto *v : {&t_update, &t_update_big, &t_panel, &t_solve, &t_search})
      for (Timed &t : *v) {
        if (t.begin) hipEventDestroy(t.begin);
        if (t.end) hipEventDestroy(t.end);
      }
  }

  // Rows per lane of the panel kernel: the panel's rows must fit in the co-resident workgroups.
  // The panel is held in registers: as few workgroups as possible (the most rows per lane the
  // panel width allows), and no more than the panel stream can keep resident together.
  bool panel_shape(int pb, int &groups, int &rpl) const {
    panel_shape_for(m - pb, nb, forced_rpl, max_groups, groups, rpl);
    return groups <= max_groups;
  }
  // A panel starting at pb with more rows than the register-resident kernel can hold on the
  // panel stream's CUs is factored by the tall path instead.
  bool tall(int pb) const {
    int groups = 0, rpl = 0;
    return pb < steps && !panel_shape(pb, groups, rpl);
  }

  // Tall panel: one column step at a time (column argmax, whole-row interchange, multipliers and
  // rank-1 update), every launch predicated on the control block; rows are interchanged across
  // all columns at once, so its swap list is the identity.  Runs in order on `st`.
  hipError_t tall_panel(int pb, int pe, int positioned, hipStream_t st) {
    int *list = w.d_lists + (launches++ % kLists) * kListStride;
    hipError_t e = profile ? mark(t_panel, st, 0.0) : hipSuccess;
    if (e == hipSuccess) e = launch_tall_begin(w.d_ctl, pb, pe, st);
    for (int k = pb; e == hipSuccess && k < pe; ++k) {
      if (!(positioned && k == pb)) {
        e = launch_active_search(a, lda, m, k + 1, k, w.d_pv, w.d_po, w.d_ctl, st);
        if (e == hipSuccess) e = launch_tall_pivot(a, lda, n, k, w.d_ipiv, w.d_ctl, st);
      }
      if (e == hipSuccess) e = launch_tall_update(a, lda, m, k, pe, w.d_ctl, st);
    }
    const unsigned long long seq = ++w.done_seq;
    if (e == hipSuccess) e = launch_tall_finish(w.d_ctl, list, nb, pb, done_slot(seq), seq, st);
    if (e == hipSuccess && profile) e = close(t_panel, st);
    if (e != hipSuccess) return e;
    ++panels;
    launched = Launch{true, pb, pe, list, -1, seq, false};
    return single ? hipSuccess : hipEventRecord(w.ev_pan[seq % kDoneSlots], st);
  }
  PanelDone *done_slot(unsigned long long seq) const {
    return w.d_done ? w.d_done + seq % kDoneSlots : nullptr;
  }

  hipError_t panel(int pb, int pe, int positioned, hipStream_t st) {
    int groups = 0, rpl = 0;
    const bool lag = panel_kernel != DLU_PANEL_MOVE &&
                     lag_shape_for(m - pb, nb, sizeof(T), forced_rpl, max_groups, groups, rpl);
    if (!lag && panel_kernel == DLU_PANEL_LAG) return hipErrorInvalidValue;
    if (!lag && !panel_shape(pb, groups, rpl)) return tall_panel(pb, pe, positioned, st);
    PanelParams<T> p{};
    p.a = a;
    p.lda = lda;
    p.m = m;
    p.pb = pb;
    p.pe = pe;
    p.positioned = positioned;
    p.groups = groups;
    // Step tags are 16 bits wide in the exchange headers: before they would wrap, every
    // exchange word is cleared (tag 0, never a valid step) and numbering restarts at 1.
    if (w.seq + static_cast<unsigned>(nb) + 1u > kMaxStepTag) {
      hipError_t e = hipMemsetAsync(w.d_slots, 0, slot_bytes(), st);
      if (e == hipSuccess) e = hipMemsetAsync(w.d_rowk, 0, rowk_bytes(), st);
      if (e != hipSuccess) return e;
      w.seq = 0;
    }
    p.seq_base = w.seq;
    w.seq += static_cast<unsigned>(nb) + 1u;
    p.ipiv = w.d_ipiv;
    p.margins = track_margins ? w.d_margins : nullptr;
    p.ctl = w.d_ctl;
    p.slots = w.d_slots;
    p.rowk = w.d_rowk;
    if (defer && events >= w.archive_cap) return hipErrorInvalidValue;  // (one list per panel)
    p.swap_list = defer ? w.d_archive + static_cast<size_t>(events) * kListStride
                        : w.d_lists + (launches % kLists) * kListStride;
    ++launches;
    const int event = defer ? events++ : -1;
    p.spin_limit = spin;
    p.phase_ticks = profile ? w.d_phase : nullptr;
    p.done_seq = ++w.done_seq;
    p.done = done_slot(p.done_seq);
    p.pred = pred;
    hipError_t e = profile ? mark(t_panel, st, 0.0) : hipSuccess;
    if (e == hipSuccess) e = lag ? launch_panel_lag(p, nb, rpl, st) : launch_panel(p, nb, rpl, st);
    if (e == hipSuccess && profile) e = close(t_panel, st);
    if (e != hipSuccess) return e;
    ++panels;
    launched = Launch{true, pb, pe, p.swap_list, event, p.done_seq, pred.go != nullptr};
    return single ? hipSuccess : hipEventRecord(w.ev_pan[p.done_seq % kDoneSlots], st);
  }

  // A panel's outcome: spin on the word the panel kernel stores into its host slot as its last
  // act (its sequence number in the top bits).  A launch whose end event completes without that
  // store has failed.
  hipError_t wait_panel(const Launch &L, Control &c) {
    if (single || w.h_done == nullptr) {  // no pinned copy: read the device's
      hipError_t e = hipMemcpyAsync(&c, w.d_ctl, sizeof(Control), hipMemcpyDeviceToHost, s);
      return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    PanelDone &d = w.h_done[L.seq % kDoneSlots];
    const unsigned long long want = L.seq & kDoneSeqMask;
    auto take = [&](unsigned long long word) {
      c = Control{};
      c.stop = static_cast<int>(static_cast<unsigned>(word));
      c.zero_column = static_cast<int>((word >> 32) & 1u);
      c.deferred = static_cast<int>((word >> 33) & 1u);
      c.error = static_cast<int>((word >> 34) & 3u);
    };
    for (unsigned it = 1;; ++it) {
      unsigned long long word = __atomic_load_n(&d.word, __ATOMIC_ACQUIRE);
      if ((word >> 36) == want) {
        take(word);
        return hipSuccess;
      }
      if ((it & 255) == 0) {
        const hipError_t q = hipEventQuery(w.ev_pan[L.seq % kDoneSlots]);
        if (q == hipSuccess) {
          word = __atomic_load_n(&d.word, __ATOMIC_ACQUIRE);
          if ((word >> 36) == want) {
            take(word);
            return hipSuccess;
          }
          return hipErrorLaunchFailure;
        }
        if (q != hipErrorNotReady) return q;
      }
      __builtin_ia32_pause();
    }
  }

  // Interchange + U block row + trailing update of columns [c0, c1) for the panel rows
  // [pb, pb+w), updating rows [pb+w, m).
  // A(r0:r0+M, c0:c0+N) -= A(r0:r0+M, kc:kc+K) * A(kr:kr+K, c0:c0+N).
  hipError_t gemm(int r0, int c0, int kc, int kr, int M, int N, int K, hipStream_t st,
                  AmaxOut amax = AmaxOut{nullptr, nullptr, 0, 0}) {
    if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
    const double flops = 2.0 * M * static_cast<double>(N) * K;
    // C read and written once, A and B read once: the least memory traffic of the launch.
    const double bytes =
        sizeof(T) * (2.0 * M * static_cast<double>(N) + static_cast<double>(K) * (M + N));
    std::vector<Timed> &tv = gemm_uses_large_tiles(M, N) ? t_update_big : t_update;
    hipError_t e = profile ? mark(tv, st, flops, bytes) : hipSuccess;
    if (e == hipSuccess)
      e = launch_gemm_sub<T>(at(a, lda, r0, c0), lda, at(a, lda, r0, kc), lda, at(a, lda, kr, c0),
                             lda, M, N, K, st, -1, pred, amax);
    if (e == hipSuccess && profile) e = close(tv, st);
    return e;
  }

  // Rows [pb, pb+w) of columns [c0, c1) become U12 = L11^{-1} (rows after the interchange of
  // `list`; no list: in place).
  hipError_t solve(int pb, int wdt, int c0, int c1, const int *list, hipStream_t st) {
    if (c1 <= c0 || wdt <= 0) return hipSuccess;
    hipError_t e = profile ? mark(t_solve, st, 0.0) : hipSuccess;
    if (e == hipSuccess) e = launch_swap_trsm(a, lda, pb, wdt, c0, c1, list, nb, st, pred);
    if (e == hipSuccess && profile) e = close(t_solve, st);
    return e;
  }

  // Interchange + U block row + trailing update of columns [c0, c1) for the panel rows
  // [pb, pb+w), updating rows [pb+w, m) (delayed_dgetrf.cpp:61-85, 101-125).
  hipError_t update(int pb, int wdt, int c0, int c1, const int *list, hipStream_t st) {
    if (c1 <= c0 || wdt <= 0) return hipSuccess;
    hipError_t e = solve(pb, wdt, c0, c1, list, st);
    if (e == hipSuccess) e = gemm(pb + wdt, c0, pb, pb, m - pb - wdt, c1 - c0, wdt, st);
    return e;
  }

  // Panel pairs.  While the trailing GEMM dominates, two consecutive full panels P0 (rows and
  // columns from b0) and P1 (from b1 = b0 + nb) share one K = 2 nb update.  After P0 only the
  // next panel's columns are brought up to date; the columns beyond wait.  After P1, for a
  // column range: P0's interchange and U block row (TRSM), P1's interchange, P0's update of
  // P1's rows, P1's U block row, and one GEMM with K = 2 nb for every row below (interchanges
  // commute with the deferred update once P0's L columns carry P1's interchange too).  The
  // next panel's block goes first, then the far columns' left half; their right half follows the
  // next step's near update, so on the in-order update stream no near update ever waits for more
  // than half of a pair's far work.
  struct Pair {
    bool open = false;  // P0 done, its far update owed
    int b0 = 0, fe = 0;  // P0's first row/column; first column still owed P0's update
    const int *l0 = nullptr;
    int e0 = -1;        // P0's event (deferred left interchanges)
  };
  Pair pair;
  struct Rest {  // the right half of a closed pair's far update, not yet queued
    bool pending = false;
    int b0 = 0, c0 = 0;
    const int *l0 = nullptr, *l1 = nullptr;
  };
  Rest rest;
  // Columns [c0, c1) up to (but without) the K = 2 nb GEMM.
  hipError_t pair_prepare(int b0, const int *l0, const int *l1, int c0, int c1, hipStream_t st) {
    const int b1 = b0 + nb;
    // P0's interchange and U block row, then P1's interchange, in one pass over the strips
    hipError_t e = profile ? mark(t_solve, st, 0.0) : hipSuccess;
    if (e == hipSuccess) e = launch_swap_trsm_pair(a, lda, b0, c0, c1, l0, l1, nb, st, pred);
    if (e == hipSuccess && profile) e = close(t_solve, st);
    if (e == hipSuccess) e = gemm(b1, c0, b0, b0, nb, c1 - c0, nb, st);
    if (e == hipSuccess) e = solve(b1, nb, c0, c1, nullptr, st);
    return e;
  }
  hipError_t pair_update(int b0, const int *l0, const int *l1, int c0, int c1, hipStream_t st) {
    if (c1 <= c0) return hipSuccess;
    hipError_t e = pair_prepare(b0, l0, l1, c0, c1, st);
    if (e == hipSuccess) e = gemm(b0 + 2 * nb, c0, b0, b0, m - b0 - 2 * nb, c1 - c0, 2 * nb, st);
    return e;
  }
  // Queue the pending right half, then the pair's left interchanges.  With next_wdt > 0 (the
  // width of the panel that runs meanwhile, when nothing else of this step touches far columns)
  // its GEMM's last columns may go to the panel stream.
  hipError_t queue_rest(int next_wdt) {
    if (!rest.pending) return hipSuccess;
    rest.pending = false;
    const int b0 = rest.b0, c0 = rest.c0, r0 = b0 + 2 * nb;
    hipError_t e = hipSuccess;
    if (c0 < n) {
      if ((e = wait_tail(us)) != hipSuccess) return e;
      if ((e = pair_prepare(b0, rest.l0, rest.l1, c0, n, us)) != hipSuccess) return e;
      const int cb = next_wdt > 0 ? tail_split(m - r0, 2 * nb, c0, next_wdt) : n;
      if (cb < n) {
        if ((e = hipEventRecord(w.ev_solved, us)) != hipSuccess) return e;
        tail = Tail{true, false, r0, b0, 2 * nb, cb};
      }
      if ((e = gemm(r0, c0, b0, b0, m - r0, cb - c0, 2 * nb, us)) != hipSuccess) return e;
    }
    if (defer) return hipSuccess;  // both movements reach the columns left of b0 at the end
    if ((e = launch_row_interchange(a, lda, 0, b0, rest.l0, nb, us)) != hipSuccess) return e;
    return launch_row_interchange(a, lda, 0, b0, rest.l1, nb, us);
  }
  // P0's owed update on its own (the pair could not be completed).
  hipError_t close_pair(hipStream_t st) {
    if (!pair.open) return hipSuccess;
    pair.open = false;
    hipError_t e = wait_tail(st);
    if (e == hipSuccess) e = update(pair.b0, nb, pair.fe, n, pair.l0, st);
    if (e == hipSuccess) e = left(pair.b0, pair.b0 + nb, pair.e0, pair.l0, st);
    return e;
  }

  // The row movement of panel `ev` (its swap list) for the columns left of its block
  // [pb, pe): interchanged now, or, deferred, recorded as the block's first owed movement.
  hipError_t left(int pb, int pe, int ev, const int *list, hipStream_t st) {
    if (defer) {
      blocks.push_back(Block{pb, pe, ev + 1});
      return hipSuccess;
    }
    return launch_row_interchange(a, lda, 0, pb, list, nb, st);
  }

  // The deferred left interchanges: every movement kept so far, composed backward, applied once
  // to the rows below each recorded block.  Afterwards interchanges are immediate again.
  hipError_t apply_deferred(hipStream_t st) {
    if (!defer) return hipSuccess;
    defer = false;
    const int E = events;
    int *snap_slot = w.h_plan;          // [0, E]
    int *col_slot = w.h_plan + E + 1;   // [0, n)
    int *col_r0 = col_slot + n;         // [0, n)
    int slots = 0, ncols = 0;
    for (int e = 0; e <= E; ++e) snap_slot[e] = -1;
    for (const Block &b : blocks) {
      if (b.first < E && snap_slot[b.first] < 0) snap_slot[b.first] = slots++;
      ncols = std::max(ncols, b.c1);
    }
    for (int c = 0; c < ncols; ++c) col_slot[c] = -1;
    for (const Block &b : blocks)
      for (int c = b.c0; c < b.c1; ++c) {
        col_slot[c] = b.first < E ? snap_slot[b.first] : -1;
        col_r0[c] = b.c1;
      }
    blocks.clear();
    if (slots == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(w.d_plan, w.h_plan, (E + 1 + 2 * static_cast<size_t>(n)) * sizeof(int),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = launch_left_compose(m, w.d_archive, kListStride, nb, E, w.d_plan, w.d_snaps, st);
    if (e == hipSuccess)
      e = launch_left_gather<T>(a, lda, m, ncols, w.d_plan + E + 1, w.d_plan + E + 1 + n,
                                w.d_snaps, st);
    return e;
  }

  // Far-update tail on the panel stream.  While the trailing GEMM outlasts the next panel, the
  // panel stream's CUs would idle once that panel is done; the GEMM's last columns [cb, n) are
  // queued there behind it instead.  The split balances the two streams with a throughput model
  // (per-CU MFMA rate of the large-tile GEMM, per-column panel step latency, U block row solve
  // per column):
  //   update stream: solve + GEMM(pe2 .. cb) on U CUs;  panel stream: panel + GEMM(cb .. n) on P.
  struct Tail {
    bool queued = false;   // a tail GEMM is waiting to be launched behind the next panel
    bool pending = false;  // a launched tail GEMM may still run (ev_tail marks its end)
    int r0 = 0, kc = 0, K = 0, cb = 0;  // A(r0:m, cb:n) -= A(r0:m, kc:kc+K) A(kc:kc+K, cb:n)
  };
  Tail tail;
  // First column of the tail for a GEMM over rows [m - rows, m), columns [c0, n), depth wdt.
  int tail_split(int rows, int wdt, int c0, int next_wdt) const {
    if (!split || c0 >= n) return n;
    if (forced_split > 0) {
      const int cols = std::min(forced_split, n - c0 - nb);
      return cols > 0 ? n - cols : n;
    }
    if (panel_cus <= 0) return n;
    const int U = std::max(1, w.cus - panel_cus), P = panel_cus;
    // Per-CU rate of the trailing GEMM (measured: deeper K reuses C better), panel step latency.
    const double rate = (sizeof(T) == 8 ? 0.145e12 : 0.27e12) * (wdt > nb ? 1.3 : 1.0);
    const double work = 2.0 * rows * static_cast<double>(n - c0) * wdt;
    const double t_panel = 4.4e-6 * next_wdt, t_solve = 7.5e-9 * (n - c0);
    const double t = (work / rate + P * t_panel + U * t_solve) / (P + U);
    const double w_tail = P * rate * (t - t_panel);
    if (w_tail <= 0.0 || rows <= 0) return n;
    const int cols = static_cast<int>(w_tail / (2.0 * rows * wdt)) / 128 * 128;
    if (cols < 128 || n - cols < c0 + nb) return n;  // the next panel's block stays on `us`
    return n - cols;
  }
  hipError_t wait_tail(hipStream_t st) {
    if (!tail.pending) return hipSuccess;
    tail.pending = false;
    return hipStreamWaitEvent(st, w.ev_tail, 0);
  }
  // Called right after the next panel is queued on `ps`.
  hipError_t launch_tail() {
    if (!tail.queued) return hipSuccess;
    tail.queued = false;
    hipError_t e = hipStreamWaitEvent(ps, w.ev_solved, 0);
    if (e == hipSuccess)
      e = gemm(tail.r0, tail.cb, tail.kc, tail.kc, m - tail.r0, n - tail.cb, tail.K, ps);
    if (e == hipSuccess) e = hipEventRecord(w.ev_tail, ps);
    if (e == hipSuccess) tail.pending = true;
    return e;
  }

  // The next panel, queued on `ps` behind everything `us` holds so far (the next panel's block),
  // as soon as that block's update is queued: the far work of the step is queued after it, so
  // the host's launch calls for the far work are off the panel chain.
  hipError_t launch_next(int pb_next, int pe_next, int positioned) {
    if (single || tall(pb_next)) return panel(pb_next, pe_next, positioned, us);
    hipError_t e = hipEventRecord(w.ev_next, us);
    if (e == hipSuccess) e = hipStreamWaitEvent(ps, w.ev_next, 0);
    if (e == hipSuccess) e = panel(pb_next, pe_next, positioned, ps);
    return e;
  }

  // What follows a panel [pb, pe) that eliminated all its columns.  Pair2: it closes a pair (its
  // near work is P0's interchange by its list and the pair's K = 2 nb update of the next block,
  // its far work the pair's left half).  Pair1: it opens one (the next block alone now; the
  // previous pair's right half follows).  Single: the next block, then every far column.
  // Whole: no look-ahead (last panel, a tall next panel, look-ahead off): everything, then the
  // next panel.
  enum class Step { Whole, Pair2, Pair1, Single };
  Step step_after(int pb, int pe, bool lookahead) const {
    const int pe2 = std::min(pe + nb, steps);
    if (pe - pb != nb || pe >= steps || !lookahead || tall(pe)) return Step::Whole;
    if (pair.open) return Step::Pair2;
    if (pe2 < n && m - pe2 >= merge_rows) return Step::Pair1;
    return Step::Single;
  }
  // The work the next panel waits for.
  hipError_t near(Step st, const Launch &L, hipStream_t q) {
    const int pe2 = std::min(L.pe + nb, steps);
    if (st == Step::Pair2) {
      hipError_t e = launch_row_interchange(a, lda, pair.b0, L.pb, L.list, nb, q, pred);
      // The far columns' interchanges and solve need P0's columns after this interchange only.
      if (e == hipSuccess && q != us) e = hipEventRecord(w.ev_pair, q);
      return e == hipSuccess ? pair_update(pair.b0, pair.l0, L.list, L.pe, pe2, q) : e;
    }
    return update(L.pb, nb, L.pe, pe2, L.list, q);
  }
  // The rest of the step, on `us`.  With the near work on the panel stream (near_on_ps), a
  // pair's second step waits there for P1's interchange on P0's columns (`ev_pair`) before its
  // interchanges and solve, and for the whole near work (`ev_near`) before its GEMM, so that the
  // next panel is resident before the GEMM fills the GPU.
  hipError_t far(Step st, const Launch &L, bool near_on_ps = false) {
    const int pb = L.pb, pe = L.pe, pe2 = std::min(pe + nb, steps);
    hipError_t e;
    if (st == Step::Pair2) {
      // The left half of the far columns (it holds the block after next); the right half is
      // queued after the next step's near work.
      const int b0 = pair.b0;
      pair.open = false;
      if ((e = queue_rest(0)) != hipSuccess) return e;  // (none pending: a pair just opened)
      if ((e = wait_tail(us)) != hipSuccess) return e;
      const int half = std::max(nb, (n - pe2) / 2 / nb * nb);
      const int cm = std::min(n, pe2 + half);
      if (near_on_ps && (e = hipStreamWaitEvent(us, w.ev_pair, 0)) != hipSuccess) return e;
      if (cm > pe2 && (e = pair_prepare(b0, pair.l0, L.list, pe2, cm, us)) != hipSuccess) return e;
      if (near_on_ps && (e = hipStreamWaitEvent(us, w.ev_near, 0)) != hipSuccess) return e;
      if (cm > pe2 && (e = gemm(b0 + 2 * nb, pe2, b0, b0, m - b0 - 2 * nb, cm - pe2, 2 * nb, us)) !=
                          hipSuccess)
        return e;
      rest = Rest{true, b0, cm, pair.l0, L.list};
      if (defer) {  // P0's columns took P1's movement in the near work
        blocks.push_back(Block{b0, pb, L.event + 1});
        blocks.push_back(Block{pb, pe, L.event + 1});
      }
      return hipSuccess;
    }
    if (st == Step::Pair1) {
      // Only the next panel's columns now (and the previous pair's right half, which the next
      // panel's block does not touch).
      if ((e = queue_rest(pe2 - pe)) != hipSuccess) return e;
      pair.open = true;
      pair.b0 = pb;
      pair.fe = pe2;
      pair.l0 = L.list;
      pair.e0 = L.event;
      return hipSuccess;
    }
    // The far columns while the next panel runs (their last columns on the panel stream behind
    // it when the GEMM outlasts it).
    if ((e = queue_rest(0)) != hipSuccess) return e;
    if ((e = wait_tail(us)) != hipSuccess) return e;
    if ((e = solve(pb, nb, pe2, n, L.list, us)) != hipSuccess) return e;
    const int cb = tail_split(m - pb - nb, nb, pe2, pe2 - pe);
    if (cb < n) {
      if ((e = hipEventRecord(w.ev_solved, us)) != hipSuccess) return e;
      tail = Tail{true, false, pb + nb, pb, nb, cb};
    }
    if ((e = gemm(pb + nb, pe2, pb, pb, m - pb - nb, cb - pe2, nb, us)) != hipSuccess) return e;
    return left(pb, pb + nb, L.event, L.list, us);
  }

  // Runs on the panel stream `ps` and the update stream `us`, forked from and joined back to the
  // caller's stream.  Per panel the host reads the control block once to learn where the panel
  // stopped; every pivot decision inside the panel is taken on the device.
  //
  // Speculation (two streams without CU partitions or panel-stream GEMM tails): every near work
  // goes on `ps` between the panels, and while panel j runs the host already queues panel j's
  // near work and panel j+1 behind it, on the guess that panel j eliminates all its columns.
  // Those launches are predicated on panel j's token (Control::go) and do nothing when the guess
  // fails, so the panel chain runs without a host round trip or a cross-stream wait between
  // panels.  `ev_far` (everything `us` holds) orders each near work after the far work before it;
  // a pair's second step records `ev_pair` after its first near kernel (P1's interchange on P0's
  // columns) and `ev_near` after the near work: its far interchanges and solve wait for the
  // first, its far GEMM for the second (the other far work reads nothing the near work writes,
  // and waits for the panel alone).  Between two panels `ps` holds one wait and one record
  // besides the kernels (three records after a pair's second panel).
  hipError_t factor(int *info, bool lookahead) {
    hipError_t e;
    if (single) lookahead = false;
    const bool speculate = lookahead && panel_cus == 0 && forced_split <= 0;
    if (!single && ((e = hipEventRecord(w.ev_fork, s)) != hipSuccess ||
                    (e = hipStreamWaitEvent(ps, w.ev_fork, 0)) != hipSuccess ||
                    (e = hipStreamWaitEvent(us, w.ev_fork, 0)) != hipSuccess ||
                    (e = hipEventRecord(w.ev_far, us)) != hipSuccess))
      return e;
    if ((e = panel(0, std::min(nb, steps), 0, tall(0) ? us : ps)) != hipSuccess) return e;
    Launch cur = launched, nxt;  // the panel awaited; the one queued behind it on speculation
    for (;;) {
      Control c;
      if ((e = wait_panel(cur, c)) != hipSuccess) return e;
      if (c.error != 0) return hipErrorLaunchFailure;
      const int pb = cur.pb, pe = cur.pe;
      const int stop = c.stop;
      const int wdt = stop - pb;
      const int pe2 = std::min(pe + nb, steps);
      if (!single && (e = hipStreamWaitEvent(us, w.ev_pan[cur.seq % kDoneSlots], 0)) != hipSuccess)
        return e;
      const bool whole = !c.deferred && !c.zero_column && wdt == pe - pb;
      const Step st = whole ? step_after(pb, pe, lookahead) : Step::Whole;
      if (nxt.valid && !(whole && st != Step::Whole)) {
        // The guess failed: the near work and the panel queued behind `cur` did nothing.  Nothing
        // else may touch the control block before they are through.
        if ((e = hipStreamWaitEvent(us, w.ev_pan[nxt.seq % kDoneSlots], 0)) != hipSuccess) return e;
        nxt = Launch{};
      }
      if (st != Step::Pair2) {
        if ((e = close_pair(us)) != hipSuccess) return e;
      }
      if (c.deferred) {
        if ((e = queue_rest(0)) != hipSuccess) return e;
        if ((e = wait_tail(us)) != hipSuccess) return e;
        *info = c.stop;
        break;
      }
      if (c.zero_column) {
        // Zero column at `stop`: bring the far columns up to date with the accepted prefix
        // (delayed_dgetrf.cpp:183-186), search the whole active block, interchange, restart.
        if ((e = queue_rest(0)) != hipSuccess) return e;
        if ((e = wait_tail(us)) != hipSuccess) return e;
        // The prefix's update of the far columns [pe, n) is the last write to that part of the
        // active block: its GEMM leaves the block's argmax there as per-workgroup partials, and
        // the search reads only the panel's own columns [stop, pe) (H4: materialize + search in
        // one pass).
        int ext = 0;
        if (wdt > 0) {
          const int gm = m - stop, gn = n - pe;
          const int wgs = gn > 0 && gm > 0 ? gemm_workgroups(tuned_gemm_config(gm, gn, wdt), gm, gn) : 0;
          if (variant == DLU_FALLBACK_FULL && w.d_gpv != nullptr && wgs > 0 &&
              wgs <= kAmaxPartials) {
            if ((e = solve(pb, wdt, pe, n, cur.list, us)) != hipSuccess) return e;
            if ((e = gemm(stop, pe, pb, pb, gm, gn, wdt, us,
                          AmaxOut{w.d_gpv, w.d_gpo, pe - stop, gm})) != hipSuccess)
              return e;
            ext = wgs;
          } else if ((e = update(pb, wdt, pe, n, cur.list, us)) != hipSuccess) {
            return e;
          }
          if ((e = left(pb, pb + wdt, cur.event, cur.list, us)) != hipSuccess) return e;
        }
        const int scan_end = ext > 0 ? pe : n;  // columns the search itself reads
        // The fallback interchanges whole rows: every owed left movement goes first.
        if ((e = apply_deferred(us)) != hipSuccess) return e;
        if (profile &&
            (e = mark(t_search, us, 0.0,
                      sizeof(T) * static_cast<double>(m - stop) *
                          static_cast<double>(scan_end - stop))) != hipSuccess)
          return e;
        e = variant == DLU_FALLBACK_COLUMN
                ? launch_column_search(a, lda, m, n, stop, w.d_ctl, us)
                : launch_active_search(a, lda, m, scan_end, stop, w.d_pv, w.d_po, w.d_ctl, us,
                                       w.d_gpv, w.d_gpo, ext);
        if (e != hipSuccess) return e;
        if (profile && (e = close(t_search, us)) != hipSuccess) return e;
        if ((e = launch_fallback_apply(a, lda, m, n, stop, w.d_ipiv, w.d_jpiv, w.d_ctl, us)) !=
            hipSuccess)
          return e;
        const int rb = stop, re = std::min(rb + nb, steps);
        // The restarted panel follows the fallback (nothing runs beside it): on the same stream,
        // or, when the panel stream owns a CU partition sized for panels, there, behind it.
        if (panel_cus > 0 && !tall(rb)) {
          if ((e = hipEventRecord(w.ev_next, us)) != hipSuccess) return e;
          if ((e = hipStreamWaitEvent(ps, w.ev_next, 0)) != hipSuccess) return e;
          if ((e = panel(rb, re, 1, ps)) != hipSuccess) return e;
        } else if ((e = panel(rb, re, 1, us)) != hipSuccess) {
          return e;
        }
        cur = launched;
      } else if (st == Step::Whole) {
        if ((e = queue_rest(0)) != hipSuccess) return e;
        if ((e = wait_tail(us)) != hipSuccess) return e;
        if ((e = update(pb, wdt, pe, n, cur.list, us)) != hipSuccess) return e;
        if ((e = left(pb, pb + wdt, cur.event, cur.list, us)) != hipSuccess) return e;
        if (pe >= steps) {
          *info = steps;
          break;
        }
        if ((e = launch_next(pe, pe2, 0)) != hipSuccess) return e;
        cur = launched;
      } else if (speculate) {
        if (!nxt.valid) {  // nothing queued behind `cur` yet: its near work and the next panel
          // (`cur` may have run on `us`: a restart after a fallback, or a tall first panel)
          if ((e = hipStreamWaitEvent(ps, w.ev_pan[cur.seq % kDoneSlots], 0)) != hipSuccess) return e;
          if ((e = hipStreamWaitEvent(ps, w.ev_far, 0)) != hipSuccess) return e;
          if ((e = near(st, cur, ps)) != hipSuccess) return e;
          if (st == Step::Pair2 && (e = hipEventRecord(w.ev_near, ps)) != hipSuccess) return e;
          if ((e = panel(pe, pe2, 0, ps)) != hipSuccess) return e;
          nxt = launched;
        }
        // Only a pair's second step has far work that reads what its near work wrote (P0's L
        // columns after P1's interchange, waited for inside); the others need the panel alone
        // (waited for above).
        if ((e = far(st, cur, true)) != hipSuccess) return e;
        if ((e = hipEventRecord(w.ev_far, us)) != hipSuccess) return e;
        cur = nxt;
        nxt = Launch{};
        // The guess for the panel now running: its near work and the panel after it.  (A tall
        // panel publishes no token, so nothing is guessed behind one.)
        const Step sn = step_after(cur.pb, cur.pe, lookahead);
        if (sn != Step::Whole && !tall(cur.pb)) {
          pred = Pred{&w.d_ctl->go, cur.seq};
          e = hipStreamWaitEvent(ps, w.ev_far, 0);
          if (e == hipSuccess) e = near(sn, cur, ps);
          if (e == hipSuccess && sn == Step::Pair2) e = hipEventRecord(w.ev_near, ps);
          if (e == hipSuccess) e = panel(cur.pe, std::min(cur.pe + nb, steps), 0, ps);
          pred = Pred{nullptr, 0};
          if (e != hipSuccess) return e;
          nxt = launched;
        }
      } else {
        // The next panel's block first (a pair's second panel: after P0's interchange by its
        // list), then the far columns while the next panel runs.
        if (st == Step::Pair2) {
          if ((e = queue_rest(0)) != hipSuccess) return e;
          if ((e = wait_tail(us)) != hipSuccess) return e;
        }
        if ((e = near(st, cur, us)) != hipSuccess) return e;
        if ((e = launch_next(pe, pe2, 0)) != hipSuccess) return e;
        const Launch next = launched;
        if ((e = far(st, cur)) != hipSuccess) return e;
        cur = next;
      }
      if (speculate && (e = hipEventRecord(w.ev_far, us)) != hipSuccess) return e;
      if ((e = launch_tail()) != hipSuccess) return e;
    }
    return apply_deferred(us);
  }

  // The caller's stream resumes after everything queued on the internal streams.
  hipError_t join() {
    if (single) return hipSuccess;
    hipError_t e = hipEventRecord(w.ev_join, ps);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, w.ev_join, 0);
    if (e == hipSuccess) e = hipEventRecord(w.ev_join, us);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, w.ev_join, 0);
    return e;
  }
};

// Every device resource of a workspace (streams, events, device and pinned host memory).
void release_workspace(Workspace &w) {
  if (!w.ready && w.d_ipiv == nullptr) return;
  for (int i = 0; i < w.nstreams; ++i) {
    if (w.streams[i].panel) hipStreamDestroy(w.streams[i].panel);
    if (w.streams[i].update) hipStreamDestroy(w.streams[i].update);
  }
  for (hipEvent_t ev : {w.ev_fork, w.ev_next, w.ev_join, w.ev_solved, w.ev_tail, w.ev_pan[0],
                        w.ev_pan[1], w.ev_pan[2], w.ev_pan[3], w.ev_near, w.ev_far, w.ev_pair})
    if (ev) hipEventDestroy(ev);
  for (void *p : {static_cast<void *>(w.d_ctl), static_cast<void *>(w.d_ipiv),
                  static_cast<void *>(w.d_jpiv), static_cast<void *>(w.d_margins),
                  static_cast<void *>(w.d_slots), static_cast<void *>(w.d_rowk),
                  static_cast<void *>(w.d_lists), static_cast<void *>(w.d_pv),
                  static_cast<void *>(w.d_po), static_cast<void *>(w.d_phase),
                  static_cast<void *>(w.d_gpv), static_cast<void *>(w.d_gpo),
                  static_cast<void *>(w.d_archive), static_cast<void *>(w.d_snaps),
                  static_cast<void *>(w.d_plan)})
    if (p) hipFree(p);
  if (w.h_plan) hipHostFree(w.h_plan);
  if (w.h_ctl) hipHostFree(w.h_ctl);
  if (w.h_done) hipHostFree(w.h_done);
  w = Workspace{};
}

}  // namespace

// Frees the idle workspaces of every device (busy ones belong to calls still running).
void release_pools() {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return;
  for (int d = 0; d < kMaxDevices; ++d) {
    DevicePool &p = g_pool[d];
    std::lock_guard<std::mutex> lk(p.mu);
    for (int i = 0; i < kPoolSize; ++i) {
      Workspace &w = p.ws[i];
      if (p.busy[i] || (!w.ready && w.d_ipiv == nullptr)) continue;
      if (hipSetDevice(w.device >= 0 ? w.device : d) != hipSuccess) continue;
      release_workspace(w);
    }
  }
  hipSetDevice(cur);
}

dlu_status to_status(hipError_t e) {
  if (e == hipSuccess) return DLU_SUCCESS;
  // DLU_VERBOSE=1 in the environment names the HIP error behind a failed call on stderr.
  static const bool verbose = [] {
    const char *v = std::getenv("DLU_VERBOSE");
    return v != nullptr && v[0] != '\0' && v[0] != '0';
  }();
  if (verbose) std::fprintf(stderr, "libdlu: HIP error %d (%s)\n", static_cast<int>(e),
                            hipGetErrorString(e));
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return DLU_ALLOCATION_FAILED;
  if (e == hipErrorLaunchFailure) return DLU_KERNEL_FAILURE;
  return DLU_RUNTIME_FAILURE;
}

// The part of a factorization shared by both entry points: control block, swap lists and step
// arrays initialised on the stream, the factorization, then the pivots, margins and report
// copied back once the stream has drained.
template <typename T>
dlu_status execute(Run<T> &run, bool lookahead, int *info, int *ipiv, int *jpiv,
                   double *margins, dlu_report *rep) {
  Workspace &w = run.w;
  run.track_margins = margins != nullptr;
  const int steps = run.steps;
  const hipStream_t stream = run.s;
  // `hc` stays alive until the stream has drained (its copies are ordered on the stream).
  Control hc{};
  hc.min_margin = 2.0;
  hc.col_found = -1;
  // Swap lists start empty with their append counters at zero (panels clear them after use;
  // this also recovers from an aborted call).
  hipError_t e;
  if ((e = hipMemcpyAsync(w.d_ctl, &hc, sizeof(Control), hipMemcpyHostToDevice, stream)) ==
          hipSuccess &&
      (e = hipMemsetAsync(w.d_lists, 0, kLists * kListStride * sizeof(int), stream)) == hipSuccess &&
      (!run.defer || (e = hipMemsetAsync(w.d_archive, 0,
                                         static_cast<size_t>(w.archive_cap) * kListStride * sizeof(int),
                                         stream)) == hipSuccess) &&
      (e = launch_init_steps(w.d_ipiv, w.d_jpiv, w.d_margins, steps, stream)) == hipSuccess) {
    int out_info = steps;
    e = run.factor(&out_info, lookahead);
    if (e == hipSuccess) *info = out_info;
  }
  // Whatever happened, leave no work of this call queued on the internal streams.
  hipError_t e2 = run.join();
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess)
    e = hipMemcpyAsync(ipiv, w.d_ipiv, steps * sizeof(int), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(jpiv, w.d_jpiv, steps * sizeof(int), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess && margins)
    e = hipMemcpyAsync(margins, w.d_margins, steps * sizeof(double), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&hc, w.d_ctl, sizeof(Control), hipMemcpyDeviceToHost, stream);
  e2 = hipStreamSynchronize(stream);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess && hc.error != 0) e = hipErrorLaunchFailure;  // a wait expired late
  if (e != hipSuccess) {
    *info = -1;
    run.release();
    return to_status(e);
  }
  if (rep) {
    const Control c = hc;
    rep->panels = run.panels;
    rep->fallbacks = c.fallbacks;
    rep->near_ties = c.near_ties;
    rep->exact_ties = c.exact_ties;
    rep->min_margin = c.min_margin;
    if (run.profile) {
      rep->large_update_launches = static_cast<int>(run.t_update_big.size());
      rep->large_update_seconds =
          Run<T>::total(run.t_update_big, &rep->large_update_flops, &rep->large_update_bytes);
      rep->update_launches = static_cast<int>(run.t_update.size()) + rep->large_update_launches;
      rep->update_seconds = Run<T>::total(run.t_update, &rep->update_flops) + rep->large_update_seconds;
      rep->update_flops += rep->large_update_flops;
      rep->panel_launches = static_cast<int>(run.t_panel.size());
      rep->panel_seconds = Run<T>::total(run.t_panel, nullptr);
      rep->solve_launches = static_cast<int>(run.t_solve.size());
      rep->solve_seconds = Run<T>::total(run.t_solve, nullptr);
      rep->search_launches = static_cast<int>(run.t_search.size());
      double search_flops = 0.0;
      rep->search_seconds = Run<T>::total(run.t_search, &search_flops, &rep->search_bytes);
      unsigned long long ticks[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (hipMemcpy(ticks, w.d_phase, sizeof ticks, hipMemcpyDeviceToHost) == hipSuccess)
        for (int i = 0; i < 8; ++i) rep->panel_phase_seconds[i] = ticks[i] * 1e-8;  // 100 MHz
    }
  }
  run.release();
  return DLU_SUCCESS;
}

template <typename T>
dlu_status getrf_device(int m, int n, T *d_a, int lda, int *ipiv, int *jpiv, int *info,
                        hipStream_t stream, const dlu_options *opt, dlu_report *rep,
                        double *margins) {
  if (info == nullptr) return DLU_INVALID_ARGUMENT;
  *info = -1;
  if (!valid_arguments(m, n, d_a, lda, ipiv, jpiv)) return DLU_INVALID_ARGUMENT;
  int nb = (opt && opt->nb) ? opt->nb : 128;
  if (nb != 64 && nb != 128 && nb != 256) return DLU_INVALID_ARGUMENT;
  const int variant = opt ? opt->fallback : DLU_FALLBACK_FULL;
  if (variant != DLU_FALLBACK_FULL && variant != DLU_FALLBACK_COLUMN) return DLU_INVALID_ARGUMENT;
  const int forced_rpl = opt ? opt->rows_per_lane : 0;
  if (forced_rpl < 0 || forced_rpl > 4) return DLU_INVALID_ARGUMENT;
  const int panel_kernel = opt ? opt->panel_kernel : DLU_PANEL_AUTO;
  if (panel_kernel != DLU_PANEL_AUTO && panel_kernel != DLU_PANEL_MOVE &&
      panel_kernel != DLU_PANEL_LAG)
    return DLU_INVALID_ARGUMENT;
  const bool lookahead = !(opt && opt->lookahead < 0);
  const long long spin =
      (opt && opt->spin_limit_ticks > 0) ? opt->spin_limit_ticks : 200000000ll;
  const int steps = std::min(m, n);
  *info = steps;
  for (int k = 0; k < steps; ++k) ipiv[k] = jpiv[k] = k;
  if (rep) {
    std::memset(rep, 0, sizeof *rep);
    rep->min_margin = 2.0;
  }
  if (margins)
    for (int k = 0; k < steps; ++k) margins[k] = -1.0;
  if (steps == 0) return DLU_SUCCESS;
  // Panels taller than 256 workgroups x 128 rows need the narrower panel kernels.
  if (nb == 256 && m > 32768) nb = 128;

  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess || device < 0 || device >= kMaxDevices) {
    *info = -1;
    return DLU_RUNTIME_FAILURE;
  }
  int cus = 0;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) !=
          hipSuccess ||
      cus <= 0) {
    *info = -1;
    return DLU_RUNTIME_FAILURE;
  }
  // CUs reserved for the panel stream: none by default, FP64 and FP32 alike (measured at
  // n = 16384: the trailing update gains more from the whole device than the panels lose by
  // waiting for update workgroups to drain from a CU, and without a partition the panel chain
  // runs speculatively; FP32 70.5 ms unpartitioned against 72.9 ms with one CU per panel
  // workgroup).  opt->panel_cus > 0 forces a size (rounded down to whole XCDs).
  // A partition is made of whole XCDs: a stream's CU mask fills the XCDs in order (32 mask bits
  // each on MI355X), and a launch's workgroups are dealt round robin over the XCDs its mask
  // touches, so a partly masked XCD would receive as many panel workgroups as a full one and
  // could not hold them all at once (measured: 40 panel workgroups on a 48-CU mask never became
  // co-resident; on 32 and 64 CUs they do).
  int groups = 0, rpl = 0;
  const bool lag0 = panel_kernel != DLU_PANEL_MOVE &&
                    lag_shape_for(m, nb, sizeof(T), forced_rpl, cus, groups, rpl);
  if (panel_kernel == DLU_PANEL_LAG && !lag0) return DLU_INVALID_ARGUMENT;
  if (!lag0) panel_shape_for(m, nb, forced_rpl, std::min(kMaxPanelGroups, cus), groups, rpl);
  const int xcd = cus % kCusPerXcd == 0 ? kCusPerXcd : cus;  // unknown layout: no partition
  int want = 0;
  if (opt && opt->panel_cus > 0) want = std::min(opt->panel_cus, cus) / xcd * xcd;
  if (opt && opt->panel_cus < 0) want = 0;
  if (want >= cus) want = 0;                // no CUs left for the updates: share the device
  if (groups > want && want > 0) want = 0;  // too tall for a partition: share the device
  Lease lease(device, std::min(groups, cus), want > 0, cus);
  Workspace &w = lease.workspace();
  if ((e = prepare(w, device, steps)) != hipSuccess) {
    *info = -1;
    return to_status(e);
  }
  Run<T> run{w, m, n, lda, steps, nb, d_a, stream, spin};
  run.profile = opt && opt->profile;
  run.forced_rpl = forced_rpl;
  run.variant = variant;
  run.panel_kernel = panel_kernel;
  if (opt && opt->merge_rows != 0) run.merge_rows = opt->merge_rows < 0 ? (1 << 30) : opt->merge_rows;
  run.split = !(opt && opt->split_update < 0);
  if (opt && opt->split_update > 0) run.forced_split = opt->split_update;
  {
    run.max_groups = std::min(kMaxPanelGroups, w.cus);
    Workspace::Streams *st = nullptr;
    if ((e = stream_pair(w, want, st)) != hipSuccess) {
      *info = -1;
      return to_status(e);
    }
    run.ps = st->panel;
    run.us = st->update;
    run.panel_cus = std::max(0, st->panel_cus);
    if (st->panel_cus > 0) run.max_groups = std::min(kMaxPanelGroups, st->panel_cus);
  }
  if (run.profile && (e = hipMemsetAsync(w.d_phase, 0, 8 * sizeof(unsigned long long), stream)) !=
                         hipSuccess) {
    *info = -1;
    return to_status(e);
  }
  // Left interchanges deferred to one pass at the end: the panels' lists stay (one per panel),
  // and each column is permuted once (through LDS: at most kDeferMaxRows rows).
  if (m <= kDeferMaxRows && !run.tall(0) && !(opt && opt->defer_left < 0)) {
    // (one more: a panel queued on speculation behind the first zero column takes a list too)
    const int lists = (steps + nb - 1) / nb + 1;
    run.defer = prepare_defer(w, m, n, lists) == hipSuccess && w.archive_cap >= lists;
  }
  return execute(run, lookahead, info, ipiv, jpiv, margins, rep);
}

// ---- The allocation-free entry point --------------------------------------------------------
// Every device buffer a factorization needs besides the caller's matrix lives in one file-scope
// device array, laid out below; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the array's name is not a device address
// on the host).  The factorization then runs in order on the caller's stream with no events, no
// pinned host memory and no look-ahead, and reads the control block with a plain copy.
constexpr int kNoAllocMaxSteps = 32768;  // min(m, n) at most
constexpr size_t na_align(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kNaCtl = 0;
constexpr size_t kNaIpiv = kNaCtl + na_align(sizeof(Control));
constexpr size_t kNaJpiv = kNaIpiv + na_align(sizeof(int) * kNoAllocMaxSteps);
constexpr size_t kNaMargins = kNaJpiv + na_align(sizeof(int) * kNoAllocMaxSteps);
constexpr size_t kNaSlots = kNaMargins + na_align(sizeof(double) * kNoAllocMaxSteps);
constexpr size_t kNaRowk = kNaSlots + na_align(slot_bytes());
constexpr size_t kNaLists = kNaRowk + na_align(rowk_bytes());
constexpr size_t kNaPv = kNaLists + na_align(sizeof(int) * kLists * kListStride);
constexpr size_t kNaPo = kNaPv + na_align(sizeof(double) * kSearchPartials);
constexpr size_t kNaPhase = kNaPo + na_align(sizeof(unsigned long long) * kSearchPartials);
constexpr size_t kNoAllocScratchBytes = kNaPhase + na_align(8 * sizeof(unsigned long long));

// About 2.7 MB per device, held until the process ends.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocDevice {
  std::mutex mu;  // calls on one device take turns: they share the array
  bool ready = false;
  Workspace w;    // views into this device's copy of the array
};
NoAllocDevice g_noalloc[kMaxDevices];

dlu_status getrf_noalloc(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv, int *info,
                         hipStream_t stream) {
  if (info == nullptr) return DLU_INVALID_ARGUMENT;
  *info = -1;
  if (!valid_arguments(m, n, d_a, lda, ipiv, jpiv)) return DLU_INVALID_ARGUMENT;
  const int steps = std::min(m, n);
  if (steps > kNoAllocMaxSteps) return DLU_INVALID_ARGUMENT;  // beyond the array's step arrays
  *info = steps;
  for (int k = 0; k < steps; ++k) ipiv[k] = jpiv[k] = k;
  if (steps == 0) return DLU_SUCCESS;
  int device = 0, cus = 0;
  hipError_t e = hipGetDevice(&device);
  if (e == hipSuccess && (device < 0 || device >= kMaxDevices)) e = hipErrorInvalidDevice;
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess || cus <= 0) {
    *info = -1;
    return to_status(e != hipSuccess ? e : hipErrorInvalidDevice);
  }
  NoAllocDevice &d = g_noalloc[device];
  std::lock_guard<std::mutex> lk(d.mu);
  if (!d.ready) {
    // Once per device, with that device current: this device's address of the array.
    void *ptr = nullptr;
    if ((e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch))) != hipSuccess || ptr == nullptr) {
      *info = -1;
      if (e == hipSuccess) e = hipErrorInvalidSymbol;  // a null address is no address
      return to_status(e);
    }
    char *b = static_cast<char *>(ptr);
    Workspace &w = d.w;
    w.device = device;
    w.cus = cus;
    w.steps_cap = kNoAllocMaxSteps;
    w.d_ctl = reinterpret_cast<Control *>(b + kNaCtl);
    w.d_ipiv = reinterpret_cast<int *>(b + kNaIpiv);
    w.d_jpiv = reinterpret_cast<int *>(b + kNaJpiv);
    w.d_margins = reinterpret_cast<double *>(b + kNaMargins);
    w.d_slots = reinterpret_cast<uint64_t *>(b + kNaSlots);
    w.d_rowk = reinterpret_cast<uint64_t *>(b + kNaRowk);
    w.d_lists = reinterpret_cast<int *>(b + kNaLists);
    w.d_pv = reinterpret_cast<double *>(b + kNaPv);
    w.d_po = reinterpret_cast<unsigned long long *>(b + kNaPo);
    w.d_phase = reinterpret_cast<unsigned long long *>(b + kNaPhase);
    // Exchange words start with step tag 0, which no step uses.
    if ((e = hipMemsetAsync(w.d_slots, 0, slot_bytes(), stream)) != hipSuccess ||
        (e = hipMemsetAsync(w.d_rowk, 0, rowk_bytes(), stream)) != hipSuccess) {
      *info = -1;
      return to_status(e);
    }
    d.ready = true;
  }
  Run<double> run{d.w, m, n, lda, steps, 128, d_a, stream, 200000000ll};
  run.single = true;
  run.split = false;
  run.merge_rows = 1 << 30;
  run.ps = run.us = stream;
  run.max_groups = std::min(kMaxPanelGroups, cus);
  return execute(run, false, info, ipiv, jpiv, nullptr, nullptr);
}

dlu_status getrf_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info,
                      hipStream_t stream, int variant) {
  if (info == nullptr) return DLU_INVALID_ARGUMENT;
  *info = -1;
  if (!valid_arguments(m, n, a, lda, ipiv, jpiv)) return DLU_INVALID_ARGUMENT;
  if (std::min(m, n) == 0) {
    *info = 0;  // delayed_lu_cuda/delayed_dgetrf_cuda.cu:456-460
    return DLU_SUCCESS;
  }
  size_t bytes = 0;
  if (!matrix_bytes(lda, n, bytes)) return DLU_INVALID_ARGUMENT;
  double *d_a = nullptr;
  if (hipMalloc(&d_a, bytes) != hipSuccess) return DLU_ALLOCATION_FAILED;
  std::unique_ptr<double, hipError_t (*)(void *)> hold(d_a, hipFree);
  if (hipMemcpyAsync(d_a, a, bytes, hipMemcpyHostToDevice, stream) != hipSuccess)
    return DLU_RUNTIME_FAILURE;
  dlu_options opt{};
  opt.fallback = variant;
  const dlu_status st = getrf_device(m, n, d_a, lda, ipiv, jpiv, info, stream, &opt, nullptr,
                                     nullptr);
  if (st != DLU_SUCCESS) return st;
  if (hipMemcpyAsync(a, d_a, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess) {
    *info = -1;
    return DLU_RUNTIME_FAILURE;
  }
  return DLU_SUCCESS;
}

template dlu_status getrf_device<float>(int, int, float *, int, int *, int *, int *, hipStream_t,
                                        const dlu_options *, dlu_report *, double *);
template dlu_status getrf_device<double>(int, int, double *, int, int *, int *, int *,
                                         hipStream_t, const dlu_options *, dlu_report *,
                                         double *);

}  // namespace dlu

extern "C" {

dlu_status dlu_getrf_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv, int *info,
                            void *stream) {
  return dlu::getrf_device(m, n, d_a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream),
                           nullptr, nullptr, nullptr);
}

dlu_status dlu_getrf_device_ex(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                               int *info, void *stream, const dlu_options *options,
                               dlu_report *report, double *margins) {
  return dlu::getrf_device(m, n, d_a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream),
                           options, report, margins);
}

dlu_status dlu_sgetrf_device(int m, int n, float *d_a, int lda, int *ipiv, int *jpiv, int *info,
                             void *stream) {
  return dlu::getrf_device(m, n, d_a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream),
                           nullptr, nullptr, nullptr);
}

dlu_status dlu_sgetrf_device_ex(int m, int n, float *d_a, int lda, int *ipiv, int *jpiv,
                                int *info, void *stream, const dlu_options *options,
                                dlu_report *report, double *margins) {
  return dlu::getrf_device(m, n, d_a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream),
                           options, report, margins);
}

dlu_status dlu_getrf_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv, int *info,
                          void *stream) {
  return dlu::getrf_host(m, n, a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream),
                         DLU_FALLBACK_FULL);
}

dlu_status dlu_getrf_native_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                   int *info, void *stream) {
  return dlu_getrf_device(m, n, d_a, lda, ipiv, jpiv, info, stream);
}

dlu_status dlu_getrf_native_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                 int *info, void *stream) {
  return dlu_getrf_host(m, n, a, lda, ipiv, jpiv, info, stream);
}

dlu_status dlu_getrf_hybrid_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                   int *info, void *stream) {
  return dlu_getrf_device(m, n, d_a, lda, ipiv, jpiv, info, stream);
}

dlu_status dlu_getrf_hybrid_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                 int *info, void *stream) {
  return dlu_getrf_host(m, n, a, lda, ipiv, jpiv, info, stream);
}

dlu_status dlu_getrf_column_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                   int *info, void *stream) {
  dlu_options opt{};
  opt.fallback = DLU_FALLBACK_COLUMN;
  return dlu::getrf_device(m, n, d_a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream),
                           &opt, nullptr, nullptr);
}

dlu_status dlu_getrf_column_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                 int *info, void *stream) {
  return dlu::getrf_host(m, n, a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream),
                         DLU_FALLBACK_COLUMN);
}

dlu_status dlu_getrf_device_noalloc(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                    int *info, void *stream) {
  return dlu::getrf_noalloc(m, n, d_a, lda, ipiv, jpiv, info, static_cast<hipStream_t>(stream));
}

void dlu_shutdown(void) {
  dlu::release_pools();
  dlu::release_thread_solve_cache();
}

const char *dlu_status_string(dlu_status status) {
  switch (status) {
    case DLU_SUCCESS: return "success";
    case DLU_INVALID_ARGUMENT: return "invalid argument";
    case DLU_ALLOCATION_FAILED: return "HIP allocation failed";
    case DLU_RUNTIME_FAILURE: return "HIP runtime failure";
    case DLU_KERNEL_FAILURE: return "kernel failure";
  }
  return "unknown status";
}

dlu_status dlu_gemm_sub_device(int M, int N, int K, double *C, const double *A, const double *B,
                               int ld, void *stream) {
  return dlu_gemm_sub_device_config(M, N, K, C, A, B, ld, -1, stream);
}

dlu_status dlu_gemm_sub_device_config(int M, int N, int K, double *C, const double *A,
                                      const double *B, int ld, int config, void *stream) {
  if (M < 0 || N < 0 || K < 0 || ld < std::max(1, M) || (M > 0 && N > 0 && C == nullptr) ||
      (M > 0 && K > 0 && A == nullptr) || (K > 0 && N > 0 && B == nullptr) || ld < K ||
      config >= dlu::kGemmConfigs)
    return DLU_INVALID_ARGUMENT;
  return dlu::to_status(dlu::launch_gemm_sub<double>(C, ld, A, ld, B, ld, M, N, K,
                                                     static_cast<hipStream_t>(stream), config));
}

int dlu_tuned_gemm_config(int M, int N, int K) {
  return (M < 1 || N < 1 || K < 1) ? -1 : dlu::tuned_gemm_config(M, N, K);
}

}  // extern "C"
