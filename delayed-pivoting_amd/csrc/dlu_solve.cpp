// Host drivers of the solve side: getrs with the delayed-pivoting factors, the FP64 backward
// error, the plain FP64 solver and the mixed-precision solver (FP32 factor + FP64 iterative
// refinement).  None of these has a counterpart in the reference, which factors only; they are
// the solve and refinement the factorization is used for.
#include "dlu.h"
#include "dlu_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace dlu {
namespace {

constexpr int kChainRows = 64;         // rows per block of the one-launch triangular solves
constexpr int kChainRhs = 16;          // right-hand sides per triangular-solve launch
constexpr long long kChainSpin = 200000000ll;  // bounded waits: 2 s of the 100 MHz clock

// Device buffers of the solve side, grown on demand and kept per device.
struct SolveCache {
  enum { A32, A64, V64A, V64B, V32A, V32B, NORMW, NORMS, PERM, WORK, CHAIN, WORDS, COUNT };
  void *buf[COUNT] = {};
  size_t cap[COUNT] = {};
  unsigned epoch = 0;                  // last epoch handed to a triangular-solve launch
  unsigned long long ticket_base = 0;  // tickets taken so far from the CHAIN counter
  bool chain_fresh = false;            // CHAIN or WORDS was (re)allocated: not cleared yet
  hipError_t get(int i, size_t bytes, void **out) {
    if (bytes == 0) bytes = 16;
    if (cap[i] < bytes) {
      if (buf[i]) hipFree(buf[i]);
      buf[i] = nullptr;
      cap[i] = 0;
      hipError_t e = hipMalloc(&buf[i], bytes);
      if (e != hipSuccess) return e;
      cap[i] = bytes;
      if (i == CHAIN || i == WORDS) chain_fresh = true;
    }
    *out = buf[i];
    return hipSuccess;
  }
  // Chain state for one triangular-solve launch over `blocks` row blocks, on stream s.  CHAIN
  // holds the ticket counter (8 bytes) and the error word; WORDS the tagged exchange words (16
  // columns x 2 words per padded row).  Fresh areas are cleared on s itself, so the clearing is
  // ordered before the launch whatever kind of stream s is.
  hipError_t chain(int blocks, ChainState &cs, hipStream_t s) {
    void *p = nullptr, *pw = nullptr;
    hipError_t e = get(CHAIN, 16, &p);
    if (e == hipSuccess)
      e = get(WORDS, sizeof(unsigned long long) * 32 * 64 * static_cast<size_t>(blocks), &pw);
    if (e != hipSuccess) return e;
    if (chain_fresh || epoch == 0xFFFFFFFFu) {  // new areas, or the tags would repeat an epoch
      if ((e = hipMemsetAsync(p, 0, cap[CHAIN], s)) != hipSuccess) return e;
      if ((e = hipMemsetAsync(pw, 0, cap[WORDS], s)) != hipSuccess) return e;
      chain_fresh = false;
      epoch = 0;
      ticket_base = 0;
    }
    cs.ticket = static_cast<unsigned long long *>(p);
    cs.err = reinterpret_cast<int *>(static_cast<char *>(p) + 8);
    cs.words = static_cast<unsigned long long *>(pw);
    cs.ticket_base = ticket_base;
    cs.epoch = ++epoch;
    cs.spin_limit = kChainSpin;
    ticket_base += static_cast<unsigned long long>(blocks);
    return hipSuccess;
  }
  // After the stream has drained: did a triangular-solve wait expire?  Clears the flag.
  hipError_t chain_error(hipStream_t s, bool *failed) {
    *failed = false;
    if (!buf[CHAIN]) return hipSuccess;
    int h = 0;
    int *d_err = reinterpret_cast<int *>(static_cast<char *>(buf[CHAIN]) + 8);
    hipError_t e = hipMemcpyAsync(&h, d_err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && h != 0) {
      *failed = true;
      chain_fresh = true;  // the counter may have drifted from ticket_base: clear before reuse
    }
    return e;
  }
};
// Per host thread (the serving workers and callers each keep their own scratch).  No
// destructor frees it: a thread-local destructor may run after the HIP runtime is gone.  The
// serving workers free theirs before they exit, callers through dlu_shutdown.
thread_local SolveCache t_cache[64];

// Net row order of P b (swaps in step order) and net order of Q y (swaps in reverse order):
// (P b)[i] = b[prow[i]],  (Q y)[i] = y[pcol[i]].
void build_perms(int n, const int *ipiv, const int *jpiv, std::vector<int> &both) {
  both.resize(2 * static_cast<size_t>(n));
  int *prow = both.data(), *pcol = both.data() + n;
  for (int i = 0; i < n; ++i) prow[i] = pcol[i] = i;
  for (int k = 0; k < n; ++k) std::swap(prow[k], prow[ipiv[k]]);
  for (int k = n - 1; k >= 0; --k) std::swap(pcol[k], pcol[jpiv[k]]);
}

bool perms_valid(int n, const int *ipiv, const int *jpiv) {
  for (int k = 0; k < n; ++k)
    if (ipiv[k] < k || ipiv[k] >= n || jpiv[k] < k || jpiv[k] >= n) return false;
  return true;
}

// X = Q U^{-1} L^{-1} P B, in place in b: a gather into work (rows padded to a multiple of
// 64), one launch per triangle and per 16 right-hand sides, a gather back.  chain(k, cs) fills
// the state of the k-th triangular-solve launch.
template <typename T, typename ChainFn>
hipError_t getrs_chains(int n, int nrhs, const T *lu, int lda, const int *d_prow,
                        const int *d_pcol, T *b, int ldb, T *work, ChainFn &&chain,
                        hipStream_t s) {
  const int npad = (n + kChainRows - 1) / kChainRows * kChainRows;
  hipError_t e = launch_gather_rows<T>(b, ldb, work, npad, d_prow, n, nrhs, s);
  int k = 0;
  for (int c0 = 0; e == hipSuccess && c0 < nrhs; c0 += kChainRhs) {
    const int nc = std::min(kChainRhs, nrhs - c0);
    T *xc = work + static_cast<size_t>(c0) * npad;
    ChainState cs{};
    if ((e = chain(k++, cs)) == hipSuccess)
      e = launch_trsv_chain<T>(lu, lda, n, xc, npad, nc, false, cs, s);
    if (e == hipSuccess && (e = chain(k++, cs)) == hipSuccess)
      e = launch_trsv_chain<T>(lu, lda, n, xc, npad, nc, true, cs, s);
  }
  if (e == hipSuccess) e = launch_gather_rows<T>(work, npad, b, ldb, d_pcol, n, nrhs, s);
  return e;
}

// getrs_chains with its scratch (work and the chain state) from the cache.
template <typename T>
hipError_t getrs_core(int n, int nrhs, const T *lu, int lda, const int *d_prow, const int *d_pcol,
                      T *b, int ldb, SolveCache &c, hipStream_t s) {
  const int npad = (n + kChainRows - 1) / kChainRows * kChainRows;
  const int blocks = npad / kChainRows;
  void *pw = nullptr;
  hipError_t e = c.get(SolveCache::WORK, sizeof(T) * static_cast<size_t>(npad) * nrhs, &pw);
  if (e != hipSuccess) return e;
  return getrs_chains<T>(n, nrhs, lu, lda, d_prow, d_pcol, b, ldb, static_cast<T *>(pw),
                         [&](int, ChainState &cs) { return c.chain(blocks, cs, s); }, s);
}

bool solve_args(int n, int nrhs, const void *lu, int lda, const int *ipiv, const int *jpiv,
                int info, const void *b, int ldb) {
  if (n < 0 || nrhs < 0 || lda < std::max(1, n) || ldb < std::max(1, n)) return false;
  if (n == 0 || nrhs == 0) return true;
  if (lu == nullptr || b == nullptr || ipiv == nullptr || jpiv == nullptr) return false;
  return info == n && perms_valid(n, ipiv, jpiv);
}

template <typename T>
dlu_status getrs_device(int n, int nrhs, const T *d_lu, int lda, const int *ipiv,
                        const int *jpiv, int info, T *d_b, int ldb, hipStream_t s) {
  if (!solve_args(n, nrhs, d_lu, lda, ipiv, jpiv, info, d_b, ldb)) return DLU_INVALID_ARGUMENT;
  if (n == 0 || nrhs == 0) return DLU_SUCCESS;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) return DLU_RUNTIME_FAILURE;
  SolveCache &c = t_cache[device];
  std::vector<int> perms;
  build_perms(n, ipiv, jpiv, perms);
  void *pp = nullptr;
  hipError_t e = c.get(SolveCache::PERM, 2 * sizeof(int) * static_cast<size_t>(n), &pp);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pp, perms.data(), 2 * sizeof(int) * static_cast<size_t>(n),
                       hipMemcpyHostToDevice, s);
  const int *d_perm = static_cast<const int *>(pp);
  if (e == hipSuccess)
    e = getrs_core<T>(n, nrhs, d_lu, lda, d_perm, d_perm + n, d_b, ldb, c, s);
  bool failed = false;
  if (e == hipSuccess) e = c.chain_error(s, &failed);
  if (e == hipSuccess && failed) return DLU_KERNEL_FAILURE;
  return to_status(e);
}

// r = b - A x and the column norms of x, r and b into norms + 1 (zeroed before).
hipError_t queue_residual_norms(int n, int nrhs, const double *a, int lda, const double *x,
                                int ldx, const double *b, int ldb, double *r, double *norms,
                                hipStream_t s) {
  double *xn = norms + 1, *rn = norms + 1 + nrhs, *bn = norms + 1 + 2 * nrhs;
  hipError_t e = launch_residual(a, lda, x, ldx, b, ldb, r, n, n, nrhs, s);
  if (e == hipSuccess) e = launch_norm_inf_columns(x, ldx, n, nrhs, xn, s);
  if (e == hipSuccess) e = launch_norm_inf_columns(r, n, n, nrhs, rn, s);
  if (e == hipSuccess) e = launch_norm_inf_columns(b, ldb, n, nrhs, bn, s);
  return e;
}

// Per column j of X: ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf), from the
// norms [||A||, ||x_0..||, ||r_0..||, ||b_0..||] on the host.
void bwd_from_norms(const double *h, int nrhs, double *bwd) {
  for (int j = 0; j < nrhs; ++j) {
    const double den = h[0] * h[1 + j] + h[1 + 2 * nrhs + j];
    bwd[j] = den > 0.0 ? h[1 + nrhs + j] / den : (h[1 + nrhs + j] > 0.0 ? 1.0 : 0.0);
  }
}

// The backward errors of X (see bwd_from_norms).  norms must hold 3*nrhs+1 doubles and r
// n*nrhs doubles; normw 32*n doubles.  Results land in host bwd.
hipError_t backward_error(int n, int nrhs, const double *a, int lda, const double *x, int ldx,
                          const double *b, int ldb, double *r, double *normw, double *norms,
                          bool have_anorm, double *bwd, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (!have_anorm) {
    e = hipMemsetAsync(norms, 0, sizeof(double), s);
    if (e == hipSuccess) e = launch_norm_inf_matrix(a, lda, n, normw, norms, s);
  }
  if (e == hipSuccess) e = hipMemsetAsync(norms + 1, 0, 3 * sizeof(double) * nrhs, s);
  if (e == hipSuccess) e = queue_residual_norms(n, nrhs, a, lda, x, ldx, b, ldb, r, norms, s);
  std::vector<double> h(3 * static_cast<size_t>(nrhs) + 1);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), norms, h.size() * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  bwd_from_norms(h.data(), nrhs, bwd);
  return hipSuccess;
}

struct Timer {
  hipEvent_t e[4] = {};
  bool ok = true;
  Timer() {
    for (hipEvent_t &x : e) ok = ok && hipEventCreate(&x) == hipSuccess;
  }
  ~Timer() {
    for (hipEvent_t x : e)
      if (x) hipEventDestroy(x);
  }
  void mark(int i, hipStream_t s) {
    if (ok) hipEventRecord(e[i], s);
  }
  double seconds(int i, int j) {
    float ms = 0.0f;
    if (!ok || hipEventElapsedTime(&ms, e[i], e[j]) != hipSuccess) return 0.0;
    return ms * 1e-3;
  }
};

bool gesv_args(int n, int nrhs, const void *a, int lda, const void *b, int ldb, const void *x,
               int ldx) {
  if (n < 0 || nrhs < 0 || lda < std::max(1, n) || ldb < std::max(1, n) ||
      ldx < std::max(1, n))
    return false;
  return n == 0 || nrhs == 0 || (a && b && x);
}

}  // namespace

// Factor an n x n FP64 matrix for later solves.  lu32 / lu64 have room for n*n floats / doubles
// (leading dimension n), d_perm for 2n ints.  fp32: factor a float copy into lu32; when that
// leaves a delayed suffix and lu64 is given, factor an FP64 copy into lu64 instead.
dlu_status make_factor(int n, const double *d_a, int lda, bool fp32, int nb, float *lu32,
                       double *lu64, int *d_perm, DeviceFactor &f, int *first_info,
                       hipStream_t s) {
  dlu_options fopt{};
  fopt.nb = nb;
  std::vector<int> ipiv(n), jpiv(n), perms;
  int info = -1;
  hipError_t e;
  f = DeviceFactor{};
  f.n = n;
  f.d_perm = d_perm;
  if (fp32) {
    if ((e = launch_to_float(d_a, lda, lu32, n, n, n, s)) != hipSuccess) return to_status(e);
    const dlu_status st = getrf_device<float>(n, n, lu32, n, ipiv.data(), jpiv.data(), &info, s,
                                              &fopt, nullptr, nullptr);
    if (st != DLU_SUCCESS) return st;
    if (first_info) *first_info = info;
    f.fp32 = true;
    f.lu = lu32;
    f.info = info;
    if (info != n && lu64 == nullptr) return DLU_SUCCESS;  // the caller decides what to do
  }
  if (!fp32 || info != n) {
    if (lu64 == nullptr) return DLU_INVALID_ARGUMENT;
    if ((e = hipMemcpy2DAsync(lu64, n * sizeof(double), d_a, lda * sizeof(double),
                              n * sizeof(double), n, hipMemcpyDeviceToDevice, s)) != hipSuccess)
      return to_status(e);
    const dlu_status st = getrf_device<double>(n, n, lu64, n, ipiv.data(), jpiv.data(), &info, s,
                                               &fopt, nullptr, nullptr);
    if (st != DLU_SUCCESS) return st;
    if (first_info && !fp32) *first_info = info;
    f.fp32 = false;
    f.lu = lu64;
    f.info = info;
  }
  if (info == n) {
    build_perms(n, ipiv.data(), jpiv.data(), perms);
    if ((e = hipMemcpyAsync(d_perm, perms.data(), 2 * sizeof(int) * static_cast<size_t>(n),
                            hipMemcpyHostToDevice, s)) != hipSuccess)
      return to_status(e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return to_status(e);  // perms is local
  }
  return DLU_SUCCESS;
}

// x := the solution of A x = b, refined in FP64 with the factor f: x = 0, r = b; then
// x += solve(r) and r = b - A x until every column's backward error is at most tol.  Scratch
// memory is per host thread.
dlu_status solve_refine(const DeviceFactor &f, int nrhs, const double *d_a, int lda,
                        const double *d_b, int ldb, double *d_x, int ldx, double tol,
                        int max_iter, hipStream_t s, int *iterations, double *bwd,
                        bool *converged) {
  const int n = f.n;
  if (f.info != n) return DLU_INVALID_ARGUMENT;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) return DLU_RUNTIME_FAILURE;
  SolveCache &c = t_cache[device];
  const size_t nr = static_cast<size_t>(n) * nrhs;
  void *p_r = nullptr, *p_r32 = nullptr, *p_nw = nullptr, *p_norms = nullptr;
  hipError_t e = c.get(SolveCache::V64A, nr * sizeof(double), &p_r);
  if (e == hipSuccess) e = c.get(SolveCache::V32A, nr * sizeof(float), &p_r32);
  if (e == hipSuccess) e = c.get(SolveCache::NORMW, 32 * static_cast<size_t>(n) * sizeof(double), &p_nw);
  if (e == hipSuccess) e = c.get(SolveCache::NORMS, (3 * static_cast<size_t>(nrhs) + 1) * sizeof(double), &p_norms);
  if (e != hipSuccess) return to_status(e);
  double *r = static_cast<double *>(p_r), *normw = static_cast<double *>(p_nw);
  double *norms = static_cast<double *>(p_norms);
  float *r32 = static_cast<float *>(p_r32);
  std::vector<double> be(nrhs, 1.0);
  if ((e = hipMemset2DAsync(d_x, ldx * sizeof(double), 0, n * sizeof(double), nrhs, s)) !=
          hipSuccess ||
      (e = hipMemcpy2DAsync(r, n * sizeof(double), d_b, ldb * sizeof(double), n * sizeof(double),
                            nrhs, hipMemcpyDeviceToDevice, s)) != hipSuccess ||
      (e = hipMemsetAsync(norms, 0, sizeof(double), s)) != hipSuccess ||
      (e = launch_norm_inf_matrix(d_a, lda, n, normw, norms, s)) != hipSuccess)
    return to_status(e);
  int it = 0;
  bool conv = false;
  while (it < max_iter) {
    ++it;
    if (f.fp32) {
      e = launch_to_float(r, n, r32, n, n, nrhs, s);
      if (e == hipSuccess)
        e = getrs_core<float>(n, nrhs, static_cast<const float *>(f.lu), n, f.d_perm,
                              f.d_perm + n, r32, n, c, s);
      if (e == hipSuccess) e = launch_add_correction<float>(d_x, ldx, r32, n, n, nrhs, s);
    } else {
      e = getrs_core<double>(n, nrhs, static_cast<const double *>(f.lu), n, f.d_perm,
                             f.d_perm + n, r, n, c, s);
      if (e == hipSuccess) e = launch_add_correction<double>(d_x, ldx, r, n, n, nrhs, s);
    }
    if (e == hipSuccess)
      e = backward_error(n, nrhs, d_a, lda, d_x, ldx, d_b, ldb, r, normw, norms, true, be.data(),
                         s);
    if (e != hipSuccess) return to_status(e);
    if (*std::max_element(be.begin(), be.end()) <= tol) {
      conv = true;
      break;
    }
  }
  bool failed = false;
  if ((e = c.chain_error(s, &failed)) != hipSuccess) return to_status(e);
  if (failed) return DLU_KERNEL_FAILURE;
  if (iterations) *iterations = it;
  if (converged) *converged = conv;
  if (bwd)
    for (int j = 0; j < nrhs; ++j) bwd[j] = be[j];
  return DLU_SUCCESS;
}

// ---- refinement sweeps replayed from a HIP graph ------------------------------------------

struct RefinePlan {
  int n = 0, nrhs = 0;
  bool fp32 = false;
  const void *lu = nullptr;
  const int *d_perm = nullptr;
  const double *a = nullptr;
  int lda = 0;
  void *mem = nullptr;  // one stream-ordered allocation holds every buffer below
  size_t bytes = 0;
  ChainClocks *st = nullptr;  // followed by the norms: one device-to-host copy reads both
  double *norms = nullptr;
  size_t report_bytes = 0;    // st .. the end of the norms
  double *B = nullptr, *X = nullptr, *R = nullptr;  // n x nrhs each, leading dimension n
  float *R32 = nullptr;
  void *work = nullptr;       // npad x nrhs of the factor's type
  unsigned long long *words = nullptr;
  size_t nwords = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<char> host;     // the report on the host
  bool broken = false;        // a wait expired: the counter may have drifted, never replay
};

namespace {

// One sweep, as solve_refine issues it, onto a capturing stream.
hipError_t record_sweep(RefinePlan &p, hipStream_t s) {
  const int n = p.n, nrhs = p.nrhs;
  const int blocks = (n + kChainRows - 1) / kChainRows;
  const int slots = 2 * ((nrhs + kChainRhs - 1) / kChainRhs);
  auto chain = [&](int k, ChainState &cs) {
    cs.words = p.words;
    cs.ticket = &p.st->ticket;
    cs.err = &p.st->err;
    cs.clock = &p.st->slot[k];
    cs.spin_limit = kChainSpin;
    return k < slots ? hipSuccess : hipErrorInvalidValue;
  };
  hipError_t e = launch_chain_clock(p.st, slots, blocks, p.words, p.nwords, p.norms + 1,
                                    3 * nrhs, s);
  if (e != hipSuccess) return e;
  if (p.fp32) {
    e = launch_to_float(p.R, n, p.R32, n, n, nrhs, s);
    if (e == hipSuccess)
      e = getrs_chains<float>(n, nrhs, static_cast<const float *>(p.lu), n, p.d_perm,
                              p.d_perm + n, p.R32, n, static_cast<float *>(p.work), chain, s);
    if (e == hipSuccess) e = launch_add_correction<float>(p.X, n, p.R32, n, n, nrhs, s);
  } else {
    e = getrs_chains<double>(n, nrhs, static_cast<const double *>(p.lu), n, p.d_perm,
                             p.d_perm + n, p.R, n, static_cast<double *>(p.work), chain, s);
    if (e == hipSuccess) e = launch_add_correction<double>(p.X, n, p.R, n, n, nrhs, s);
  }
  if (e == hipSuccess)
    e = queue_residual_norms(n, nrhs, p.a, p.lda, p.X, n, p.B, n, p.R, p.norms, s);
  return e;
}

}  // namespace

bool refine_plan_matches(const RefinePlan *p, int nrhs, const double *d_a, int lda) {
  return p && !p->broken && p->nrhs == nrhs && p->a == d_a && p->lda == lda;
}

void refine_plan_destroy(RefinePlan *p, hipStream_t s) {
  if (!p) return;
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  if (p->mem) hipFreeAsync(p->mem, s);
  delete p;
}

dlu_status refine_plan_create(const DeviceFactor &f, int nrhs, const double *d_a, int lda,
                              hipStream_t s, RefinePlan **out, size_t *bytes) {
  *out = nullptr;
  const int n = f.n;
  if (f.info != n || n < 1 || nrhs < 1 || nrhs > kPlanMaxRhs || !d_a || lda < n)
    return DLU_INVALID_ARGUMENT;
  RefinePlan *p = new RefinePlan;
  p->n = n;
  p->nrhs = nrhs;
  p->fp32 = f.fp32;
  p->lu = f.lu;
  p->d_perm = f.d_perm;
  p->a = d_a;
  p->lda = lda;
  const size_t nr = static_cast<size_t>(n) * nrhs;
  const int npad = (n + kChainRows - 1) / kChainRows * kChainRows;
  const int blocks = npad / kChainRows;
  auto al = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
  const size_t o_norms = al(sizeof(ChainClocks));
  p->report_bytes = o_norms + sizeof(double) * (3 * static_cast<size_t>(nrhs) + 1);
  const size_t o_b = al(p->report_bytes);
  const size_t o_x = o_b + al(nr * sizeof(double));
  const size_t o_r = o_x + al(nr * sizeof(double));
  const size_t o_r32 = o_r + al(nr * sizeof(double));
  const size_t o_work = o_r32 + (f.fp32 ? al(nr * sizeof(float)) : 0);
  p->nwords = static_cast<size_t>(32) * 64 * blocks;
  const size_t o_words =
      o_work + al(static_cast<size_t>(npad) * nrhs * (f.fp32 ? sizeof(float) : sizeof(double)));
  p->bytes = o_words + al(p->nwords * sizeof(unsigned long long));
  double *normw = nullptr;  // the matrix norm's scratch, for the creation only
  hipError_t e = hipMallocAsync(&p->mem, p->bytes, s);
  if (e == hipSuccess) {
    e = hipMallocAsync(reinterpret_cast<void **>(&normw), 32 * sizeof(double) * n, s);
    if (e != hipSuccess) hipFreeAsync(p->mem, s);
  }
  if (e != hipSuccess) {
    delete p;
    return DLU_ALLOCATION_FAILED;
  }
  char *m = static_cast<char *>(p->mem);
  p->st = reinterpret_cast<ChainClocks *>(m);
  p->norms = reinterpret_cast<double *>(m + o_norms);
  p->B = reinterpret_cast<double *>(m + o_b);
  p->X = reinterpret_cast<double *>(m + o_x);
  p->R = reinterpret_cast<double *>(m + o_r);
  p->R32 = f.fp32 ? reinterpret_cast<float *>(m + o_r32) : nullptr;
  p->work = m + o_work;
  p->words = reinterpret_cast<unsigned long long *>(m + o_words);
  p->host.resize(p->report_bytes);
  // Eager set-up on s: the counter, epochs and words start at zero, ||A|| goes to norms[0].
  e = hipMemsetAsync(m, 0, p->report_bytes, s);
  if (e == hipSuccess) e = hipMemsetAsync(p->words, 0, p->nwords * sizeof(unsigned long long), s);
  if (e == hipSuccess) e = launch_norm_inf_matrix(d_a, lda, n, normw, p->norms, s);
  const hipError_t ef = hipFreeAsync(normw, s);
  if (e == hipSuccess) e = ef;
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  // The sweep, recorded on a stream of its own (the plan's graph can then run on any stream).
  hipStream_t cap = nullptr;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (e == hipSuccess) {
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t er = record_sweep(*p, cap);
      e = hipStreamEndCapture(cap, &p->graph);
      if (er != hipSuccess) e = er;
    }
    hipStreamDestroy(cap);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    refine_plan_destroy(p, s);
    return to_status(e);
  }
  *out = p;
  if (bytes) *bytes = p->bytes;
  return DLU_SUCCESS;
}

dlu_status refine_plan_solve(RefinePlan *p, const double *d_b, int ldb, double *d_x, int ldx,
                             double tol, int max_iter, hipStream_t s, int *iterations,
                             double *bwd, bool *converged) {
  if (!p || p->broken || !d_b || !d_x || ldb < p->n || ldx < p->n) return DLU_INVALID_ARGUMENT;
  const int n = p->n, nrhs = p->nrhs;
  const size_t col = static_cast<size_t>(n) * sizeof(double);
  hipError_t e = hipMemcpy2DAsync(p->B, col, d_b, ldb * sizeof(double), col, nrhs,
                                  hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(p->X, 0, col * nrhs, s);
  if (e == hipSuccess) e = hipMemcpyAsync(p->R, p->B, col * nrhs, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return to_status(e);
  std::vector<double> be(nrhs, 1.0), h(3 * static_cast<size_t>(nrhs) + 1);
  const size_t o_norms = reinterpret_cast<char *>(p->norms) - reinterpret_cast<char *>(p->st);
  int it = 0;
  bool conv = false;
  while (it < max_iter) {
    ++it;
    if ((e = hipGraphLaunch(p->exec, s)) != hipSuccess ||
        (e = hipMemcpyAsync(p->host.data(), p->st, p->report_bytes, hipMemcpyDeviceToHost, s)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return to_status(e);
    if (reinterpret_cast<const ChainClocks *>(p->host.data())->err != 0) {
      p->broken = true;
      return DLU_KERNEL_FAILURE;
    }
    std::memcpy(h.data(), p->host.data() + o_norms, h.size() * sizeof(double));
    bwd_from_norms(h.data(), nrhs, be.data());
    if (*std::max_element(be.begin(), be.end()) <= tol) {
      conv = true;
      break;
    }
  }
  e = hipMemcpy2DAsync(d_x, ldx * sizeof(double), p->X, col, col, nrhs, hipMemcpyDeviceToDevice,
                       s);
  if (e != hipSuccess) return to_status(e);
  if (iterations) *iterations = it;
  if (converged) *converged = conv;
  if (bwd)
    for (int j = 0; j < nrhs; ++j) bwd[j] = be[j];
  return DLU_SUCCESS;
}

namespace {

// Plain FP64 (use_fp32 = false) or mixed precision (use_fp32 = true): factor, solve, refine.
dlu_status gesv(int n, int nrhs, const double *d_a, int lda, const double *d_b, int ldb,
                double *d_x, int ldx, hipStream_t s, const dlu_mixed_options *opt,
                dlu_mixed_report *rep, double *bwd, bool use_fp32) {
  if (rep) std::memset(rep, 0, sizeof *rep);
  if (!gesv_args(n, nrhs, d_a, lda, d_b, ldb, d_x, ldx)) return DLU_INVALID_ARGUMENT;
  const double tol = (opt && opt->tol > 0.0) ? opt->tol : 1e-13;
  const int max_iter = (opt && opt->max_iter > 0) ? opt->max_iter : 30;
  const bool fallback = !(opt && opt->fp64_fallback < 0);
  if (n == 0 || nrhs == 0) {
    if (rep) rep->converged = 1;
    return DLU_SUCCESS;
  }
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) return DLU_RUNTIME_FAILURE;
  SolveCache &c = t_cache[device];
  const size_t nn = static_cast<size_t>(n) * n;
  void *p32 = nullptr, *p64 = nullptr, *pperm = nullptr;
  hipError_t e = c.get(SolveCache::PERM, 2 * sizeof(int) * static_cast<size_t>(n), &pperm);
  if (e == hipSuccess && use_fp32) e = c.get(SolveCache::A32, nn * sizeof(float), &p32);
  if (e == hipSuccess && (!use_fp32 || fallback))
    e = c.get(SolveCache::A64, nn * sizeof(double), &p64);
  if (e != hipSuccess) return to_status(e);
  Timer tm;
  tm.mark(0, s);
  DeviceFactor f;
  int first_info = -1;
  dlu_status st = make_factor(n, d_a, lda, use_fp32, opt ? opt->nb : 0,
                              static_cast<float *>(p32), static_cast<double *>(p64),
                              static_cast<int *>(pperm), f, &first_info, s);
  if (st != DLU_SUCCESS) return st;
  if (rep) {
    rep->factor_info = first_info;
    rep->used_fp64 = f.fp32 ? 0 : 1;
  }
  if (f.info != n) return DLU_INVALID_ARGUMENT;  // singular to working precision: no solve
  tm.mark(1, s);
  int it = 0;
  bool conv = false;
  std::vector<double> be(nrhs, 1.0);
  st = solve_refine(f, nrhs, d_a, lda, d_b, ldb, d_x, ldx, tol, max_iter, s, &it, be.data(),
                    &conv);
  if (st != DLU_SUCCESS) return st;
  tm.mark(2, s);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return to_status(e);
  if (bwd)
    for (int j = 0; j < nrhs; ++j) bwd[j] = be[j];
  if (rep) {
    rep->iterations = it;
    rep->converged = conv ? 1 : 0;
    rep->factor_seconds = tm.seconds(0, 1);
    rep->refine_seconds = tm.seconds(1, 2);
    rep->total_seconds = tm.seconds(0, 2);
    rep->max_backward_error = *std::max_element(be.begin(), be.end());
  }
  return DLU_SUCCESS;
}

}  // namespace

void release_thread_solve_cache() {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return;
  for (int d = 0; d < 64; ++d) {
    SolveCache &c = t_cache[d];
    bool any = false;
    for (void *p : c.buf) any = any || p != nullptr;
    if (!any || hipSetDevice(d) != hipSuccess) continue;
    hipDeviceSynchronize();  // a solve of this thread may still be queued
    for (void *p : c.buf)
      if (p) hipFree(p);
    c = SolveCache{};
  }
  hipSetDevice(cur);
}

size_t thread_solve_cache_bytes(int device) {
  if (device < 0 || device >= 64) return 0;
  size_t total = 0;
  for (size_t b : t_cache[device].cap) total += b;
  return total;
}

}  // namespace dlu

extern "C" {

dlu_status dlu_getrs_device(int n, int nrhs, const double *d_lu, int lda, const int *ipiv,
                            const int *jpiv, int info, double *d_b, int ldb, void *stream) {
  return dlu::getrs_device<double>(n, nrhs, d_lu, lda, ipiv, jpiv, info, d_b, ldb,
                                   static_cast<hipStream_t>(stream));
}

dlu_status dlu_sgetrs_device(int n, int nrhs, const float *d_lu, int lda, const int *ipiv,
                             const int *jpiv, int info, float *d_b, int ldb, void *stream) {
  return dlu::getrs_device<float>(n, nrhs, d_lu, lda, ipiv, jpiv, info, d_b, ldb,
                                  static_cast<hipStream_t>(stream));
}

dlu_status dlu_backward_error_device(int n, int nrhs, const double *d_a, int lda,
                                     const double *d_x, int ldx, const double *d_b, int ldb,
                                     double *bwd, void *stream) {
  if (!dlu::gesv_args(n, nrhs, d_a, lda, d_b, ldb, d_x, ldx) || (nrhs > 0 && n > 0 && !bwd))
    return DLU_INVALID_ARGUMENT;
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) bwd[j] = 0.0;
    return DLU_SUCCESS;
  }
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) return DLU_RUNTIME_FAILURE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  dlu::SolveCache &c = dlu::t_cache[device];
  void *p_r = nullptr, *p_nw = nullptr, *p_norms = nullptr;
  hipError_t e = c.get(dlu::SolveCache::V64A, sizeof(double) * static_cast<size_t>(n) * nrhs, &p_r);
  if (e == hipSuccess) e = c.get(dlu::SolveCache::NORMW, 32 * static_cast<size_t>(n) * sizeof(double), &p_nw);
  if (e == hipSuccess) e = c.get(dlu::SolveCache::NORMS, (3 * static_cast<size_t>(nrhs) + 1) * sizeof(double), &p_norms);
  if (e == hipSuccess)
    e = dlu::backward_error(n, nrhs, d_a, lda, d_x, ldx, d_b, ldb, static_cast<double *>(p_r),
                            static_cast<double *>(p_nw), static_cast<double *>(p_norms), false,
                            bwd, s);
  return dlu::to_status(e);
}

dlu_status dlu_gesv_device(int n, int nrhs, const double *d_a, int lda, const double *d_b,
                           int ldb, double *d_x, int ldx, void *stream,
                           const dlu_mixed_options *options, dlu_mixed_report *report,
                           double *bwd) {
  return dlu::gesv(n, nrhs, d_a, lda, d_b, ldb, d_x, ldx, static_cast<hipStream_t>(stream),
                   options, report, bwd, false);
}

dlu_status dlu_gesv_mixed_device(int n, int nrhs, const double *d_a, int lda, const double *d_b,
                                 int ldb, double *d_x, int ldx, void *stream,
                                 const dlu_mixed_options *options, dlu_mixed_report *report,
                                 double *bwd) {
  return dlu::gesv(n, nrhs, d_a, lda, d_b, ldb, d_x, ldx, static_cast<hipStream_t>(stream),
                   options, report, bwd, true);
}

}  // extern "C"
