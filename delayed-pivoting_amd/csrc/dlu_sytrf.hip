// Delayed Bunch-Kaufman LDL^T of a symmetric matrix on gfx950: the GPU form of the reference's
// blocked delayed_sytrf_blocked, which exists on the CPU only.
//
// Reference behaviour restated here (delayed-bunch-kaufman/):
//   the Bunch-Kaufman constant (1 + sqrt 17) / 8, the 2x2 tolerance 64 eps   delayed_sytrf.cpp:16-23
//   lower-only symmetric read and the four-segment symmetric interchange      :28-67
//   the three tests choosing a 1x1 pivot, a 1x1 after an interchange, or a
//   2x2 block; a 2x2 block accepted only when its scaled determinant clears
//   the tolerance; a node that takes neither goes to the front of the
//   delayed suffix [active_end, n)                                             :71-106, :197-277
//   the lazy panel S = A - L W^T (W = L D), a column or an active row formed
//   on demand, W's rows interchanged with A's                                  delayed_sytrf_blocked.cpp:61-123
//   panels of at most 64 columns that never split a 2x2 block, then one
//   lower update A22 -= 1/2 (L21 W21^T + W21 L21^T) over every trailing row
//   (the delayed nodes keep receiving their Schur complement)                  :142-164, :192-355
// Only the lower triangle is read or written.
//
// Layout: the caller's column-major matrix, in place; W (n x nb, leading dimension n), three
// length-n vectors (two columns and one row of the Schur complement), perm and piv_size on the
// device, and a control block.  One panel is one launch of a single 1024-thread workgroup, which
// makes every decision on the device: at a step it forms column k over rows [k, n) from the
// stored values and the panel's L W^T (one thread per row, the panel columns in order), reduces
// the argmax, forms the candidate row when the tests need it, interchanges (every element of
// the four segments, and W's rows, in one pass), and writes L, D and W.  The host reads where
// the panel ended, then launches the lower trailing update: one workgroup per 64 x 64 tile on
// or below the diagonal, both products on v_mfma_f64_16x16x4_f64, diagonal tiles written on
// and below the diagonal only.
#include "dlu.h"
#include "dlu_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace dlu {
namespace {

constexpr int kSyPanelThreads = 1024;
constexpr int kSyMaxNb = 64;

struct SyControl {
  int k;           // the next position to factor (the panel's first column at launch)
  int active_end;  // [active_end, n) is the delayed suffix
  int pw;          // columns the panel accepted
  int pad;
};

__device__ __forceinline__ double bk_alpha() { return (1.0 + sqrt(17.0)) / 8.0; }

__device__ __forceinline__ bool bk_stable_2x2(double a, double b, double c) {
  const double scale = fmax(fmax(fabs(a), fabs(b)), fabs(c));
  if (!(scale > 0.0)) return false;
  const double as = a / scale, bs = b / scale, cs = c / scale;
  const double det = as * cs - bs * bs;
  const double nrm = fmax(fabs(as) + fabs(bs), fabs(bs) + fabs(cs));
  return fabs(det) > 64.0 * 2.220446049250313e-16 * nrm * nrm;
}

__device__ __forceinline__ double *at(double *a, int lda, int i, int j) {
  return a + static_cast<size_t>(i) + static_cast<size_t>(j) * static_cast<size_t>(lda);
}

// out[r] = A(r, c) - sum_{j < pw} L(r, pb + j) W(c, j) for rows r in [r0, n), A's column c
// read below the diagonal (r >= c).  One thread per row; the panel's columns in order, from the
// stored value down (the order of a column-major dgemv).
__device__ void form_column(const double *a, int lda, int n, int pb, int pw, int c, int r0,
                            double *out, const double *s_wc) {
  for (int r = r0 + static_cast<int>(threadIdx.x); r < n; r += kSyPanelThreads) {
    double v = a[static_cast<size_t>(r) + static_cast<size_t>(c) * lda];
    for (int j = 0; j < pw; ++j)
      v = fma(-a[static_cast<size_t>(r) + static_cast<size_t>(pb + j) * lda], s_wc[j], v);
    out[r] = v;
  }
}

// Block-wide reduction to thread 0's view: the largest |x| with the lowest index among equals
// (a NaN never wins), or the plain maximum of |x|.
__device__ void block_argmax(double v, int idx, double *s_v, int *s_i, double &best, int &best_i) {
  const int t = threadIdx.x;
  s_v[t] = v;
  s_i[t] = idx;
  __syncthreads();
  for (int s = kSyPanelThreads / 2; s > 0; s >>= 1) {
    if (t < s) {
      const double o = s_v[t + s];
      const int oi = s_i[t + s];
      if (o > s_v[t] || (o == s_v[t] && oi < s_i[t])) {
        s_v[t] = o;
        s_i[t] = oi;
      }
    }
    __syncthreads();
  }
  best = s_v[0];
  best_i = s_i[0];
  __syncthreads();
}

// Symmetric interchange of nodes x < y in lower storage (rows x, y left of x; column x below x
// against row y between them; the diagonals; columns x, y below y), W's rows over the panel's
// pw columns, and perm.  The element sets are disjoint, so the whole workgroup does all of them
// in one pass.
__device__ void sym_swap(double *a, int lda, int n, int x, int y, double *w, int ldw, int pw,
                         int *perm) {
  if (x == y) return;
  if (x > y) {
    const int t = x;
    x = y;
    y = t;
  }
  const int t = threadIdx.x;
  for (int j = t; j < x; j += kSyPanelThreads) {
    double *p = at(a, lda, x, j), *q = at(a, lda, y, j);
    const double v = *p;
    *p = *q;
    *q = v;
  }
  for (int i = x + 1 + t; i < y; i += kSyPanelThreads) {
    double *p = at(a, lda, i, x), *q = at(a, lda, y, i);
    const double v = *p;
    *p = *q;
    *q = v;
  }
  for (int i = y + 1 + t; i < n; i += kSyPanelThreads) {
    double *p = at(a, lda, i, x), *q = at(a, lda, i, y);
    const double v = *p;
    *p = *q;
    *q = v;
  }
  for (int j = t; j < pw; j += kSyPanelThreads) {
    double *p = w + x + static_cast<size_t>(j) * ldw, *q = w + y + static_cast<size_t>(j) * ldw;
    const double v = *p;
    *p = *q;
    *q = v;
  }
  if (t == 0) {
    double *p = at(a, lda, x, x), *q = at(a, lda, y, y);
    const double v = *p;
    *p = *q;
    *q = v;
    const int pv = perm[x];
    perm[x] = perm[y];
    perm[y] = pv;
  }
  __syncthreads();
}

// One panel of the blocked delayed Bunch-Kaufman factorization (delayed_sytrf_blocked.cpp:196-349)
// starting at ctl->k, every decision on the device.  Leaves the panel's end in ctl.
__global__ __launch_bounds__(kSyPanelThreads) void sytrf_panel_kernel(
    double *a, int lda, int n, double *w, int ldw, double *col0, double *col1, double *rowv,
    int *perm, int *piv, SyControl *ctl, int nb) {
  __shared__ double s_v[kSyPanelThreads];
  __shared__ int s_i[kSyPanelThreads];
  __shared__ double s_wc[kSyMaxNb];  // the W row of the column being formed
  __shared__ int s_kind, s_to;
  const int t = threadIdx.x;
  const int pb = ctl->k;
  int active_end = ctl->active_end;
  int k = pb, pw = 0;
  auto load_wrow = [&](int row) {  // W's row `row` to LDS, once every reader of the last is done
    __syncthreads();
    if (t < pw) s_wc[t] = w[row + static_cast<size_t>(t) * ldw];
    __syncthreads();
  };
  while (k < active_end && pw < nb) {
    // The current column over rows [k, n).
    load_wrow(k);
    form_column(a, lda, n, pb, pw, k, k, col0, s_wc);
    __syncthreads();
    // Its largest |value| below the diagonal inside the active part (idamax: first index).
    double bv = -1.0;
    int bi = k + 1;
    for (int r = k + 1 + t; r < active_end; r += kSyPanelThreads) {
      const double v = fabs(col0[r]);
      if (v > bv) {
        bv = v;
        bi = r;
      }
    }
    double colmax;
    int p;
    block_argmax(bv, bv < 0.0 ? 0x7fffffff : bi, s_v, s_i, colmax, p);
    if (active_end - k <= 1 || colmax < 0.0) {
      colmax = 0.0;
      p = k;
    }
    const double akk = fabs(col0[k]);
    const double alpha = bk_alpha();
    int kind = -1, to = k;  // -1: the candidate row is needed
    if (!(fmax(akk, colmax) > 0.0)) {
      kind = 0;
    } else if (akk >= alpha * colmax) {
      kind = 1;
    }
    if (kind < 0) {
      // Row p over the active columns [k, active_end): A(p, c) read symmetrically, minus
      // sum_j L(c, pb + j) W(p, j).
      load_wrow(p);
      for (int c = k + t; c < active_end; c += kSyPanelThreads) {
        double v = c <= p ? *at(a, lda, p, c) : *at(a, lda, c, p);
        for (int j = 0; j < pw; ++j)
          v = fma(-a[static_cast<size_t>(c) + static_cast<size_t>(pb + j) * lda], s_wc[j], v);
        rowv[c] = v;
      }
      __syncthreads();
      double rv = 0.0;
      for (int c = k + t; c < active_end; c += kSyPanelThreads)
        if (c != p) rv = fmax(rv, fabs(rowv[c]));
      double rowmax;
      int unused;
      block_argmax(rv, 0, s_v, s_i, rowmax, unused);
      if (t == 0) {
        const double app = fabs(rowv[p]);
        if (akk >= alpha * colmax * (colmax / rowmax)) {
          s_kind = 1;
          s_to = k;
        } else if (app >= alpha * rowmax) {
          s_kind = 1;
          s_to = p;
        } else if (bk_stable_2x2(col0[k], col0[p], rowv[p])) {
          s_kind = 2;
          s_to = p;
        } else {
          s_kind = 0;
          s_to = k;
        }
      }
      __syncthreads();
      kind = s_kind;
      to = s_to;
      __syncthreads();
    }
    if (kind == 0) {  // to the front of the delayed suffix; position k is examined again
      sym_swap(a, lda, n, k, active_end - 1, w, ldw, pw, perm);
      --active_end;
      continue;
    }
    if (kind == 2 && pw + 2 > nb) break;  // a 2x2 block never straddles two panels
    if (kind == 1) {
      if (to != k) {
        sym_swap(a, lda, n, k, to, w, ldw, pw, perm);
        load_wrow(k);
        form_column(a, lda, n, pb, pw, k, k, col0, s_wc);
        __syncthreads();
      }
      const double d = col0[k];
      for (int r = k + t; r < n; r += kSyPanelThreads) {
        const double b = col0[r];
        *at(a, lda, r, k) = r == k ? d : b / d;
        w[r + static_cast<size_t>(pw) * ldw] = b;
      }
      if (t == 0) piv[k] = 1;
      __syncthreads();
      ++k;
      ++pw;
      continue;
    }
    // 2x2: node `to` to k + 1, both columns formed before either is overwritten.
    sym_swap(a, lda, n, k + 1, to, w, ldw, pw, perm);
    load_wrow(k);
    form_column(a, lda, n, pb, pw, k, k, col0, s_wc);
    load_wrow(k + 1);
    form_column(a, lda, n, pb, pw, k + 1, k + 1, col1, s_wc);
    __syncthreads();
    const double d00 = col0[k], d10 = col0[k + 1], d11 = col1[k + 1];
    const double sc = fmax(fmax(fabs(d00), fabs(d10)), fabs(d11));
    const double s00 = d00 / sc, s10 = d10 / sc, s11 = d11 / sc;
    const double inv = 1.0 / (sc * (s00 * s11 - s10 * s10));
    double *w0 = w + static_cast<size_t>(pw) * ldw, *w1 = w + static_cast<size_t>(pw + 1) * ldw;
    for (int r = k + 2 + t; r < n; r += kSyPanelThreads) {
      const double b0 = col0[r], b1 = col1[r];
      *at(a, lda, r, k) = (s11 * b0 - s10 * b1) * inv;
      *at(a, lda, r, k + 1) = (s00 * b1 - s10 * b0) * inv;
      w0[r] = b0;
      w1[r] = b1;
    }
    if (t == 0) {
      *at(a, lda, k, k) = d00;
      *at(a, lda, k + 1, k) = d10;
      *at(a, lda, k + 1, k + 1) = d11;
      w0[k] = d00;
      w0[k + 1] = d10;
      w1[k] = d10;
      w1[k + 1] = d11;
      piv[k] = 2;
      piv[k + 1] = 0;
    }
    __syncthreads();
    k += 2;
    pw += 2;
  }
  if (t == 0) {
    ctl->k = k;
    ctl->active_end = active_end;
    ctl->pw = pw;
  }
}

// A22 -= 1/2 (L21 W21^T + W21 L21^T) on and below the diagonal of the trailing part [end, n):
// L21 = A(end:n, pb:pb+pw), W21 = W(end:n, 0:pw).  One workgroup per 64 x 64 tile (ti >= tj),
// wave w its 16 rows, four 16-column MFMA tiles; the tile index is decoded from the linear
// block index over the lower-triangular tile set.
__global__ __launch_bounds__(256) void sytrf_update_kernel(double *a, int lda, int n, int end,
                                                           int pb, int pw, const double *w,
                                                           int ldw) {
  typedef Mfma<double> MF;
  const long long b = blockIdx.x;
  long long ti = static_cast<long long>((sqrt(8.0 * static_cast<double>(b) + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > b) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
  const long long tj = b - ti * (ti + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int r0 = end + 64 * static_cast<int>(ti) + 16 * wave;
  const int c0 = end + 64 * static_cast<int>(tj);
  const int row = r0 + lr;
  const bool rok = row < n;
  MF::acc_t acc[4];
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) acc[ct] = MF::acc_t{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < pw; k0 += 4) {
    const int kk = k0 + lk;
    const bool kok = kk < pw;
    const double al = (rok && kok) ? a[static_cast<size_t>(row) + static_cast<size_t>(pb + kk) * lda] : 0.0;
    const double aw = (rok && kok) ? w[static_cast<size_t>(row) + static_cast<size_t>(kk) * ldw] : 0.0;
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const int col = c0 + 16 * ct + lr;
      const bool cok = col < n && kok;
      const double bw = cok ? w[static_cast<size_t>(col) + static_cast<size_t>(kk) * ldw] : 0.0;
      const double bl = cok ? a[static_cast<size_t>(col) + static_cast<size_t>(pb + kk) * lda] : 0.0;
      acc[ct] = MF::mma(al, bw, acc[ct]);
      acc[ct] = MF::mma(aw, bl, acc[ct]);
    }
  }
#pragma unroll
  for (int ct = 0; ct < 4; ++ct)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int orow = r0 + MF::acc_row(lane, r), ocol = c0 + 16 * ct + lr;
      if (orow < n && ocol <= orow) *at(a, lda, orow, ocol) -= 0.5 * acc[ct][r];
    }
}

__global__ void sytrf_init_kernel(int *perm, int *piv, int n, SyControl *ctl) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    perm[i] = i;
    piv[i] = 0;
  }
  if (i == 0) *ctl = SyControl{0, n, 0, 0};
}

dlu_status sytrf_device(int n, double *d_a, int lda, int *perm, int *piv_size, int *info,
                        int nb, hipStream_t s) {
  if (info == nullptr) return DLU_INVALID_ARGUMENT;
  *info = -1;
  if (n < 0 || lda < std::max(1, n) || nb < 2 || nb > kSyMaxNb ||
      (n > 0 && (d_a == nullptr || perm == nullptr || piv_size == nullptr)))
    return DLU_INVALID_ARGUMENT;
  if (static_cast<double>(lda) * n * 8.0 > 9.0e18) return DLU_INVALID_ARGUMENT;
  if (n == 0) {
    *info = 0;
    return DLU_SUCCESS;
  }
  const size_t nn = static_cast<size_t>(n);
  const size_t bytes = sizeof(double) * nn * (static_cast<size_t>(nb) + 3) +
                       2 * sizeof(int) * nn + 256;
  char *mem = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&mem), bytes, s);
  if (e != hipSuccess) return DLU_ALLOCATION_FAILED;
  SyControl *d_ctl = reinterpret_cast<SyControl *>(mem);
  double *w = reinterpret_cast<double *>(mem + 256);
  double *col0 = w + nn * nb, *col1 = col0 + nn, *rowv = col1 + nn;
  int *d_perm = reinterpret_cast<int *>(rowv + nn), *d_piv = d_perm + nn;
  hipLaunchKernelGGL(sytrf_init_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_perm, d_piv, n,
                     d_ctl);
  e = hipGetLastError();
  SyControl h{0, n, 0, 0};
  while (e == hipSuccess && h.k < h.active_end) {
    const int pb = h.k;
    hipLaunchKernelGGL(sytrf_panel_kernel, dim3(1), dim3(kSyPanelThreads), 0, s, d_a, lda, n, w,
                       n, col0, col1, rowv, d_perm, d_piv, d_ctl, nb);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipMemcpyAsync(&h, d_ctl, sizeof h, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
    const int end = h.k;
    if (h.pw > 0 && end < n) {
      const long long tiles = (n - end + 63) / 64;
      hipLaunchKernelGGL(sytrf_update_kernel, dim3(static_cast<unsigned>(tiles * (tiles + 1) / 2)),
                         dim3(256), 0, s, d_a, lda, n, end, pb, h.pw, w, n);
      if ((e = hipGetLastError()) != hipSuccess) break;
    }
    // The panel left its end in the control block, where the next panel starts.  A panel ends
    // with no column accepted only when nothing active is left (k == active_end).
  }
  if (e == hipSuccess) {
    if ((e = hipMemcpyAsync(perm, d_perm, sizeof(int) * nn, hipMemcpyDeviceToHost, s)) ==
            hipSuccess &&
        (e = hipMemcpyAsync(piv_size, d_piv, sizeof(int) * nn, hipMemcpyDeviceToHost, s)) ==
            hipSuccess)
      e = hipStreamSynchronize(s);
  }
  const hipError_t ef = hipFreeAsync(mem, s);
  if (e == hipSuccess) e = ef;
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return to_status(e);
  *info = h.active_end;
  return DLU_SUCCESS;
}

}  // namespace
}  // namespace dlu

extern "C" {

dlu_status dlu_sytrf_device(int n, double *d_a, int lda, int *perm, int *piv_size, int *info,
                            void *stream) {
  return dlu::sytrf_device(n, d_a, lda, perm, piv_size, info, dlu::kSyMaxNb,
                           static_cast<hipStream_t>(stream));
}

dlu_status dlu_sytrf_device_nb(int n, double *d_a, int lda, int *perm, int *piv_size, int *info,
                               int nb, void *stream) {
  return dlu::sytrf_device(n, d_a, lda, perm, piv_size, info, nb,
                           static_cast<hipStream_t>(stream));
}

dlu_status dlu_sytrf_host(int n, double *a, int lda, int *perm, int *piv_size, int *info,
                          void *stream) {
  if (info == nullptr) return DLU_INVALID_ARGUMENT;
  *info = -1;
  if (n < 0 || lda < std::max(1, n) || (n > 0 && (a == nullptr || perm == nullptr ||
                                                  piv_size == nullptr)))
    return DLU_INVALID_ARGUMENT;
  if (n == 0) {
    *info = 0;
    return DLU_SUCCESS;
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = sizeof(double) * static_cast<size_t>(lda) * n;
  double *d_a = nullptr;
  if (hipMalloc(&d_a, bytes) != hipSuccess) return DLU_ALLOCATION_FAILED;
  dlu_status st = DLU_SUCCESS;
  if (hipMemcpyAsync(d_a, a, bytes, hipMemcpyHostToDevice, s) != hipSuccess) {
    st = DLU_RUNTIME_FAILURE;
  } else {
    st = dlu_sytrf_device(n, d_a, lda, perm, piv_size, info, stream);
    if (st == DLU_SUCCESS && (hipMemcpyAsync(a, d_a, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
                              hipStreamSynchronize(s) != hipSuccess)) {
      *info = -1;
      st = DLU_RUNTIME_FAILURE;
    }
  }
  hipFree(d_a);
  return st;
}

}  // extern "C"
