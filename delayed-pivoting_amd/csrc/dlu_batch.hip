// Batched delayed-pivoting LU and solve for small matrices: one workgroup per matrix, the matrix
// resident in LDS for the whole factorization and solve.
//
// Reference behaviour restated per matrix (delayed_lu/delayed_dgetrf.cpp:129-251, unblocked: the
// reference's panel materialization makes its blocked order equivalent):
//   column pivot, largest |a| and the lowest row on ties              :170-172
//   a zero (or NaN-only) column falls back to the largest |a| of the whole active block, lowest
//   column-major offset on ties, row swap then column swap            :32-47, 177-209
//   an all-zero active block ends the elimination (info = k)          :189-192
//   multipliers by the reciprocal of the pivot, rank-1 update         :221-239
// followed, when info == n, by the solve x = Q U^{-1} L^{-1} P b.
#include "dlu_internal.hpp"

namespace dlu {
namespace {

template <int CTRL>
__device__ __forceinline__ unsigned bdpp(unsigned x) {
  return __builtin_amdgcn_update_dpp(0u, x, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ unsigned wave_umax(unsigned u) {
  u = max(u, bdpp<0xB1>(u));
  u = max(u, bdpp<0x4E>(u));
  u = max(u, bdpp<0x141>(u));
  u = max(u, bdpp<0x140>(u));
  {
    const auto a = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    u = max(a[0], a[1]);
  }
  {
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    u = max(a[0], a[1]);
  }
  return u;
}
__device__ __forceinline__ unsigned wave_umin(unsigned u) {
  u = min(u, bdpp<0xB1>(u));
  u = min(u, bdpp<0x4E>(u));
  u = min(u, bdpp<0x141>(u));
  u = min(u, bdpp<0x140>(u));
  {
    const auto a = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    u = min(a[0], a[1]);
  }
  {
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    u = min(a[0], a[1]);
  }
  return u;
}
// Largest 64-bit key over the wave (high halves first).
__device__ __forceinline__ uint64_t wave_kmax(uint64_t k) {
  const unsigned hi = wave_umax(static_cast<unsigned>(k >> 32));
  const unsigned lo = wave_umax(static_cast<unsigned>(k >> 32) == hi ? static_cast<unsigned>(k) : 0u);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
// |x| as an ordered key; NaN is never a pivot (key 0, like an exact zero).
__device__ __forceinline__ uint64_t bkey(double x) {
  const double a = fabs(x);
  return a == a ? static_cast<uint64_t>(__double_as_longlong(a)) : 0ull;
}

template <int NMAX>
__global__ __launch_bounds__(256) void batch_gesv_kernel(const BatchItem *__restrict__ items) {
  constexpr int LS = NMAX + 1;  // LDS column stride
  __shared__ double As[NMAX * LS];
  __shared__ double Bs[NMAX * kBatchMaxRhs];
  __shared__ int s_piv[NMAX], s_qiv[NMAX];
  __shared__ uint64_t s_key[4];
  __shared__ unsigned s_off[4];
  __shared__ int s_p, s_q, s_stop;
  const BatchItem it = items[blockIdx.x];
  const int n = it.n, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  auto A = [&](int i, int j) -> double & { return As[j * LS + i]; };

  for (int idx = t; idx < n * n; idx += 256) {
    const int i = idx % n, j = idx / n;
    A(i, j) = it.a[static_cast<size_t>(i) + static_cast<size_t>(j) * it.lda];
  }
  int info = n;
  if (it.solve_only) {
    for (int k = t; k < n; k += 256) {
      s_piv[k] = it.ipiv[k];
      s_qiv[k] = it.jpiv[k];
    }
    info = it.info_in;
  } else {
    if (t == 0) s_stop = n;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
      // Column pivot of column k (wave 0): largest |a|, lowest row on ties.
      if (wave == 0) {
        uint64_t best = 0;
        int brow = 0x7fffffff;
        for (int i = k + lane; i < n; i += 64) {
          const uint64_t key = bkey(A(i, k));
          if (key > best) {
            best = key;
            brow = i;
          }
        }
        const uint64_t g = wave_kmax(best);
        const int p = static_cast<int>(wave_umin(g != 0 && best == g ? static_cast<unsigned>(brow)
                                                                    : 0x7fffffffu));
        if (lane == 0) {
          s_p = g != 0 ? p : -1;
          s_q = k;
        }
      }
      __syncthreads();
      if (s_p < 0) {
        // Zero column: the largest |a| of the active block, lowest column-major offset on ties.
        const int rows = n - k, cnt = rows * (n - k);
        uint64_t best = 0;
        unsigned boff = 0xffffffffu;
        for (int o = t; o < cnt; o += 256) {
          const uint64_t key = bkey(A(k + o % rows, k + o / rows));
          if (key > best) {  // offsets rise with o, so the first maximum seen is the lowest
            best = key;
            boff = static_cast<unsigned>(o);
          }
        }
        const uint64_t g = wave_kmax(best);
        const unsigned off = wave_umin(g != 0 && best == g ? boff : 0xffffffffu);
        if (lane == 0) {
          s_key[wave] = g;
          s_off[wave] = off;
        }
        __syncthreads();
        if (t == 0) {
          uint64_t gb = 0;
          unsigned go = 0xffffffffu;
          for (int w = 0; w < 4; ++w)
            if (s_key[w] > gb || (s_key[w] == gb && s_off[w] < go)) {
              gb = s_key[w];
              go = s_off[w];
            }
          if (gb == 0) {
            s_stop = k;  // the whole active block is zero: defer the rest
            s_p = -1;
          } else {
            s_p = k + static_cast<int>(go % static_cast<unsigned>(rows));
            s_q = k + static_cast<int>(go / static_cast<unsigned>(rows));
          }
        }
        __syncthreads();
        if (s_stop == k) break;
      }
      const int p = s_p, q = s_q;
      if (t == 0) {
        s_piv[k] = p;
        s_qiv[k] = q;
      }
      // Row swap over all n columns, then column swap over all n rows.
      if (p != k)
        for (int j = t; j < n; j += 256) {
          const double x = A(k, j);
          A(k, j) = A(p, j);
          A(p, j) = x;
        }
      __syncthreads();
      if (q != k)
        for (int i = t; i < n; i += 256) {
          const double x = A(i, k);
          A(i, k) = A(i, q);
          A(i, q) = x;
        }
      __syncthreads();
      const double recip = 1.0 / A(k, k);
      for (int i = k + 1 + t; i < n; i += 256) A(i, k) *= recip;
      __syncthreads();
      const int m = n - k - 1;
      for (int idx = t; idx < m * m; idx += 256) {
        const int i = k + 1 + idx % m, j = k + 1 + idx / m;
        A(i, j) = fma(-A(i, k), A(k, j), A(i, j));
      }
      __syncthreads();
    }
    __syncthreads();
    info = s_stop;
    for (int k = t; k < n; k += 256) {
      if (k >= info) {
        s_piv[k] = k;
        s_qiv[k] = k;
      }
    }
    __syncthreads();
    for (int k = t; k < n; k += 256) {
      it.ipiv[k] = s_piv[k];
      it.jpiv[k] = s_qiv[k];
    }
    for (int idx = t; idx < n * n; idx += 256) {
      const int i = idx % n, j = idx / n;
      it.lu[static_cast<size_t>(i) + static_cast<size_t>(j) * it.ldlu] = A(i, j);
    }
    if (t == 0) it.info[0] = info;
  }
  __syncthreads();
  if (info != n || it.nrhs <= 0) return;

  // Solve: X = Q U^{-1} L^{-1} P B, one thread per (row, right-hand side) where parallel.
  const int nr = it.nrhs;
  auto Bv = [&](int i, int j) -> double & { return Bs[j * NMAX + i]; };
  for (int idx = t; idx < n * nr; idx += 256) {
    const int i = idx % n, j = idx / n;
    Bv(i, j) = it.b[static_cast<size_t>(i) + static_cast<size_t>(j) * it.ldb];
  }
  __syncthreads();
  if (t < nr)
    for (int k = 0; k < n; ++k) {
      const int p = s_piv[k];
      const double x = Bv(k, t);
      Bv(k, t) = Bv(p, t);
      Bv(p, t) = x;
    }
  __syncthreads();
  for (int k = 0; k < n; ++k) {  // unit lower
    for (int idx = t; idx < (n - k - 1) * nr; idx += 256) {
      const int i = k + 1 + idx % (n - k - 1), j = idx / (n - k - 1);
      Bv(i, j) = fma(-A(i, k), Bv(k, j), Bv(i, j));
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {  // upper
    if (t < nr) Bv(k, t) = Bv(k, t) / A(k, k);
    __syncthreads();
    for (int idx = t; idx < k * nr; idx += 256) {
      const int i = idx % k, j = idx / k;
      Bv(i, j) = fma(-A(i, k), Bv(k, j), Bv(i, j));
    }
    __syncthreads();
  }
  if (t < nr)
    for (int k = n - 1; k >= 0; --k) {
      const int q = s_qiv[k];
      const double x = Bv(k, t);
      Bv(k, t) = Bv(q, t);
      Bv(q, t) = x;
    }
  __syncthreads();
  for (int idx = t; idx < n * nr; idx += 256) {
    const int i = idx % n, j = idx / n;
    it.x[static_cast<size_t>(i) + static_cast<size_t>(j) * it.ldx] = Bv(i, j);
  }
}

// Deterministic values in [-1, 1) from a counter hash (splitmix64): inputs for the cutoff tuner.
__global__ void fill_uniform_kernel(double *a, size_t count, unsigned long long seed) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += gridDim.x * 256ull) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    a[i] = static_cast<double>(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

}  // namespace

hipError_t launch_fill_uniform(double *a, size_t count, unsigned long long seed, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const size_t blocks = (count + 255) / 256 < 4096 ? (count + 255) / 256 : 4096;
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                     a, count, seed);
  return hipGetLastError();
}

hipError_t launch_batch_gesv(const BatchItem *d_items, int count, int nmax, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (nmax <= 64)
    hipLaunchKernelGGL(batch_gesv_kernel<64>, dim3(count), dim3(256), 0, s, d_items);
  else if (nmax <= 128)
    hipLaunchKernelGGL(batch_gesv_kernel<128>, dim3(count), dim3(256), 0, s, d_items);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace dlu
