// Wave-level helpers of the panel kernels (gfx950, 64-lane waves): 64-bit halves, cross-lane
// reads, DPP / permlane reductions and the magnitude keys of the pivot search.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dlu {
__device__ __forceinline__ unsigned lo32(double x) {
  return static_cast<unsigned>(static_cast<unsigned long long>(__double_as_longlong(x)));
}
__device__ __forceinline__ unsigned hi32(double x) {
  return static_cast<unsigned>(static_cast<unsigned long long>(__double_as_longlong(x)) >> 32);
}
__device__ __forceinline__ double mk(unsigned lo, unsigned hi) {
  return __longlong_as_double(
      static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Value of x held by lane `src` (wave-uniform result).
__device__ __forceinline__ double lane_value(double x, int src) {
  return mk(static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(lo32(x)), src)),
            static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(hi32(x)), src)));
}
__device__ __forceinline__ float lane_value(float x, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}

// An element travels as Words<T>::n exchange words of 32-bit payload each.
template <typename T>
struct Words;
template <>
struct Words<double> {
  static constexpr int n = 2;
  static __device__ __forceinline__ unsigned part(double x, int i) { return i ? hi32(x) : lo32(x); }
  static __device__ __forceinline__ double join(const unsigned *w) { return mk(w[0], w[1]); }
};
template <>
struct Words<float> {
  static constexpr int n = 1;
  static __device__ __forceinline__ unsigned part(float x, int) { return __float_as_uint(x); }
  static __device__ __forceinline__ float join(const unsigned *w) { return __uint_as_float(w[0]); }
};

template <int CTRL>
__device__ __forceinline__ unsigned dpp(unsigned x) {
  return __builtin_amdgcn_update_dpp(0u, x, CTRL, 0xF, 0xF, false);
}

// Wave-wide max / min: four DPP steps inside each row of 16 lanes, then the gfx950 permlane
// swaps across rows (with both operands equal, the pair returned holds the value of this lane
// and of its partner 16 or 32 lanes away).
__device__ __forceinline__ double wave_max(double x) {
#define DLU_DPP_MAX(CTRL) x = fmax(x, mk(dpp<CTRL>(lo32(x)), dpp<CTRL>(hi32(x))));
  DLU_DPP_MAX(0xB1)   // quad_perm [1,0,3,2]
  DLU_DPP_MAX(0x4E)   // quad_perm [2,3,0,1]
  DLU_DPP_MAX(0x141)  // row_half_mirror
  DLU_DPP_MAX(0x140)  // row_mirror
#undef DLU_DPP_MAX
  {
    const auto a = __builtin_amdgcn_permlane16_swap(lo32(x), lo32(x), false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi32(x), hi32(x), false, false);
    x = fmax(mk(a[0], b[0]), mk(a[1], b[1]));
  }
  {
    const auto a = __builtin_amdgcn_permlane32_swap(lo32(x), lo32(x), false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi32(x), hi32(x), false, false);
    x = fmax(mk(a[0], b[0]), mk(a[1], b[1]));
  }
  return x;
}

__device__ __forceinline__ int wave_min(int x) {
  unsigned u = static_cast<unsigned>(x);
#define DLU_DPP_MIN(CTRL) u = static_cast<unsigned>(min(static_cast<int>(u), static_cast<int>(dpp<CTRL>(u))));
  DLU_DPP_MIN(0xB1)
  DLU_DPP_MIN(0x4E)
  DLU_DPP_MIN(0x141)
  DLU_DPP_MIN(0x140)
#undef DLU_DPP_MIN
  {
    const auto a = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    u = static_cast<unsigned>(min(static_cast<int>(a[0]), static_cast<int>(a[1])));
  }
  {
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    u = static_cast<unsigned>(min(static_cast<int>(a[0]), static_cast<int>(a[1])));
  }
  return static_cast<int>(u);
}

__device__ __forceinline__ unsigned wave_max_u32(unsigned u) {
#define DLU_DPP_UMAX(CTRL) u = max(u, dpp<CTRL>(u));
  DLU_DPP_UMAX(0xB1)
  DLU_DPP_UMAX(0x4E)
  DLU_DPP_UMAX(0x141)
  DLU_DPP_UMAX(0x140)
#undef DLU_DPP_UMAX
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

// Magnitude keys.  For non-negative doubles the IEEE bit pattern, read as an unsigned 64-bit
// integer, orders exactly like the value, so |x| is compared through its bits; NaN and rows that
// may not pivot get key 0, which is also the key of an exact zero (neither can be accepted).
__device__ __forceinline__ uint64_t mag_key(double x, bool eligible) {
  const double a = fabs(x);
  return (eligible && a == a) ? static_cast<uint64_t>(__double_as_longlong(a)) : 0ull;
}
// Float magnitudes widen to double exactly, so both precisions share the 64-bit key.
__device__ __forceinline__ uint64_t mag_key(float x, bool eligible) {
  return mag_key(static_cast<double>(x), eligible);
}
__device__ __forceinline__ double key_value(uint64_t key) {
  return __longlong_as_double(static_cast<long long>(key));
}

// Wave-wide maximum of a 64-bit key: the high halves first, then the low halves of the lanes
// holding the largest high half.
__device__ __forceinline__ uint64_t wave_max_key(uint64_t key) {
  const unsigned hi = wave_max_u32(static_cast<unsigned>(key >> 32));
  const unsigned lo =
      wave_max_u32(static_cast<unsigned>(key >> 32) == hi ? static_cast<unsigned>(key) : 0u);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// The same maximum, skipping the pass over the low halves when a single lane holds the largest
// high half (for magnitudes: unless two candidates agree in sign, exponent and the top 20 bits).
__device__ __forceinline__ uint64_t wave_max_key_fast(uint64_t key) {
  const unsigned kh = static_cast<unsigned>(key >> 32);
  const unsigned hi = wave_max_u32(kh);
  const uint64_t holders = __ballot(kh == hi);
  if (__popcll(holders) == 1) {
    const int src = __ffsll(static_cast<long long>(holders)) - 1;
    return (static_cast<uint64_t>(hi) << 32) |
           static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(key), src));
  }
  const unsigned lo = wave_max_u32(kh == hi ? static_cast<unsigned>(key) : 0u);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace dlu
