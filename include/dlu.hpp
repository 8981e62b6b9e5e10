// dlu.hpp -- C++ convenience layer over the C ABI in dlu.h, shaped like the reference's
// delayed_lu_cuda/delayed_dgetrf_cuda.hpp:7-63 (int& info, stream defaulting to the null stream,
// native_* and hybrid_* spellings).  Header-only; link against libdlu.so.
#pragma once

#include <hip/hip_runtime_api.h>

#include "dlu.h"

namespace dlu {

// delayed_lu_cuda/delayed_dgetrf_cuda.hpp:7-13, same ordinals.
enum class Status {
  success = DLU_SUCCESS,
  invalid_argument = DLU_INVALID_ARGUMENT,
  allocation_failed = DLU_ALLOCATION_FAILED,
  runtime_failure = DLU_RUNTIME_FAILURE,
  kernel_failure = DLU_KERNEL_FAILURE,
};

// delayed_dgetrf_cuda.hpp:24-26
inline Status delayed_getrf_device(int m, int n, double *d_a, int lda, int *ipiv, int *jpiv,
                                   int &info, hipStream_t stream = nullptr) {
  return static_cast<Status>(dlu_getrf_device(m, n, d_a, lda, ipiv, jpiv, &info, stream));
}

// delayed_dgetrf_cuda.hpp:35-37
inline Status delayed_getrf_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                 int &info, hipStream_t stream = nullptr) {
  return static_cast<Status>(dlu_getrf_host(m, n, a, lda, ipiv, jpiv, &info, stream));
}

// delayed_dgetrf_cuda.hpp:40-46
inline Status delayed_getrf_native_device(int m, int n, double *d_a, int lda, int *ipiv,
                                          int *jpiv, int &info, hipStream_t stream = nullptr) {
  return delayed_getrf_device(m, n, d_a, lda, ipiv, jpiv, info, stream);
}
inline Status delayed_getrf_native_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                        int &info, hipStream_t stream = nullptr) {
  return delayed_getrf_host(m, n, a, lda, ipiv, jpiv, info, stream);
}

// delayed_dgetrf_cuda.hpp:55-61 (the panel always runs on the GPU here)
inline Status delayed_getrf_hybrid_device(int m, int n, double *d_a, int lda, int *ipiv,
                                          int *jpiv, int &info, hipStream_t stream = nullptr) {
  return delayed_getrf_device(m, n, d_a, lda, ipiv, jpiv, info, stream);
}
inline Status delayed_getrf_hybrid_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                        int &info, hipStream_t stream = nullptr) {
  return delayed_getrf_host(m, n, a, lda, ipiv, jpiv, info, stream);
}

// delayed_lu/delayed_dgetrf_column.hpp:19 (the column-delay variant, on the GPU)
inline Status delayed_getrf_column_device(int m, int n, double *d_a, int lda, int *ipiv,
                                          int *jpiv, int &info, hipStream_t stream = nullptr) {
  return static_cast<Status>(dlu_getrf_column_device(m, n, d_a, lda, ipiv, jpiv, &info, stream));
}
inline Status delayed_getrf_column_host(int m, int n, double *a, int lda, int *ipiv, int *jpiv,
                                        int &info, hipStream_t stream = nullptr) {
  return static_cast<Status>(dlu_getrf_column_host(m, n, a, lda, ipiv, jpiv, &info, stream));
}

// delayed-bunch-kaufman/delayed_sytrf_blocked.hpp:16 (the delayed Bunch-Kaufman LDL^T, on the
// GPU)
inline Status delayed_sytrf_blocked_device(int n, double *d_a, int lda, int *perm, int *piv_size,
                                           int &info, hipStream_t stream = nullptr) {
  return static_cast<Status>(dlu_sytrf_device(n, d_a, lda, perm, piv_size, &info, stream));
}
inline Status delayed_sytrf_blocked_host(int n, double *a, int lda, int *perm, int *piv_size,
                                         int &info, hipStream_t stream = nullptr) {
  return static_cast<Status>(dlu_sytrf_host(n, a, lda, perm, piv_size, &info, stream));
}

// delayed_dgetrf_cuda.hpp:63
inline const char *status_string(Status s) {
  return dlu_status_string(static_cast<dlu_status>(s));
}

}  // namespace dlu
