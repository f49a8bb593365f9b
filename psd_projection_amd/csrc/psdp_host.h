// Host-side error reporting shared by the C ABI translation units of libpsdp.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/psd_projection/psdp.h"

namespace psdp {

// Records a printf-style message as the calling thread's last error
// (psdp_last_error) and returns st.
psdp_status_t fail(psdp_status_t st, const char* fmt, ...);

}  // namespace psdp

#define HIPCHK(call)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return ::psdp::fail(PSDP_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                                   \
  } while (0)

#define STCHK(call)               \
  do {                            \
    psdp_status_t s_ = (call);    \
    if (s_ != PSDP_OK) return s_; \
  } while (0)

// rocBLAS / rocSOLVER status (rocblas_status_success == 0).
#define RBCHK(call)                                                                                  \
  do {                                                                                               \
    const int r_ = (int)(call);                                                                      \
    if (r_ != 0)                                                                                     \
      return ::psdp::fail(PSDP_ERR_DEVICE, "%s failed with status %d (%s:%d)", #call, r_, __FILE__, \
                          __LINE__);                                                                 \
  } while (0)
