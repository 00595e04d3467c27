// cumf_internal.h -- shared helpers of the host side of libGPUALS.so.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/als.h"

namespace cumf {

// Thread-local last-error text read by cumf_last_error().
void set_last_error(const std::string& msg);
void clear_last_error();

}  // namespace cumf

// Evaluates a HIP call; on failure records "<what>: <hip error>" and
// returns CUMF_ERR_HIP from the enclosing function.
#define CUMF_HIP_TRY(expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::cumf::set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_));        \
      return CUMF_ERR_HIP;                                                              \
    }                                                                                   \
  } while (0)

namespace cumf {
// Device allocation owned by the host driver.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t n) {
    reset();
    if (n == 0) return CUMF_OK;
    if (hipMalloc(&p, n) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      cumf::set_last_error("hipMalloc of " + std::to_string(n) + " bytes failed");
      return CUMF_ERR_NOMEM;
    }
    bytes = n;
    return CUMF_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};
}  // namespace cumf

#define CUMF_TRY(expr)          \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != CUMF_OK) return rc_; \
  } while (0)
