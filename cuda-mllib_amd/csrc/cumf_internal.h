// cumf_internal.h -- shared helpers of the host side of libGPUALS.so.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

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
