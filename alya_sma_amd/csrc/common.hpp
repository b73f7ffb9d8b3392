// Error reporting shared by the translation units of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

namespace sma_abi {
// Text of the last error on this thread (returned by sma_last_error()).
inline thread_local std::string last_error;
inline int fail(int code, const std::string& msg) {
  last_error = msg;
  return code;
}
}  // namespace sma_abi

#define SMA_HIPCHK(x)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return ::sma_abi::fail(SMA_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct sma_mlp;
// Fused MFMA inference over the step's list of points, segment counters
// on the device (surrogate.hip).
int sma_mlp_infer_listed(sma_mlp* m, const float* X, const unsigned* seg_count, int64_t seg_cap,
                         float* Y, hipStream_t s);
