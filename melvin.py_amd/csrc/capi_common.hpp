// Helpers shared by the extern "C" translation units of libmelvin_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace mlv {
namespace abi {

// Record a failure message (read back by mlv_last_error) and return code.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// log2 of n when n is a power of two, else -1.
int ilog2_exact(int n);
// FP64-tabulated twiddle buffer of the device, in the precision of dtype.
int get_twiddles(int device, int dtype, const void** out);
size_t max_lds_per_block(int device);
size_t cplx_bytes(int dtype);
// Copy up to max_records tracker records, oldest first, from the device
// diagnostics header at hdr (diag.hpp layout) to out; the device must be
// idle.  *nrec receives the count.
int copy_tracker(const double* hdr, double* out, int max_records, int* nrec);

}  // namespace abi
}  // namespace mlv

#define HIP_TRY(call, what)                                        \
  do {                                                             \
    hipError_t e_ = (call);                                        \
    if (e_ != hipSuccess) return ::mlv::abi::hip_fail(e_, what);   \
  } while (0)
