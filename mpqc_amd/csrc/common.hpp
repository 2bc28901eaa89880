// Common definitions for the MI355X (gfx950) CCSD library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace mpqc_hip {

// A failed HIP call becomes a C++ exception; the C-ABI layer turns it into a
// status code plus a message retrievable with mpqc_ccsd_last_error().
struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define MPQC_HIP_CHECK(expr)                                                    \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      throw ::mpqc_hip::HipError(std::string(#expr) + " failed: " +             \
                                 hipGetErrorString(_e) + " at " + __FILE__ +    \
                                 ":" + std::to_string(__LINE__));               \
  } while (0)

inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw HipError(std::string("kernel launch failed (") + what + "): " + hipGetErrorString(e));
}

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

inline long ceil_div(long a, long b) { return (a + b - 1) / b; }

}  // namespace mpqc_hip
