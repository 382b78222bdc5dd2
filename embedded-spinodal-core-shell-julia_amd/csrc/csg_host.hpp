// Host-side helpers shared by the C ABI units: error reporting, the kernel
// table lookup and the loop constants.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>

#include "csg_ops.hpp"

namespace csg {

// Set the calling thread's last error message (csg_last_error) and return code.
int fail(int code, const std::string& msg);

#define CSG_HIPCHK(expr)                                                                 \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::csg::fail(CSG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
bool ops_for(int n, Ops* o) {
  if constexpr (std::is_same<T, double>::value)
    return ops_fp64_small(n, o) || ops_fp64_mid(n, o) || ops_fp64_big(n, o) || ops_fp64_pow3(n, o);
  else
    return ops_mixed_small(n, o) || ops_mixed_mid(n, o) || ops_mixed_big(n, o) || ops_mixed_pow3(n, o);
}

inline std::string unsupported_size_message(int n) {
  return "grid size " + std::to_string(n) +
         " not supported (N = 2^m with 16 <= N <= 4096, or N = 3*2^m with 48 <= N <= 768)";
}

constexpr int FIN_NT = 256;           // threads of the standalone history reduction
constexpr int STEPS_PER_GRAPH = 32;   // loop iterations per captured graph

}  // namespace csg
