// Host-side helpers shared by the C ABI units: error reporting, the kernel
// table lookup and the loop constants.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

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

// Twiddle buffer of an N-point grid (layout: stab_offset), written to
// w[0 .. twiddle_len(n)): W_N^m, m in [0, N), from extended-precision sines
// and cosines rounded once, then each stage-table entry taken from the first
// half of that table (negated past the half circle), so both twiddle paths
// use the same values.  Allocates nothing.
template <class T>
void fill_twiddles(int n, cplx<T>* w) {
  for (int m = 0; m < n; ++m) {
    const long double ang = -2.0L * 3.141592653589793238462643383279502884L * m / n;
    w[m] = cplx<T>{static_cast<T>(cosl(ang)), static_cast<T>(sinl(ang))};
  }
  const int h = n / 2;
  for (int hh = 0; hh < 2; ++hh) {
    const int len = hh == 0 ? n : n / 2;
    for (int i = 0; i < stab_e_count; ++i) {
      const int e = stab_e_at(i);
      if (!stab_e_ok(len, e)) continue;
      const int off = stab_offset(n, len, e), q1 = len - 1;
      for (int q = 0; q < stab_len(len, e); ++q) {
        const int x = (q < q1 ? stab_exp_rt(len, e, false, q) : stab_exp_rt(len, e, true, q - q1)) * (n / len);
        const cplx<T> v = w[x < h ? x : x - h];
        w[off + q] = x < h ? v : cplx<T>{-v.x, -v.y};
      }
    }
  }
}

template <class T>
std::vector<cplx<T>> make_twiddles(int n) {
  std::vector<cplx<T>> w(twiddle_len(n));
  fill_twiddles<T>(n, w.data());
  return w;
}

// Cache hints of the step passes (NT_STORE / NT_LOAD / NT_PHI) by device
// footprint, measured in the captured loop.
// - Non-temporal pass outputs pay while a step's working set stays well
//   inside the 256 MB Infinity Cache and cost beyond it: 1024^2 FP64 40 MiB
//   -5 %, mixed 28 MiB -6 %; 2048^2 mixed 113 MiB +6 %, FP64 161 MiB +22 %;
//   4096^2 and the 256-particle ensemble, 450-950 MiB, +3 to +19 %.
// - Non-temporal column-pass loads are the reverse: 4096^2 mixed -6 %, FP64
//   and ensembles -1 %, 1024^2 mixed +2 %.
// - Non-temporal phi loads pay for one field far beyond the cache (4096^2
//   FP64, 641 MiB: -7 %) and cost where part of phi stays cached (4096^2
//   mixed, 449 MiB: +3 %; ensembles: +1 to +3 %).
// So: footprint <= NT_MAX_BYTES streams the stores, larger the column loads,
// and a single field above NT_PHI_MIN_BYTES also phi (a bit that only the FP64
// row passes of N >= PHI_HINT_MIN_N read).  CSG_NONTEMPORAL=0..7 (the hint
// bits) overrides.
constexpr size_t NT_MAX_BYTES = size_t(64) << 20;
constexpr size_t NT_PHI_MIN_BYTES = size_t(512) << 20;
inline int nontemporal_for(size_t bytes, bool ensemble) {
  if (const char* e = std::getenv("CSG_NONTEMPORAL"))
    if (e[0] >= '0' && e[0] <= '7' && e[1] == '\0') return e[0] - '0';
  int h = bytes <= NT_MAX_BYTES ? NT_STORE : NT_LOAD;
  if (!ensemble && bytes > NT_PHI_MIN_BYTES) h |= NT_PHI;
  return h;
}

constexpr int FIN_NT = 256;           // threads of the standalone history reduction
constexpr int STEPS_PER_GRAPH = 32;   // loop iterations per captured graph
constexpr int CHECK_GRAPHS = 64;      // graph replays between non-finite checks of a long run

}  // namespace csg
