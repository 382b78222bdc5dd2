// Kernel instantiations: FP64, N in {16, 32, 64, 128, 256}.
#include "csg_ops.hpp"

namespace csg {

bool ops_fp64_small(int n, Ops* o) {
  switch (n) {
    case 16: *o = make_ops<16, double>(); return true;
    case 32: *o = make_ops<32, double>(); return true;
    case 64: *o = make_ops<64, double>(); return true;
    case 128: *o = make_ops<128, double>(); return true;
    case 256: *o = make_ops<256, double>(); return true;
    default: return false;
  }
}

}  // namespace csg
