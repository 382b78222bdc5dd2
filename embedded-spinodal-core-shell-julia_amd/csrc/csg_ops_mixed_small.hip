// Kernel instantiations: mixed precision, N in {16, 32, 64, 128, 256}.
#include "csg_ops.hpp"

namespace csg {

bool ops_mixed_small(int n, Ops* o) {
  switch (n) {
    case 16: *o = make_ops<16, float>(); return true;
    case 32: *o = make_ops<32, float>(); return true;
    case 64: *o = make_ops<64, float>(); return true;
    case 128: *o = make_ops<128, float>(); return true;
    case 256: *o = make_ops<256, float>(); return true;
    default: return false;
  }
}

}  // namespace csg
