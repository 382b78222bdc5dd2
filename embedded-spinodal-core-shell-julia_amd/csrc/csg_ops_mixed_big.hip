// Kernel instantiations: mixed precision, N in {2048, 4096}.
#include "csg_ops.hpp"

namespace csg {

bool ops_mixed_big(int n, Ops* o) {
  switch (n) {
    case 2048: *o = make_ops<2048, float>(); return true;
    case 4096: *o = make_ops<4096, float>(); return true;
    default: return false;
  }
}

}  // namespace csg
