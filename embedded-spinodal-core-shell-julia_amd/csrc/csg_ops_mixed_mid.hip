// Kernel instantiations: mixed precision, N in {512, 1024}.
#include "csg_ops.hpp"

namespace csg {

bool ops_mixed_mid(int n, Ops* o) {
  switch (n) {
    case 512: *o = make_ops<512, float>(); return true;
    case 1024: *o = make_ops<1024, float>(); return true;
    default: return false;
  }
}

}  // namespace csg
