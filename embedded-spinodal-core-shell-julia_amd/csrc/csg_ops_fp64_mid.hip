// Kernel instantiations: FP64, N in {512, 1024}.
#include "csg_ops.hpp"

namespace csg {

bool ops_fp64_mid(int n, Ops* o) {
  switch (n) {
    case 512: *o = make_ops<512, double>(); return true;
    case 1024: *o = make_ops<1024, double>(); return true;
    default: return false;
  }
}

}  // namespace csg
