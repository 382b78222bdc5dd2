// Kernel instantiations: mixed precision, N in {48, 96, 192, 384, 768}.
#include "csg_ops.hpp"

namespace csg {

bool ops_mixed_pow3(int n, Ops* o) {
  switch (n) {
    case 48: *o = make_ops<48, float>(); return true;
    case 96: *o = make_ops<96, float>(); return true;
    case 192: *o = make_ops<192, float>(); return true;
    case 384: *o = make_ops<384, float>(); return true;
    case 768: *o = make_ops<768, float>(); return true;
    default: return false;
  }
}

}  // namespace csg
