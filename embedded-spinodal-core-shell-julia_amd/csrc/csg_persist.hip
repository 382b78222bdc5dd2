// Instantiations and launch helpers of the persistent ensemble kernel
// (csg_persist.hpp), FP64 and mixed.
#include <hip/hip_runtime.h>

#include "csg_persist.hpp"

namespace csg {

// Completion counters to zero at the start of a launch (a kernel, so that
// captured graphs hold no memset node).
__global__ void __launch_bounds__(256) persist_zero(unsigned* ctr, int n) {
  const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (i < n) ctr[i] = 0u;
}

bool persist_size_supported(int n) {
  for (int s : PERSIST_SIZES)
    if (s == n) return true;
  return false;
}

template <class T, int N>
void items_of(int items[4]) {
  for (int p = 0; p < 4; ++p) items[p] = persist_items<T, N>(p + 1);
}

void persist_items_of(int n, bool f64, int items[4]) {
  switch (n) {
    case 128: f64 ? items_of<double, 128>(items) : items_of<float, 128>(items); break;
    case 256: f64 ? items_of<double, 256>(items) : items_of<float, 256>(items); break;
    case 384: f64 ? items_of<double, 384>(items) : items_of<float, 384>(items); break;
    default: f64 ? items_of<double, 512>(items) : items_of<float, 512>(items); break;
  }
}

int persist_blocks_per_cu(bool f64, int* blocks) {
  const void* fn = f64 ? reinterpret_cast<const void*>(&persist_kernel<double>)
                       : reinterpret_cast<const void*>(&persist_kernel<float>);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, fn, PERSIST_NT, 0) == hipSuccess ? 0 : 1;
}

hipError_t persist_launch(bool f64, const PersistArgs& a, int G, int nctr, hipStream_t s) {
  hipLaunchKernelGGL(persist_zero, dim3(static_cast<unsigned>((nctr + 255) / 256)), dim3(256), 0, s, a.ctr, nctr);
  if (f64)
    hipLaunchKernelGGL(persist_kernel<double>, dim3(static_cast<unsigned>(G)), dim3(PERSIST_NT), 0, s, a);
  else
    hipLaunchKernelGGL(persist_kernel<float>, dim3(static_cast<unsigned>(G)), dim3(PERSIST_NT), 0, s, a);
  return hipGetLastError();
}

}  // namespace csg
