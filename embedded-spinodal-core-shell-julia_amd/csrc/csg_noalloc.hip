// csg_step_noalloc: nsteps applications of step (core_shell.jl:72-76) to a
// composition field that already lives in device memory, for callers that
// must not allocate (code under graph capture elsewhere in the program, tight
// serving loops, harnesses that forbid hipMalloc).  It runs the same kernels,
// schedule (candidate 0) and twiddle tables as csg_create + csg_set_fields +
// csg_step + csg_get_field, so the result is the same bit for bit.
//
// Everything the kernels need beyond the caller's c and phi -- the packed
// state, the two bulk work arrays, an FP32 copy of phi (mixed form), the
// per-workgroup partials and the twiddle tables -- lives in `scratch`, one
// file-scope device array of SCRATCH_BYTES per GPU that the process holds
// until it ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).  Calls on one device are
// serialised by a host mutex; calls from other processes use their own copy.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/csg.h"
#include "csg_host.hpp"

using namespace csg;

namespace {

constexpr int NOALLOC_MAX_N = CSG_NOALLOC_MAX_N;
constexpr int MAX_DEVICES = 64;

// Largest twiddle buffer of the supported grids up to NOALLOC_MAX_N.
constexpr int max_twiddle_len() {
  int m = 0;
  for (int n = 16; n <= NOALLOC_MAX_N; n *= 2) m = twiddle_len(n) > m ? twiddle_len(n) : m;
  for (int n = 48; n <= 768 && n <= NOALLOC_MAX_N; n *= 2) m = twiddle_len(n) > m ? twiddle_len(n) : m;
  return m;
}
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr int TW_MAX = max_twiddle_len();
constexpr size_t PART_MAX = 64 * 1024;   // doubles of per-workgroup partials
constexpr size_t NN_MAX = static_cast<size_t>(NOALLOC_MAX_N) * NOALLOC_MAX_N;

// Layout (byte offsets): twiddles (FP64, FP32), partials, state S, bulk
// arrays A and Bb, FP32 phi.  The FP64 form uses 8-byte bulk arrays and the
// caller's phi; the mixed form 4-byte bulk arrays and the FP32 copy.
constexpr size_t OFF_TW64 = 0;
constexpr size_t OFF_TW32 = OFF_TW64 + align256(TW_MAX * sizeof(cplx<double>));
constexpr size_t OFF_PART = OFF_TW32 + align256(TW_MAX * sizeof(cplx<float>));
constexpr size_t OFF_S = OFF_PART + align256(PART_MAX * sizeof(double));
constexpr size_t OFF_A = OFF_S + NN_MAX * sizeof(double);
constexpr size_t OFF_B = OFF_A + NN_MAX * sizeof(double);
constexpr size_t SCRATCH_BYTES = OFF_B + NN_MAX * sizeof(double);
// Mixed layout: A and Bb are FP32, so the FP32 phi follows them inside the
// FP64 footprint.
constexpr size_t mixed_off_b(size_t nn) { return OFF_A + align256(nn * sizeof(float)); }
constexpr size_t mixed_off_phi(size_t nn) { return mixed_off_b(nn) + align256(nn * sizeof(float)); }
static_assert(mixed_off_phi(NN_MAX) + align256(NN_MAX * sizeof(float)) <= SCRATCH_BYTES, "mixed layout");
static_assert(SCRATCH_BYTES == CSG_NOALLOC_SCRATCH_BYTES, "csg.h states the scratch size");

}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

struct DeviceSlot {
  std::mutex mu;
  unsigned char* base = nullptr;   // device address of `scratch` on this device
  int tw_n = 0;                    // grid whose twiddles the scratch holds (0: none)
};
DeviceSlot g_slots[MAX_DEVICES];

// Host staging of the twiddle tables (static: nothing is allocated).
cplx<double> g_tw64[TW_MAX];
cplx<float> g_tw32[TW_MAX];

__global__ void __launch_bounds__(256) phi_to_float(const double* __restrict__ in, float* __restrict__ out,
                                                    long long nn) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < nn) out[i] = static_cast<float>(in[i]);
}

int supported(int n) {
  Ops probe{};
  return ops_for<double>(n, &probe) ? CSG_OK : fail(CSG_ERR_UNSUPPORTED, unsupported_size_message(n));
}

}  // namespace

extern "C" {

int csg_step_noalloc(int32_t n, double spacing, const csg_params* p, int32_t precision, int32_t device, double dt,
                     int64_t nsteps, double* c_dev, const double* phi_dev, void* stream) {
  if (!c_dev || !phi_dev) return fail(CSG_ERR_ARG, "c_dev and phi_dev must be non-NULL");
  if (!(spacing > 0.0) || !std::isfinite(spacing)) return fail(CSG_ERR_ARG, "spacing must be > 0");
  if (precision != CSG_FP64 && precision != CSG_MIXED) return fail(CSG_ERR_ARG, "unknown precision");
  if (nsteps < 0 || !std::isfinite(dt)) return fail(CSG_ERR_ARG, "bad nsteps/dt");
  if (device < 0 || device >= MAX_DEVICES) return fail(CSG_ERR_ARG, "device ordinal out of range");
  if (int rc = supported(n)) return rc;
  if (n > NOALLOC_MAX_N)
    return fail(CSG_ERR_ARG, "grid size " + std::to_string(n) + " exceeds the allocation-free maximum N = " +
                                 std::to_string(NOALLOC_MAX_N) + " (scratch of " + std::to_string(SCRATCH_BYTES) +
                                 " bytes); use csg_create");
  Ops ops{};
  if (precision == CSG_FP64) ops_for<double>(n, &ops); else ops_for<float>(n, &ops);
  const size_t npart = 5 * static_cast<size_t>(ops.max_row1) + ops.max_col + ops.max_row3;
  if (npart > PART_MAX) return fail(CSG_ERR_ARG, "partials exceed the allocation-free scratch");
  csg_params pp;
  if (p) pp = *p; else csg_default_params(&pp);
  const DevParams dp{pp.A, pp.B, pp.chi, pp.P, pp.ca, pp.cb, pp.cc, pp.kappa, pp.mobility, pp.Q,
                     spacing, (1.0 / spacing) / n};
  hipStream_t s = static_cast<hipStream_t>(stream);

  DeviceSlot& slot = g_slots[device];
  std::lock_guard<std::mutex> lock(slot.mu);
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  if (!slot.base) {
    // once per device, with that device current: this GPU's copy of scratch
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(CSG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    slot.base = static_cast<unsigned char*>(ptr);
  }
  unsigned char* base = slot.base;
  const size_t nn = static_cast<size_t>(n) * n;
  const int twl = twiddle_len(n);
  if (slot.tw_n != n) {
    // the twiddles of this grid (both precisions), copied from static host
    // staging; the copies finish before the staging is written again
    fill_twiddles<double>(n, g_tw64);
    fill_twiddles<float>(n, g_tw32);
    slot.tw_n = 0;
    e = hipMemcpyAsync(base + OFF_TW64, g_tw64, twl * sizeof(cplx<double>), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(base + OFF_TW32, g_tw32, twl * sizeof(cplx<float>), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("twiddle upload: ") + hipGetErrorString(e));
    slot.tw_n = n;
  }
  void* tw64 = base + OFF_TW64;
  void* tw = precision == CSG_FP64 ? tw64 : static_cast<void*>(base + OFF_TW32);
  double* part = reinterpret_cast<double*>(base + OFF_PART);
  void* S = base + OFF_S;
  void* A = base + OFF_A;
  void* Bb = base + (precision == CSG_FP64 ? OFF_B : mixed_off_b(nn));
  const void* phi = phi_dev;
  // cache hints as a context of this size would choose them
  const size_t tsize = precision == CSG_FP64 ? sizeof(double) : sizeof(float);
  const int nt = nontemporal_for(nn * (2 * sizeof(double) + 3 * tsize), false);
  if (precision == CSG_MIXED) {
    float* phi32 = reinterpret_cast<float*>(base + mixed_off_phi(nn));
    hipLaunchKernelGGL(phi_to_float, dim3(static_cast<unsigned>((nn + 255) / 256)), dim3(256), 0, s, phi_dev, phi32,
                       static_cast<long long>(nn));
    phi = phi32;
  }
  auto args = [&]() {
    PassArgs a{};
    a.p = dp;
    a.phi = phi;
    a.nt = nt;
    return a;
  };
  double* part1 = part;
  double* part2 = part + 5 * ops.n_row1;
  double* part3 = part2 + ops.n_col;
  // c -> packed state (as csg_set_fields)
  PassArgs a = args();
  a.in0 = c_dev;
  a.out0 = S;
  ops.fwd(a, tw64, s);
  for (int64_t k = 0; k < nsteps; ++k) {
    a = args();
    a.in0 = S; a.out0 = A; a.part = part1;
    ops.p1(a, tw, s);
    a = args();
    a.in0 = S; a.in1 = A; a.out0 = A; a.out1 = Bb; a.part = part2;
    ops.p2(a, tw, s);
    a = args();
    a.in0 = A; a.in1 = Bb; a.out0 = A; a.out1 = Bb; a.part = part3;
    ops.p3(a, tw, s);
    a = args();
    a.in0 = A; a.in1 = Bb; a.out0 = S; a.dt = dt;   // no history row: fin.hist stays NULL
    ops.p4(a, tw, s);
  }
  // packed state -> c (as csg_get_field)
  a = args();
  a.in0 = S;
  a.out0 = c_dev;
  ops.inv_state(a, tw64, s);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("csg_step_noalloc: ") + hipGetErrorString(e));
  return CSG_OK;
}

}  // extern "C"
