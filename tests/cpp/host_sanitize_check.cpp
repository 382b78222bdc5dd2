// Host-side check of the C ABI under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/test_sanitize.py, CPU only):
// the Julia-exact RNG and initial_fields over edge sizes, and every entry
// point's argument rejection and error-cleanup paths that run before or
// without a device.  Outputs of the RNG and initial_fields go to the file
// named by argv[1] (raw float64) for the test to compare with the oracle.
// With a HIP device present, the cases that would reach the device with
// placeholder pointers are skipped.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/csg.h"

static int failures = 0;
#define EXPECT(cond)                                                    \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::printf("FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond, csg_last_error()); \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* out = std::fopen(argv[1], "wb");
  if (!out) return 2;

  // ---- host RNG and initial fields ----
  const long long seeds[] = {5749, 0, 1, 123456789012LL};
  const long long counts[] = {1, 2, 7, 382, 383, 384, 385, 1000, 1001, 1002, 1003, 4096};
  for (long long s : seeds)
    for (long long c : counts) {
      std::vector<double> a(c), b(c);
      EXPECT(csg_julia_rand(s, c, a.data()) == CSG_OK);
      EXPECT(csg_julia_rand(s, c, b.data()) == CSG_OK);
      EXPECT(std::memcmp(a.data(), b.data(), c * sizeof(double)) == 0);
      std::fwrite(a.data(), sizeof(double), c, out);
    }
  EXPECT(csg_julia_rand(5749, 0, nullptr) == CSG_OK || csg_julia_rand(5749, 0, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_julia_rand(5749, -1, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_julia_rand(5749, 4, nullptr) == CSG_ERR_ARG);
  const int sizes[] = {1, 2, 3, 16, 17, 64};
  for (int n : sizes) {
    std::vector<double> c(static_cast<size_t>(n) * n), phi(static_cast<size_t>(n) * n);
    EXPECT(csg_initial_fields(n, 1.0, 0.3 * n, 2.0, 5749, c.data(), phi.data()) == CSG_OK);
    std::fwrite(c.data(), sizeof(double), c.size(), out);
    std::fwrite(phi.data(), sizeof(double), phi.size(), out);
  }
  EXPECT(csg_initial_fields(0, 1.0, 1.0, 1.0, 1, nullptr, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_initial_fields(8, 1.0, 1.0, 1.0, 1, nullptr, nullptr) == CSG_ERR_ARG);
  std::fclose(out);

  // ---- argument rejection (no device touched) ----
  csg_params p;
  csg_default_params(&p);
  csg_default_params(nullptr);
  EXPECT(p.A == 2.0 && p.Q == 0.5);
  csg_ctx* ctx = reinterpret_cast<csg_ctx*>(0x1);
  EXPECT(csg_create(nullptr, 64, 1.0, &p, 0, 0) == CSG_ERR_ARG);
  EXPECT(csg_create(&ctx, 64, -1.0, &p, 0, 0) == CSG_ERR_ARG && ctx == nullptr);
  EXPECT(csg_create(&ctx, 64, NAN, &p, 0, 0) == CSG_ERR_ARG);
  EXPECT(csg_create(&ctx, 64, 1.0, &p, 7, 0) == CSG_ERR_ARG);
  for (int n : {0, -4, 8, 24, 100, 1536, 8192}) EXPECT(csg_create(&ctx, n, 1.0, &p, 0, 0) == CSG_ERR_UNSUPPORTED);
  EXPECT(std::strstr(csg_last_error(), "not supported") != nullptr);
  csg_batch* b = nullptr;
  const int32_t ns_bad[] = {64, 100};
  const int32_t ns_ok[] = {64, 128};
  const double sp_bad[] = {1.0, -2.0};
  EXPECT(csg_batch_create(nullptr, 2, ns_ok, nullptr, nullptr, 0, 0) == CSG_ERR_ARG);
  EXPECT(csg_batch_create(&b, 0, ns_ok, nullptr, nullptr, 0, 0) == CSG_ERR_ARG);
  EXPECT(csg_batch_create(&b, 2, nullptr, nullptr, nullptr, 0, 0) == CSG_ERR_ARG);
  EXPECT(csg_batch_create(&b, 2, ns_bad, nullptr, nullptr, 0, 0) == CSG_ERR_UNSUPPORTED);
  EXPECT(csg_batch_create(&b, 2, ns_ok, sp_bad, nullptr, 0, 0) == CSG_ERR_ARG);
  EXPECT(csg_batch_create(&b, 2, ns_ok, nullptr, nullptr, 3, 0) == CSG_ERR_ARG);
  // null handles
  double x = 0.0;
  int32_t sel[4] = {0, 0, 0, 0}, cnt = 0;
  EXPECT(csg_destroy(nullptr) == CSG_OK);
  EXPECT(csg_batch_destroy(nullptr) == CSG_OK);
  EXPECT(csg_set_fields(nullptr, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_run(nullptr, 0.01, 1, 1, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_step(nullptr, 0.01, 1) == CSG_ERR_ARG);
  EXPECT(csg_get_field(nullptr, &x) == CSG_ERR_ARG);
  EXPECT(csg_nonfinite_step(nullptr) == -1);
  EXPECT(csg_energy(nullptr, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_chemical_potential(nullptr, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_transport_rate(nullptr, &x, &x, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_geometry(nullptr, &x, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_bulk_energy(nullptr, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_bulk_derivative(nullptr, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_time_passes(nullptr, 0.01, 1, &x) == CSG_ERR_ARG);
  EXPECT(csg_device_bytes(nullptr) == 0);
  EXPECT(csg_set_params(nullptr, &p) == CSG_ERR_ARG);
  EXPECT(csg_adjoint_forward(nullptr, 0.01, 1, 1) == CSG_ERR_ARG);
  EXPECT(csg_adjoint_backward(nullptr, &x, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_schedule_info(nullptr, 0, &cnt, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_get_schedule(nullptr, sel) == CSG_ERR_ARG);
  EXPECT(csg_set_schedule(nullptr, sel) == CSG_ERR_ARG);
  EXPECT(csg_autotune(nullptr, 0.01, 1, sel, &x) == CSG_ERR_ARG);
  EXPECT(csg_batch_set_fields(nullptr, 0, &x, &x) == CSG_ERR_ARG);
  EXPECT(csg_batch_run(nullptr, 0.01, 1, 1, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_batch_get_field(nullptr, 0, &x) == CSG_ERR_ARG);
  EXPECT(csg_batch_nonfinite_step(nullptr, 0) == -1);
  EXPECT(csg_batch_device_bytes(nullptr) == 0);
  EXPECT(csg_batch_get_schedule(nullptr, 64, sel) == CSG_ERR_ARG);
  EXPECT(csg_batch_set_schedule(nullptr, 64, sel) == CSG_ERR_ARG);
  EXPECT(csg_batch_autotune(nullptr, 0.01, 1) == CSG_ERR_ARG);
  EXPECT(csg_copy_bandwidth(0, 15, 1, &x) == CSG_ERR_ARG);
  EXPECT(csg_copy_bandwidth(0, 1 << 20, 0, &x) == CSG_ERR_ARG);
  EXPECT(csg_copy_bandwidth(0, 1 << 20, 1, nullptr) == CSG_ERR_ARG);
  void* dummy = reinterpret_cast<void*>(0x10);
  double* dd = static_cast<double*>(dummy);
  EXPECT(csg_step_noalloc(2048, 1.0, nullptr, 0, 0, 0.01, 1, dd, dd, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_step_noalloc(100, 1.0, nullptr, 0, 0, 0.01, 1, dd, dd, nullptr) == CSG_ERR_UNSUPPORTED);
  EXPECT(csg_step_noalloc(64, 1.0, nullptr, 0, 0, 0.01, 1, nullptr, dd, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_step_noalloc(64, 1.0, nullptr, 0, -1, 0.01, 1, dd, dd, nullptr) == CSG_ERR_ARG);
  EXPECT(csg_step_noalloc(64, 1.0, nullptr, 0, 0, NAN, 1, dd, dd, nullptr) == CSG_ERR_ARG);

  // ---- error cleanup paths: valid arguments, no device ----
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
  (void)hipGetLastError();
  if (ndev == 0) {
    ctx = reinterpret_cast<csg_ctx*>(0x1);
    EXPECT(csg_create(&ctx, 64, 1.0, &p, 0, 0) == CSG_ERR_HIP && ctx == nullptr);
    EXPECT(csg_create(&ctx, 256, 1.0, nullptr, 1, 0) == CSG_ERR_HIP && ctx == nullptr);
    b = reinterpret_cast<csg_batch*>(0x1);
    EXPECT(csg_batch_create(&b, 2, ns_ok, nullptr, nullptr, 1, 0) == CSG_ERR_HIP && b == nullptr);
    EXPECT(csg_step_noalloc(64, 1.0, &p, 0, 0, 0.01, 1, dd, dd, nullptr) == CSG_ERR_HIP);
    EXPECT(csg_copy_bandwidth(0, 1 << 20, 1, &x) == CSG_ERR_HIP);
  } else {
    std::printf("device present: no-device cleanup cases skipped\n");
  }
  if (failures == 0) std::printf("ALL OK\n");
  return failures == 0 ? 0 : 1;
}
