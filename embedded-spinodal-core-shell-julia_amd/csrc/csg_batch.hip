// Ensembles: many independent instances of the time loop of run_simulation
// (core_shell.jl:84-103) advanced together on one GPU.
//
// Instances are grouped by grid size into classes.  A class keeps each of its
// device arrays as one slab (instance q at q times the per-instance size) and
// runs every pass as ONE launch over all its instances (instance = grid y),
// with per-instance parameters read from a device table; each launch then has
// (tiles or lines) x instances workgroups.  The classes are independent: the
// captured graph of 32 loop iterations forks one branch per class.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <memory>

#include "../../include/csg.h"
#include "csg_host.hpp"
#include "csg_persist.hpp"

using namespace csg;

#define HIPCHK CSG_HIPCHK

namespace {

struct BatchClass {
  int n = 0;
  int m = 0;                     // instances
  int nt = 0;                    // non-temporal pass outputs (the ensemble's choice)
  Ops ops{};
  std::vector<int> members;      // batch index of slab slot q
  std::vector<DevParams> hparams;
  DevParams* dparams = nullptr;  // device copy, m entries
  hipStream_t stream = nullptr;  // branch stream
  size_t nn = 0;                 // sites per instance
  size_t tsize = 8;              // bytes per bulk real
  void* S = nullptr;             // state slab (double)
  void* A = nullptr;             // bulk slabs
  void* Bb = nullptr;
  void* phi = nullptr;
  double* real = nullptr;        // one instance of host I/O staging
  void* tw = nullptr;
  void* tw64 = nullptr;
  double* part = nullptr;        // m partial records
  long long part_stride = 0;     // doubles per instance
  long long* ctr = nullptr;      // 2 per instance
  double* hist = nullptr;        // m x hist_cap x 6
  size_t hist_cap = 0;           // rows per instance
  int pitems[4] = {0, 0, 0, 0};  // work items per member and pass of the persistent kernel (0: no such kernel)
};

}  // namespace

struct csg_batch {
  int count = 0;
  int precision = 0;
  int device = 0;
  std::vector<BatchClass> cls;
  std::vector<int> cls_of, slot_of;
  std::vector<char> fields_set;
  std::vector<long long> nonfinite;
  hipStream_t stream = nullptr;  // origin of the fork/join
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t gexec = nullptr;
  double gdt = 0.0;
  int64_t gevery = 1;   // history interval captured in gexec
  size_t bytes = 0;
  int nt = 0;   // non-temporal pass outputs (nontemporal_for the ensemble's footprint)
  // persistent mode (csg_batch_set_persistent): one launch per block of steps
  bool persistent = false;
  int pG = 0;                              // resident workgroups of the persistent kernel
  int ninst = 0;                           // members (counters)
  std::unique_ptr<PersistPlan> hplan;      // host copy of the device plan
  PersistPlan* dplan = nullptr;
  long long* dranges = nullptr;
  unsigned* dctr = nullptr;
  int* dflag = nullptr;
  hipGraphExec_t pexec = nullptr;          // captured block of STEPS_PER_GRAPH iterations
};

namespace {

int bmalloc(csg_batch* b, void** p, size_t bytes) {
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();   // leave no stale error for later checks
    *p = nullptr;
    return fail(CSG_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
  }
  b->bytes += bytes;
  return CSG_OK;
}

// Pass arguments covering all instances of class k.
PassArgs class_args(const BatchClass& k) {
  PassArgs a{};
  a.nt = k.nt;
  a.p = k.hparams[0];
  a.phi = k.phi;
  a.ninst = k.m;
  a.is.params = k.dparams;
  a.is.phi = static_cast<long long>(k.nn * k.tsize);
  a.is.part = k.part_stride;
  return a;
}

long long slab_stride(const BatchClass& k, const void* buf) {
  return static_cast<long long>(k.nn * (buf == k.S ? sizeof(double) : k.tsize));
}

void bind_in(const BatchClass& k, PassArgs& a, const void* in0, const void* in1) {
  a.in0 = in0;
  a.in1 = in1;
  a.is.in0 = in0 ? slab_stride(k, in0) : 0;
  a.is.in1 = in1 ? slab_stride(k, in1) : 0;
}
void bind_out(const BatchClass& k, PassArgs& a, void* out0, void* out1) {
  a.out0 = out0;
  a.out1 = out1;
  a.is.out0 = out0 ? slab_stride(k, out0) : 0;
  a.is.out1 = out1 ? slab_stride(k, out1) : 0;
}

FinalArgs class_final(const BatchClass& k, double* hist, double dt, int64_t every) {
  FinalArgs f{};
  f.every = every;
  f.p1 = k.part;
  f.p2 = k.part + 5 * k.ops.n_row1;
  f.p3 = k.part + 5 * k.ops.n_row1 + k.ops.n_col;
  f.n1 = k.ops.n_row1;
  f.n2 = k.ops.n_col;
  f.n3 = k.ops.n_row3;
  f.hist = hist;
  f.ctr = k.ctr;
  f.dt = dt;
  f.spacing2 = k.hparams[0].spacing * k.hparams[0].spacing;
  f.inv_n2 = 1.0 / (static_cast<double>(k.n) * k.n);
  f.kappa = k.hparams[0].kappa;
  f.params = k.dparams;
  f.s_part = k.part_stride;
  f.s_hist = static_cast<long long>(k.hist_cap) * 6;
  return f;
}

// One loop iteration (core_shell.jl:90-95) of every instance of class k.
void enqueue_class_iteration(const BatchClass& k, double dt, bool advance, hipStream_t s, int64_t every) {
  PassArgs a = class_args(k);
  bind_in(k, a, k.S, nullptr);
  bind_out(k, a, k.A, nullptr);
  a.part = k.part;
  k.ops.p1(a, k.tw, s);
  a = class_args(k);
  bind_in(k, a, k.S, k.A);
  bind_out(k, a, k.A, k.Bb);
  a.part = k.part + 5 * k.ops.n_row1;
  k.ops.p2(a, k.tw, s);
  a = class_args(k);
  bind_in(k, a, k.A, k.Bb);
  bind_out(k, a, k.A, k.Bb);
  a.part = k.part + 5 * k.ops.n_row1 + k.ops.n_col;
  k.ops.p3(a, k.tw, s);
  if (advance) {
    a = class_args(k);
    bind_in(k, a, k.A, k.Bb);
    bind_out(k, a, k.S, nullptr);
    a.dt = dt;
    a.fin = class_final(k, k.hist, dt, every);
    k.ops.p4(a, k.tw, s);
  } else {
    hipLaunchKernelGGL(finalize_kernel<FIN_NT>, dim3(k.m), dim3(FIN_NT), 0, s, class_final(k, k.hist, dt, every));
  }
}

// `iters` loop iterations of every class, one branch per class, forked from
// and joined back into b->stream (captured or eager).
int enqueue_iterations(csg_batch* b, double dt, int iters, bool last_advances, int64_t every) {
  HIPCHK(hipEventRecord(b->fork, b->stream));
  for (size_t ci = 0; ci < b->cls.size(); ++ci) {
    const BatchClass& k = b->cls[ci];
    HIPCHK(hipStreamWaitEvent(k.stream, b->fork, 0));
    for (int it = 0; it < iters; ++it)
      enqueue_class_iteration(k, dt, it + 1 < iters || last_advances, k.stream, every);
    HIPCHK(hipEventRecord(b->join[ci], k.stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->join[ci], 0));
  }
  HIPCHK(hipGetLastError());
  return CSG_OK;
}

int build_batch_graph(csg_batch* b, double dt, int64_t every) {
  if (b->gexec && b->gdt == dt && b->gevery == every) return CSG_OK;
  if (b->gexec) {
    HIPCHK(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_iterations(b, dt, STEPS_PER_GRAPH, true, every);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) return rc;
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->gexec = nullptr;
    return fail(CSG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  b->gdt = dt;
  b->gevery = every;
  return CSG_OK;
}

int ensure_batch_hist(csg_batch* b, size_t rows) {
  for (BatchClass& k : b->cls) {
    if (rows <= k.hist_cap) continue;
    if (b->gexec) {  // the captured graph holds the old history pointers
      HIPCHK(hipGraphExecDestroy(b->gexec));
      b->gexec = nullptr;
    }
    if (k.hist) {
      HIPCHK(hipFree(k.hist));
      b->bytes -= k.hist_cap * 6 * sizeof(double) * k.m;
    }
    k.hist = nullptr;
    k.hist_cap = 0;
    if (int rc = bmalloc(b, reinterpret_cast<void**>(&k.hist), rows * 6 * sizeof(double) * k.m)) return rc;
    k.hist_cap = rows;
  }
  return CSG_OK;
}

// ---- persistent mode --------------------------------------------------------

// Wall-time limit of one dependency wait of the persistent kernel: far above
// any step (milliseconds), far below the harness' silence limit.
constexpr long long PERSIST_WAIT_TICKS = 200000000LL;   // 2 s at 100 MHz

// The plan of one persistent launch: per class its pass arguments (as the
// batched launches bind them, with the persistent schedules' partial
// layout) and the flattened item offsets per pass.
void build_persist_plan(const csg_batch* b, double dt, int64_t every, PersistPlan* pl) {
  std::memset(static_cast<void*>(pl), 0, sizeof(PersistPlan));
  pl->ncls = static_cast<int>(b->cls.size());
  int base = 0;
  for (int c = 0; c < pl->ncls; ++c) {
    const BatchClass& k = b->cls[c];
    PersistClass& pc = pl->cls[c];
    pc.n = k.n;
    pc.ninst = k.m;
    pc.inst_base = base;
    base += k.m;
    pc.per_step = 0;
    for (int q = 0; q < 4; ++q) {
      pc.items[q] = k.pitems[q];
      pc.per_step += k.pitems[q];
    }
    pc.tw = k.tw;
    double* p2 = k.part + 5 * k.pitems[0];
    double* p3 = p2 + k.pitems[1];
    PassArgs a = class_args(k);
    bind_in(k, a, k.S, nullptr);
    bind_out(k, a, k.A, nullptr);
    a.part = k.part;
    pc.args[0] = a;
    a = class_args(k);
    bind_in(k, a, k.S, k.A);
    bind_out(k, a, k.A, k.Bb);
    a.part = p2;
    pc.args[1] = a;
    a = class_args(k);
    bind_in(k, a, k.A, k.Bb);
    bind_out(k, a, k.A, k.Bb);
    a.part = p3;
    pc.args[2] = a;
    a = class_args(k);
    bind_in(k, a, k.A, k.Bb);
    bind_out(k, a, k.S, nullptr);
    a.dt = dt;
    FinalArgs f = class_final(k, k.hist, dt, every);
    f.p2 = p2;
    f.p3 = p3;
    f.n1 = k.pitems[0];
    f.n2 = k.pitems[1];
    f.n3 = k.pitems[2];
    a.fin = f;
    pc.args[3] = a;
  }
  for (int q = 0; q < 4; ++q) {
    pl->off[q][0] = 0;
    for (int c = 0; c < pl->ncls; ++c)
      pl->off[q][c + 1] = pl->off[q][c] + static_cast<long long>(pl->cls[c].ninst) * pl->cls[c].items[q];
  }
}

// Contiguous item ranges per workgroup and pass, of equal estimated cost
// (sites per item: an item of an N x N member covers N^2 / items sites).
std::vector<long long> persist_ranges(const PersistPlan& pl, int G) {
  std::vector<long long> r(4 * static_cast<size_t>(G + 1), 0);
  for (int q = 0; q < 4; ++q) {
    double total = 0.0;
    for (int c = 0; c < pl.ncls; ++c) total += static_cast<double>(pl.cls[c].ninst) * pl.cls[c].n * pl.cls[c].n;
    long long* rq = r.data() + static_cast<size_t>(q) * (G + 1);
    double acc = 0.0;
    int w = 1;
    for (int c = 0; c < pl.ncls; ++c) {
      const PersistClass& k = pl.cls[c];
      const double cost = static_cast<double>(k.n) * k.n / k.items[q];
      for (long long g = pl.off[q][c]; g < pl.off[q][c + 1]; ++g) {
        acc += cost;
        while (w < G && acc >= total * w / G) rq[w++] = g + 1;
      }
    }
    while (w <= G) rq[w++] = pl.off[q][pl.ncls];
  }
  return r;
}

int check_batch(const csg_batch* b, int32_t i) {
  if (!b) return fail(CSG_ERR_ARG, "null ensemble");
  if (i < 0 || i >= b->count) return fail(CSG_ERR_ARG, "instance index out of range");
  return CSG_OK;
}

}  // namespace

extern "C" {

int csg_batch_create(csg_batch** out, int32_t count, const int32_t* n, const double* spacing,
                     const csg_params* params, int32_t precision, int32_t device) {
  if (!out) return fail(CSG_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (count < 1 || !n) return fail(CSG_ERR_ARG, "count must be >= 1 and n non-NULL");
  if (precision != CSG_FP64 && precision != CSG_MIXED) return fail(CSG_ERR_ARG, "unknown precision");
  for (int i = 0; i < count; ++i) {
    const double sp = spacing ? spacing[i] : 1.0;
    if (!(sp > 0.0) || !std::isfinite(sp)) return fail(CSG_ERR_ARG, "spacing must be > 0");
    Ops probe{};
    const bool ok = precision == CSG_FP64 ? ops_for<double>(n[i], &probe) : ops_for<float>(n[i], &probe);
    if (!ok) return fail(CSG_ERR_UNSUPPORTED, unsupported_size_message(n[i]));
  }
  csg_batch* b = new csg_batch();
  b->count = count;
  b->precision = precision;
  b->device = device;
  b->cls_of.assign(count, -1);
  b->slot_of.assign(count, -1);
  b->fields_set.assign(count, 0);
  b->nonfinite.assign(count, -1);
  for (int i = 0; i < count; ++i) {
    int ci = -1;
    for (size_t k = 0; k < b->cls.size(); ++k)
      if (b->cls[k].n == n[i]) ci = static_cast<int>(k);
    if (ci < 0) {
      b->cls.emplace_back();
      ci = static_cast<int>(b->cls.size()) - 1;
      b->cls[ci].n = n[i];
    }
    BatchClass& k = b->cls[ci];
    b->cls_of[i] = ci;
    b->slot_of[i] = k.m++;
    k.members.push_back(i);
    csg_params p;
    if (params) p = params[i]; else csg_default_params(&p);
    const double sp = spacing ? spacing[i] : 1.0;
    k.hparams.push_back(DevParams{p.A, p.B, p.chi, p.P, p.ca, p.cb, p.cc, p.kappa, p.mobility, p.Q, sp,
                                  (1.0 / sp) / n[i]});
  }
  auto cleanup = [&](int code) { csg_batch_destroy(b); return code; };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess)
    return cleanup(fail(CSG_ERR_HIP, std::string("stream/event: ") + hipGetErrorString(e)));
  b->join.assign(b->cls.size(), nullptr);
  for (size_t ci = 0; ci < b->cls.size(); ++ci) {
    BatchClass& k = b->cls[ci];
    const int nk = k.n;
    if (precision == CSG_FP64) ops_for<double>(nk, &k.ops); else ops_for<float>(nk, &k.ops);
    k.nn = static_cast<size_t>(nk) * nk;
    k.tsize = precision == CSG_FP64 ? sizeof(double) : sizeof(float);
    k.part_stride = 5LL * k.ops.max_row1 + k.ops.max_col + k.ops.max_row3;
    if (persist_size_supported(nk)) {
      persist_items_of(nk, precision == CSG_FP64, k.pitems);
      const long long need = 5LL * k.pitems[0] + k.pitems[1] + k.pitems[2];
      if (need > k.part_stride) k.part_stride = need;
    }
    if ((e = hipStreamCreateWithFlags(&k.stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->join[ci], hipEventDisableTiming)) != hipSuccess)
      return cleanup(fail(CSG_ERR_HIP, std::string("stream/event: ") + hipGetErrorString(e)));
    int rc;
    const size_t m = static_cast<size_t>(k.m);
    if ((rc = bmalloc(b, &k.S, m * k.nn * sizeof(double)))) return cleanup(rc);
    if ((rc = bmalloc(b, &k.A, m * k.nn * k.tsize))) return cleanup(rc);
    if ((rc = bmalloc(b, &k.Bb, m * k.nn * k.tsize))) return cleanup(rc);
    if ((rc = bmalloc(b, &k.phi, m * k.nn * k.tsize))) return cleanup(rc);
    if ((rc = bmalloc(b, reinterpret_cast<void**>(&k.real), k.nn * sizeof(double)))) return cleanup(rc);
    if ((rc = bmalloc(b, reinterpret_cast<void**>(&k.part), m * k.part_stride * sizeof(double)))) return cleanup(rc);
    if ((rc = bmalloc(b, reinterpret_cast<void**>(&k.ctr), m * 2 * sizeof(long long)))) return cleanup(rc);
    if ((rc = bmalloc(b, reinterpret_cast<void**>(&k.dparams), m * sizeof(DevParams)))) return cleanup(rc);
    e = hipMemcpy(k.dparams, k.hparams.data(), m * sizeof(DevParams), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
    // twiddles exp(-2 pi i j / N) and the FFT stage tables
    const std::vector<cplx<double>> tw64 = make_twiddles<double>(nk);
    const std::vector<cplx<float>> tw32 = precision == CSG_FP64 ? std::vector<cplx<float>>() : make_twiddles<float>(nk);
    if ((rc = bmalloc(b, &k.tw64, tw64.size() * sizeof(cplx<double>)))) return cleanup(rc);
    e = hipMemcpy(k.tw64, tw64.data(), tw64.size() * sizeof(cplx<double>), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
    if (precision == CSG_FP64) {
      k.tw = k.tw64;
    } else {
      if ((rc = bmalloc(b, &k.tw, tw32.size() * sizeof(cplx<float>)))) return cleanup(rc);
      e = hipMemcpy(k.tw, tw32.data(), tw32.size() * sizeof(cplx<float>), hipMemcpyHostToDevice);
      if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
    }
  }
  b->nt = nontemporal_for(b->bytes, true);
  for (BatchClass& k : b->cls) k.nt = b->nt;
  *out = b;
  return CSG_OK;
}

int csg_batch_destroy(csg_batch* b) {
  if (!b) return CSG_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (BatchClass& k : b->cls)
    if (k.stream) (void)hipStreamSynchronize(k.stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->pexec) (void)hipGraphExecDestroy(b->pexec);
  for (void* p : {static_cast<void*>(b->dplan), static_cast<void*>(b->dranges), static_cast<void*>(b->dctr),
                  static_cast<void*>(b->dflag)})
    if (p) (void)hipFree(p);
  for (BatchClass& k : b->cls) {
    void* bufs[] = {k.S, k.A, k.Bb, k.phi, k.real, k.part, k.ctr, k.hist, k.dparams, k.tw64};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    if (k.tw && k.tw != k.tw64) (void)hipFree(k.tw);
    if (k.stream) (void)hipStreamDestroy(k.stream);
  }
  for (hipEvent_t ev : b->join)
    if (ev) (void)hipEventDestroy(ev);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CSG_OK;
}

int64_t csg_batch_device_bytes(const csg_batch* b) { return b ? static_cast<int64_t>(b->bytes) : 0; }

int csg_batch_set_fields(csg_batch* b, int32_t i, const double* c0, const double* phi) {
  if (int rc = check_batch(b, i)) return rc;
  if (!c0 || !phi) return fail(CSG_ERR_ARG, "c0 and phi must be non-NULL");
  HIPCHK(hipSetDevice(b->device));
  BatchClass& k = b->cls[b->cls_of[i]];
  const size_t q = static_cast<size_t>(b->slot_of[i]);
  char* phi_q = static_cast<char*>(k.phi) + q * k.nn * k.tsize;
  if (b->precision == CSG_FP64) {
    HIPCHK(hipMemcpyAsync(phi_q, phi, k.nn * sizeof(double), hipMemcpyHostToDevice, k.stream));
  } else {
    std::vector<float> tmp(k.nn);
    for (size_t j = 0; j < k.nn; ++j) tmp[j] = static_cast<float>(phi[j]);
    HIPCHK(hipMemcpyAsync(phi_q, tmp.data(), k.nn * sizeof(float), hipMemcpyHostToDevice, k.stream));
    HIPCHK(hipStreamSynchronize(k.stream));
  }
  HIPCHK(hipMemcpyAsync(k.real, c0, k.nn * sizeof(double), hipMemcpyHostToDevice, k.stream));
  PassArgs a{};
  a.p = k.hparams[q];
  a.in0 = k.real;
  a.out0 = static_cast<char*>(k.S) + q * k.nn * sizeof(double);
  k.ops.fwd(a, k.tw64, k.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(k.stream));
  b->fields_set[i] = 1;
  return CSG_OK;
}

int csg_batch_run(csg_batch* b, double dt, int64_t steps, int64_t hist_every, double* const* history) {
  if (!b) return fail(CSG_ERR_ARG, "null ensemble");
  for (int i = 0; i < b->count; ++i)
    if (!b->fields_set[i]) return fail(CSG_ERR_STATE, "fields of instance " + std::to_string(i) + " not set");
  if (steps < 0 || hist_every < 1 || !std::isfinite(dt)) return fail(CSG_ERR_ARG, "bad steps/hist_every/dt");
  HIPCHK(hipSetDevice(b->device));
  // the device keeps rows s = 0, hist_every, 2 hist_every, ... only
  const int64_t rows = steps / hist_every + 1;
  if (int rc = ensure_batch_hist(b, static_cast<size_t>(rows))) return rc;
  for (BatchClass& k : b->cls) {
    std::vector<long long> init(2 * static_cast<size_t>(k.m));
    for (int q = 0; q < k.m; ++q) {
      init[2 * q] = 0;
      init[2 * q + 1] = -1;
    }
    HIPCHK(hipMemcpyAsync(k.ctr, init.data(), init.size() * sizeof(long long), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  int64_t done = 0;
  if (b->persistent) {
    // one launch per block of iterations; the plan (pass arguments, dt,
    // history binding) is device data, so the captured block stays valid
    build_persist_plan(b, dt, hist_every, b->hplan.get());
    HIPCHK(hipMemcpyAsync(b->dplan, b->hplan.get(), sizeof(PersistPlan), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemsetAsync(b->dflag, 0, sizeof(int), b->stream));
    PersistArgs pa{};
    pa.plan = b->dplan;
    pa.ranges = b->dranges;
    pa.ctr = b->dctr;
    pa.flag = b->dflag;
    pa.spin_ticks = PERSIST_WAIT_TICKS;
    const bool f64 = b->precision == CSG_FP64;
    if (steps >= 2 * STEPS_PER_GRAPH) {
      if (!b->pexec) {
        pa.steps = STEPS_PER_GRAPH;
        pa.last_advances = 1;
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = persist_launch(f64, pa, b->pG, b->ninst, b->stream);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (le != hipSuccess) return fail(CSG_ERR_HIP, std::string("persistent launch: ") + hipGetErrorString(le));
        if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->pexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          b->pexec = nullptr;
          return fail(CSG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
      }
      while (steps - done >= STEPS_PER_GRAPH) {
        HIPCHK(hipGraphLaunch(b->pexec, b->stream));
        done += STEPS_PER_GRAPH;
      }
    }
    pa.steps = static_cast<int>(steps - done) + 1;
    pa.last_advances = 0;
    HIPCHK(persist_launch(f64, pa, b->pG, b->ninst, b->stream));
  } else {
    if (steps >= 2 * STEPS_PER_GRAPH) {
      if (int rc = build_batch_graph(b, dt, hist_every)) return rc;
      while (steps - done >= STEPS_PER_GRAPH) {
        HIPCHK(hipGraphLaunch(b->gexec, b->stream));
        done += STEPS_PER_GRAPH;
      }
    }
    // the remaining iterations, then the last history row without a step
    if (int rc = enqueue_iterations(b, dt, static_cast<int>(steps - done) + 1, false, hist_every)) return rc;
  }
  std::vector<std::vector<long long>> ctr(b->cls.size());
  std::vector<std::vector<double>> h(b->cls.size());
  for (size_t ci = 0; ci < b->cls.size(); ++ci) {
    const BatchClass& k = b->cls[ci];
    ctr[ci].resize(2 * static_cast<size_t>(k.m));
    HIPCHK(hipMemcpyAsync(ctr[ci].data(), k.ctr, ctr[ci].size() * sizeof(long long), hipMemcpyDeviceToHost,
                          b->stream));
    if (history) {
      // instance q's rows [0, rows) are at q * hist_cap * 6: one copy up to
      // the last instance's last row
      const size_t len = (static_cast<size_t>(k.m) - 1) * k.hist_cap * 6 + static_cast<size_t>(rows) * 6;
      h[ci].resize(len);
      HIPCHK(hipMemcpyAsync(h[ci].data(), k.hist, len * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    }
  }
  int stalled = 0;
  if (b->persistent) HIPCHK(hipMemcpyAsync(&stalled, b->dflag, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (stalled)
    return fail(CSG_ERR_HIP, "persistent ensemble launch stalled (a dependency wait exceeded its time limit); "
                             "the members' state is undefined");
  bool any_bad = false;
  long long first_bad = -1;
  int bad_inst = -1;
  for (int i = 0; i < b->count; ++i) {
    const size_t ci = static_cast<size_t>(b->cls_of[i]);
    const size_t q = static_cast<size_t>(b->slot_of[i]);
    b->nonfinite[i] = ctr[ci][2 * q + 1];
    if (b->nonfinite[i] >= 0 && !any_bad) {
      any_bad = true;
      first_bad = b->nonfinite[i];
      bad_inst = i;
    }
    if (history && history[i]) {
      const double* src = h[ci].data() + q * b->cls[ci].hist_cap * 6;
      for (int64_t r = 0; r < rows; ++r)
        for (int col = 0; col < 6; ++col)
          history[i][r + col * rows] = src[static_cast<size_t>(r) * 6 + col];
    }
  }
  if (any_bad)
    return fail(CSG_ERR_NONFINITE, "Non-finite composition at step " + std::to_string(first_bad) + " (instance " +
                                       std::to_string(bad_inst) + ")");
  return CSG_OK;
}

int csg_batch_set_persistent(csg_batch* b, int32_t enable) {
  if (!b) return fail(CSG_ERR_ARG, "null ensemble");
  if (!enable) {
    b->persistent = false;
    return CSG_OK;
  }
  if (b->persistent) return CSG_OK;
  if (b->cls.size() > static_cast<size_t>(PERSIST_MAX_CLASSES))
    return fail(CSG_ERR_UNSUPPORTED, "the persistent kernel takes at most 4 grid sizes");
  for (const BatchClass& k : b->cls)
    if (!k.pitems[0])
      return fail(CSG_ERR_UNSUPPORTED, "no persistent kernel for grid size " + std::to_string(k.n) +
                                           " (compiled: 128, 256, 384, 512)");
  HIPCHK(hipSetDevice(b->device));
  int per_cu = 0, dev = 0;
  if (persist_blocks_per_cu(b->precision == CSG_FP64, &per_cu) || per_cu < 1)
    return fail(CSG_ERR_HIP, "persistent kernel: occupancy query failed");
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  b->pG = per_cu * prop.multiProcessorCount;
  b->ninst = 0;
  for (const BatchClass& k : b->cls) b->ninst += k.m;
  if (!b->hplan) b->hplan.reset(new PersistPlan());
  build_persist_plan(b, 0.0, 1, b->hplan.get());
  const std::vector<long long> ranges = persist_ranges(*b->hplan, b->pG);
  if (!b->dplan && bmalloc(b, reinterpret_cast<void**>(&b->dplan), sizeof(PersistPlan))) return CSG_ERR_HIP;
  if (b->dranges) {
    HIPCHK(hipFree(b->dranges));
    b->dranges = nullptr;
  }
  if (bmalloc(b, reinterpret_cast<void**>(&b->dranges), ranges.size() * sizeof(long long))) return CSG_ERR_HIP;
  if (!b->dctr && bmalloc(b, reinterpret_cast<void**>(&b->dctr), static_cast<size_t>(b->ninst) * sizeof(unsigned)))
    return CSG_ERR_HIP;
  if (!b->dflag && bmalloc(b, reinterpret_cast<void**>(&b->dflag), sizeof(int))) return CSG_ERR_HIP;
  HIPCHK(hipMemcpy(b->dranges, ranges.data(), ranges.size() * sizeof(long long), hipMemcpyHostToDevice));
  if (b->pexec) {
    HIPCHK(hipGraphExecDestroy(b->pexec));
    b->pexec = nullptr;
  }
  b->persistent = true;
  return CSG_OK;
}

int csg_batch_get_field(csg_batch* b, int32_t i, double* out) {
  if (int rc = check_batch(b, i)) return rc;
  if (!b->fields_set[i]) return fail(CSG_ERR_STATE, "fields not set");
  if (!out) return fail(CSG_ERR_ARG, "out is NULL");
  HIPCHK(hipSetDevice(b->device));
  BatchClass& k = b->cls[b->cls_of[i]];
  const size_t q = static_cast<size_t>(b->slot_of[i]);
  PassArgs a{};
  a.p = k.hparams[q];
  a.in0 = static_cast<const char*>(k.S) + q * k.nn * sizeof(double);
  a.out0 = k.real;
  k.ops.inv_state(a, k.tw64, k.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, k.real, k.nn * sizeof(double), hipMemcpyDeviceToHost, k.stream));
  HIPCHK(hipStreamSynchronize(k.stream));
  return CSG_OK;
}

int64_t csg_batch_nonfinite_step(const csg_batch* b, int32_t i) {
  if (!b || i < 0 || i >= b->count) return -1;
  return b->nonfinite[i];
}

}  // extern "C"

// ---- kernel schedules per size class ------------------------------------------

namespace {

BatchClass* find_class(csg_batch* b, int n) {
  for (BatchClass& k : b->cls)
    if (k.n == n) return &k;
  return nullptr;
}

constexpr int BATCH_TUNE_REPS = 3;

// Mean device time per loop iteration of class k alone (its own captured
// 32-iteration loop) with schedules sel.
int time_class(BatchClass& k, double dt, const int sel[4], hipEvent_t e0, hipEvent_t e1, double* us) {
  select_schedule(k.ops, sel);
  std::vector<long long> init(2 * static_cast<size_t>(k.m));
  for (int q = 0; q < k.m; ++q) {
    init[2 * q] = 0;
    init[2 * q + 1] = -1;
  }
  hipGraph_t g;
  hipGraphExec_t ge = nullptr;
  HIPCHK(hipStreamBeginCapture(k.stream, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < STEPS_PER_GRAPH; ++it) enqueue_class_iteration(k, dt, true, k.stream, 1);
  HIPCHK(hipStreamEndCapture(k.stream, &g));
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  auto reset = [&]() {
    return hipMemcpyAsync(k.ctr, init.data(), init.size() * sizeof(long long), hipMemcpyHostToDevice, k.stream);
  };
  float ms = 0.f;
  if ((e = reset()) == hipSuccess && (e = hipGraphLaunch(ge, k.stream)) == hipSuccess && (e = reset()) == hipSuccess &&
      (e = hipEventRecord(e0, k.stream)) == hipSuccess) {
    for (int r = 0; r < BATCH_TUNE_REPS && e == hipSuccess; ++r) e = hipGraphLaunch(ge, k.stream);
    if (e == hipSuccess) e = hipEventRecord(e1, k.stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  }
  (void)hipGraphExecDestroy(ge);
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("tuning run: ") + hipGetErrorString(e));
  *us = 1e3 * ms / (BATCH_TUNE_REPS * STEPS_PER_GRAPH);
  return CSG_OK;
}

}  // namespace

extern "C" {

int csg_batch_get_schedule(const csg_batch* b, int32_t n, int32_t* sel) {
  if (!b || !sel) return fail(CSG_ERR_ARG, "null ensemble or sel");
  const BatchClass* k = find_class(const_cast<csg_batch*>(b), n);
  if (!k) return fail(CSG_ERR_ARG, "no member of size " + std::to_string(n));
  for (int q = 0; q < 4; ++q) sel[q] = k->ops.sel[q];
  return CSG_OK;
}

int csg_batch_set_schedule(csg_batch* b, int32_t n, const int32_t* sel) {
  if (!b || !sel) return fail(CSG_ERR_ARG, "null ensemble or sel");
  BatchClass* k = find_class(b, n);
  if (!k) return fail(CSG_ERR_ARG, "no member of size " + std::to_string(n));
  const int s[4] = {sel[0], sel[1], sel[2], sel[3]};
  if (!select_schedule(k->ops, s)) return fail(CSG_ERR_ARG, "schedule index out of range");
  if (b->gexec) {
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  return CSG_OK;
}

int csg_batch_autotune(csg_batch* b, double dt, int32_t rounds) {
  if (!b) return fail(CSG_ERR_ARG, "null ensemble");
  for (int i = 0; i < b->count; ++i)
    if (!b->fields_set[i]) return fail(CSG_ERR_STATE, "fields of instance " + std::to_string(i) + " not set");
  if (rounds < 1 || !std::isfinite(dt)) return fail(CSG_ERR_ARG, "rounds >= 1 and finite dt required");
  HIPCHK(hipSetDevice(b->device));
  if (b->gexec) {
    HIPCHK(hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  if (int rc = ensure_batch_hist(b, static_cast<size_t>(STEPS_PER_GRAPH) * BATCH_TUNE_REPS + 1)) return rc;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = CSG_OK;
  for (BatchClass& k : b->cls) {
    // the resident fields of the class are kept and restored afterwards
    const size_t bytes = static_cast<size_t>(k.m) * k.nn * sizeof(double);
    void* saved = nullptr;
    hipError_t e = hipMalloc(&saved, bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(saved, k.S, bytes, hipMemcpyDeviceToDevice, k.stream);
    if (e != hipSuccess) {
      rc = fail(CSG_ERR_HIP, std::string("tuning buffer: ") + hipGetErrorString(e));
      if (saved) (void)hipFree(saved);
      break;
    }
    int best[4] = {k.ops.sel[0], k.ops.sel[1], k.ops.sel[2], k.ops.sel[3]};
    double best_us = 0.0;
    rc = time_class(k, dt, best, e0, e1, &best_us);
    for (int r = 0; r < rounds && rc == CSG_OK; ++r)
      for (int q = 0; q < 4 && rc == CSG_OK; ++q)
        for (int i = 0; i < k.ops.cand[q].n && rc == CSG_OK; ++i) {
          if (i == best[q]) continue;
          int trial[4] = {best[0], best[1], best[2], best[3]};
          trial[q] = i;
          double us = 0.0;
          rc = time_class(k, dt, trial, e0, e1, &us);
          if (rc == CSG_OK && us < best_us) {
            best_us = us;
            best[q] = i;
          }
        }
    select_schedule(k.ops, best);
    e = hipMemcpyAsync(k.S, saved, bytes, hipMemcpyDeviceToDevice, k.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(k.stream);
    (void)hipFree(saved);
    if (rc == CSG_OK && e != hipSuccess) rc = fail(CSG_ERR_HIP, std::string("restore: ") + hipGetErrorString(e));
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // extern "C"
