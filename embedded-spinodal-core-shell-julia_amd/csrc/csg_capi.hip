// C ABI of the core-shell step library: context, device buffers, per-size
// kernel dispatch, the graph-captured time loop (core_shell.jl:84-103) and the
// single-evaluation operators (energy, chemical potential, transport rate).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/csg.h"
#include "csg_host.hpp"

using namespace csg;

namespace {
thread_local std::string g_err;
}  // namespace

int csg::fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK CSG_HIPCHK

struct csg_ctx {
  int n = 0;
  int precision = 0;
  int device = 0;
  double spacing = 1.0;
  csg_params p{};
  DevParams dp{};
  Ops ops{};
  hipStream_t stream = nullptr;
  size_t tsize = 8;       // bytes per bulk real
  void* S = nullptr;      // state (double, line layout)
  void* S2 = nullptr;     // scratch state for single evaluations
  void* A = nullptr;      // F' / Gx / Jx (bulk)
  void* Bb = nullptr;     // Gy / Jy (bulk)
  void* phi = nullptr;    // real rows (bulk)
  double* real = nullptr; // host I/O staging, N*N doubles
  void* tw = nullptr;     // twiddles (bulk)
  void* tw64 = nullptr;   // twiddles (double) for state transforms
  double* part = nullptr; // partials: P1 (5 each), P2, P3
  long long* ctr = nullptr;
  double* hist = nullptr;
  size_t hist_cap = 0;    // rows
  long long* hstage = nullptr;     // pinned host staging of ctr (2 words) + history rows
  size_t hstage_cap = 0;           // 8-byte words
  // non-finite checks during long runs: pinned slots for ctr[1] and events
  long long* flag_pin = nullptr;
  hipEvent_t flag_ev[2] = {nullptr, nullptr};
  // uploads from pageable host memory: two pinned chunks and the events of
  // their last transfers (made on first use)
  void* up_pin[2] = {nullptr, nullptr};
  hipEvent_t up_ev[2] = {nullptr, nullptr};
  bool up_busy[2] = {false, false};
  int up_next = 0;
  double* scratch_row = nullptr;
  double* rate = nullptr; // (N/2+1) x N complex doubles
  bool fields_set = false;
  long long nonfinite = -1;
  hipGraphExec_t gexec = nullptr;   // STEPS_PER_GRAPH loop iterations of csg_run
  double gdt = 0.0;
  int64_t gevery = 1;     // history interval captured in gexec
  DevParams gp{};         // parameters captured in gexec
  size_t bytes = 0;
  int nt = 0;             // non-temporal pass outputs (nontemporal_for the footprint at creation)
  // adjoint: checkpoints of the forward run, segment states, adjoint field
  std::vector<void*> ckpt;
  std::vector<void*> seg;
  void* G = nullptr;
  void* Xr = nullptr;
  double* adj_acc = nullptr;
  int64_t adj_steps = -1, adj_every = 0;
  double adj_dt = 0.0;
  // captured full-segment graphs of the adjoint: forward steps, recompute,
  // backward sweep (valid for adj_gK steps of adj_gdt with the current kernels)
  hipGraphExec_t adj_fwd = nullptr, adj_rec = nullptr, adj_bwd = nullptr;
  // every state kept (ckpt_every = 1): forward and backward graphs per block
  // of STEPS_PER_GRAPH steps, bound to the checkpoint buffers of a run of
  // adj_all_T steps
  std::vector<hipGraphExec_t> adj_all_fwd, adj_all_bwd;
  int64_t adj_all_T = 0;
  int64_t adj_gK = 0;
  double adj_gdt = 0.0;
  DevParams adj_gp{};     // parameters captured in the adjoint graphs
};

namespace {

// Test hook: CSG_TEST_DEVICE_BUDGET=<bytes> in the environment makes any
// allocation that would take a context past that many bytes fail as an
// out-of-memory hipMalloc would, so fallback paths can be exercised without
// exhausting the device.
bool over_test_budget(const csg_ctx* c, size_t bytes) {
  const char* e = std::getenv("CSG_TEST_DEVICE_BUDGET");
  if (!e || !*e) return false;
  const unsigned long long cap = std::strtoull(e, nullptr, 10);
  return c->bytes + bytes > cap;
}

int dmalloc(csg_ctx* c, void** p, size_t bytes) {
  if (over_test_budget(c, bytes)) {
    *p = nullptr;
    return fail(CSG_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + " bytes): over CSG_TEST_DEVICE_BUDGET");
  }
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    // clear the runtime's last error, so a caller that falls back to a
    // smaller allocation is not failed by a later hipGetLastError check
    (void)hipGetLastError();
    *p = nullptr;
    return fail(CSG_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
  }
  c->bytes += bytes;
  return CSG_OK;
}

PassArgs base_args(const csg_ctx* c) {
  PassArgs a{};
  a.p = c->dp;
  a.phi = c->phi;
  a.nt = c->nt;
  return a;
}

// Partials sub-buffers.
double* part1(const csg_ctx* c) { return c->part; }
double* part2(const csg_ctx* c) { return c->part + 5 * c->ops.n_row1; }
double* part3(const csg_ctx* c) { return c->part + 5 * c->ops.n_row1 + c->ops.n_col; }

FinalArgs final_args(const csg_ctx* c, double* hist, double dt, int64_t every = 1) {
  FinalArgs f{};
  f.every = every;
  f.p1 = part1(c);
  f.p2 = part2(c);
  f.p3 = part3(c);
  f.n1 = c->ops.n_row1;
  f.n2 = c->ops.n_col;
  f.n3 = c->ops.n_row3;
  f.hist = hist;
  f.ctr = c->ctr;
  f.dt = dt;
  f.spacing2 = c->spacing * c->spacing;
  f.inv_n2 = 1.0 / (static_cast<double>(c->n) * c->n);
  f.kappa = c->p.kappa;
  return f;
}

// P1..P3 on state `S`: partials + Gx/Gy (A/Bb) then Jx/Jy (A/Bb).
void enqueue_p123(csg_ctx* c, const void* S) {
  PassArgs a = base_args(c);
  a.in0 = S;
  a.out0 = c->A;
  a.part = part1(c);
  c->ops.p1(a, c->tw, c->stream);
  a = base_args(c);
  a.in0 = S;
  a.in1 = c->A;
  a.out0 = c->A;
  a.out1 = c->Bb;
  a.part = part2(c);
  c->ops.p2(a, c->tw, c->stream);
  a = base_args(c);
  a.in0 = c->A;
  a.in1 = c->Bb;
  a.out0 = c->A;
  a.out1 = c->Bb;
  a.part = part3(c);
  c->ops.p3(a, c->tw, c->stream);
}

// P4 on state `S` (read from `Sin` instead when given); with `hist` non-NULL
// its workgroup 0 also writes the history row of the step from the P1..P3
// partials.
void enqueue_p4(csg_ctx* c, void* S, double dt, double* hist, const void* Sin = nullptr, int64_t every = 1) {
  PassArgs a = base_args(c);
  a.in0 = c->A;
  a.in1 = c->Bb;
  a.out0 = S;
  a.sin = Sin;
  a.dt = dt;
  a.fin = final_args(c, hist, dt, every);
  c->ops.p4(a, c->tw, c->stream);
}

void enqueue_finalize(csg_ctx* c, double* hist, double dt, int64_t every = 1) {
  hipLaunchKernelGGL(finalize_kernel<FIN_NT>, dim3(1), dim3(FIN_NT), 0, c->stream,
                     final_args(c, hist, dt, every));
}

// One loop iteration of run_simulation with history (core_shell.jl:90-95);
// rows s with s % every == 0 are kept.
void enqueue_iteration(csg_ctx* c, double dt, bool advance, int64_t every) {
  enqueue_p123(c, c->S);
  if (advance) enqueue_p4(c, c->S, dt, c->hist, nullptr, every);
  else enqueue_finalize(c, c->hist, dt, every);
}

// Uploads from pageable host memory go through two pinned chunks of
// UPLOAD_CHUNK bytes: the host fills one (converting to T on the way) while
// the copy engine reads the other.  hipMemcpyAsync straight from pageable
// memory moved the 16 MB of csg_set_fields at 1024^2 in 6 ms, with the GPU
// idle and its clocks falling meanwhile, so the run that followed started
// slower.  A chunk is refilled only once the event after its last transfer
// has completed, also across calls; on return the copies are queued on the
// context's stream and `src` is no longer read.
constexpr size_t UPLOAD_CHUNK = size_t(2) << 20;

int ensure_staging(csg_ctx* c) {
  for (int b = 0; b < 2; ++b) {
    if (!c->up_pin[b]) HIPCHK(hipHostMalloc(&c->up_pin[b], UPLOAD_CHUNK, hipHostMallocDefault));
    if (!c->up_ev[b]) HIPCHK(hipEventCreateWithFlags(&c->up_ev[b], hipEventDisableTiming));
  }
  return CSG_OK;
}

template <class T>
int upload(csg_ctx* c, T* dst, const double* src, size_t count) {
  if (int rc = ensure_staging(c)) return rc;
  constexpr size_t PER = UPLOAD_CHUNK / sizeof(T);
  for (size_t off = 0; off < count; off += PER) {
    const int b = c->up_next;
    c->up_next ^= 1;
    if (c->up_busy[b]) {
      HIPCHK(hipEventSynchronize(c->up_ev[b]));
      c->up_busy[b] = false;
    }
    const size_t len = std::min(PER, count - off);
    T* pin = static_cast<T*>(c->up_pin[b]);
    if constexpr (std::is_same<T, double>::value) {
      std::memcpy(pin, src + off, len * sizeof(double));
    } else {
      for (size_t i = 0; i < len; ++i) pin[i] = static_cast<T>(src[off + i]);
    }
    c->up_busy[b] = true;   // until its event completes, also if a call below fails
    HIPCHK(hipMemcpyAsync(dst + off, pin, len * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->up_ev[b], c->stream));
  }
  return CSG_OK;
}

int upload_real(csg_ctx* c, const double* host) {
  return upload(c, c->real, host, static_cast<size_t>(c->n) * c->n);
}

// Downloads to pageable host memory through the same pinned chunks, the
// copy engine filling one while the host empties the other (stream-ordered
// after the work queued before).  Returns once `dst` holds all `count` values.
int download(csg_ctx* c, double* dst, const double* src, size_t count) {
  if (int rc = ensure_staging(c)) return rc;
  for (int b = 0; b < 2; ++b)
    if (c->up_busy[b]) {
      HIPCHK(hipEventSynchronize(c->up_ev[b]));
      c->up_busy[b] = false;
    }
  constexpr size_t PER = UPLOAD_CHUNK / sizeof(double);
  const size_t chunks = (count + PER - 1) / PER;
  for (size_t k = 0; k <= chunks; ++k) {
    if (k < chunks) {
      const int b = static_cast<int>(k & 1);
      const size_t len = std::min(PER, count - k * PER);
      c->up_busy[b] = true;   // until its event completes, also if a call below fails
      HIPCHK(hipMemcpyAsync(c->up_pin[b], src + k * PER, len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipEventRecord(c->up_ev[b], c->stream));
    }
    if (k >= 1) {
      const int b = static_cast<int>((k - 1) & 1);
      const size_t len = std::min(PER, count - (k - 1) * PER);
      HIPCHK(hipEventSynchronize(c->up_ev[b]));
      c->up_busy[b] = false;
      std::memcpy(dst + (k - 1) * PER, c->up_pin[b], len * sizeof(double));
    }
  }
  return CSG_OK;
}

int set_state_from_real(csg_ctx* c, void* S) {
  PassArgs a = base_args(c);
  a.in0 = c->real;
  a.out0 = S;
  c->ops.fwd(a, c->tw64, c->stream);
  HIPCHK(hipGetLastError());
  return CSG_OK;
}

// Destroys the captured loop of csg_run (after a change of the kernels, the
// history buffer or the loop's arguments).
int drop_step_graphs(csg_ctx* c) {
  if (c->gexec) {
    const hipError_t e = hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("hipGraphExecDestroy: ") + hipGetErrorString(e));
  }
  return CSG_OK;
}

// History rows on the device: at least HIST_MIN_ROWS, so that short runs of
// different lengths keep one buffer (and the loops captured with it).
constexpr size_t HIST_MIN_ROWS = 1024;

int ensure_hist(csg_ctx* c, size_t rows) {
  if (rows <= c->hist_cap) return CSG_OK;
  rows = std::max(rows, HIST_MIN_ROWS);
  // the captured loops hold the old history pointer
  if (int rc = drop_step_graphs(c)) return rc;
  if (c->hist) {
    HIPCHK(hipFree(c->hist));
    c->bytes -= c->hist_cap * 6 * sizeof(double);
  }
  c->hist = nullptr;
  c->hist_cap = 0;
  int rc = dmalloc(c, reinterpret_cast<void**>(&c->hist), rows * 6 * sizeof(double));
  if (rc) return rc;
  c->hist_cap = rows;
  return CSG_OK;
}

__global__ void __launch_bounds__(64) reset_ctr_kernel(long long* ctr) {
  if (threadIdx.x < 2) ctr[threadIdx.x] = threadIdx.x == 0 ? 0 : -1;
}

// Step counter 0, no non-finite step yet (a kernel: no host copy to wait for).
int reset_ctr(csg_ctx* c) {
  hipLaunchKernelGGL(reset_ctr_kernel, dim3(1), dim3(64), 0, c->stream, c->ctr);
  HIPCHK(hipGetLastError());
  return CSG_OK;
}

bool same_params(const DevParams& x, const DevParams& y) { return std::memcmp(&x, &y, sizeof(DevParams)) == 0; }

// The captured block of STEPS_PER_GRAPH loop iterations for (dt, every) and
// the current parameters.
int build_graph(csg_ctx* c, double dt, int64_t every) {
  if (c->gexec && c->gdt == dt && c->gevery == every && same_params(c->gp, c->dp)) return CSG_OK;
  if (int rc = drop_step_graphs(c)) return rc;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < STEPS_PER_GRAPH; ++s) enqueue_iteration(c, dt, true, every);
  HIPCHK(hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    return fail(CSG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  c->gdt = dt;
  c->gevery = every;
  c->gp = c->dp;
  return CSG_OK;
}

// Host staging for ctr (2 words) and `rows` history rows, kept between runs.
// Pinned host slots and events for the non-finite checks of csg_run (made
// on first use; a context that never runs a long loop allocates none).
int ensure_flag_slots(csg_ctx* c) {
  if (c->flag_pin) return CSG_OK;
  for (hipEvent_t& ev : c->flag_ev)
    if (!ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  void* p = nullptr;
  HIPCHK(hipHostMalloc(&p, 2 * sizeof(long long), hipHostMallocDefault));
  c->flag_pin = static_cast<long long*>(p);
  return CSG_OK;
}

// Pinned host staging for ctr and the history rows of a run (at least
// HIST_MIN_ROWS rows), so the copies back are direct transfers.
int ensure_hstage(csg_ctx* c, size_t rows) {
  const size_t need = 2 + std::max(rows, HIST_MIN_ROWS) * 6;   // 8-byte words
  if (c->hstage_cap >= need) return CSG_OK;
  if (c->hstage) {
    HIPCHK(hipHostFree(c->hstage));
    c->hstage = nullptr;
    c->hstage_cap = 0;
  }
  void* p = nullptr;
  HIPCHK(hipHostMalloc(&p, need * sizeof(long long), hipHostMallocDefault));
  c->hstage = static_cast<long long*>(p);
  c->hstage_cap = need;
  return CSG_OK;
}

int check_ctx(const csg_ctx* c) {
  if (!c) return fail(CSG_ERR_ARG, "null context");
  return CSG_OK;
}

void drop_adjoint_graphs(csg_ctx* c);

}  // namespace

extern "C" {

void csg_default_params(csg_params* p) {
  if (!p) return;
  p->A = 2.0;
  p->B = 4.0;
  p->chi = 0.5;
  p->P = 5.0;
  p->ca = 0.5;
  p->cb = 0.0;
  p->cc = 1.0;
  p->kappa = 2.0;
  p->mobility = 1.0;
  p->Q = 0.5;
}

const char* csg_last_error(void) { return g_err.c_str(); }

int csg_create(csg_ctx** out, int32_t n, double spacing, const csg_params* p, int32_t precision,
               int32_t device) {
  if (!out) return fail(CSG_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!(spacing > 0.0) || !std::isfinite(spacing)) return fail(CSG_ERR_ARG, "spacing must be > 0");
  if (precision != CSG_FP64 && precision != CSG_MIXED) return fail(CSG_ERR_ARG, "unknown precision");
  csg_ctx* c = new csg_ctx();
  c->n = n;
  c->precision = precision;
  c->device = device;
  c->spacing = spacing;
  if (p) c->p = *p; else csg_default_params(&c->p);
  const bool ok = (precision == CSG_FP64) ? ops_for<double>(n, &c->ops) : ops_for<float>(n, &c->ops);
  if (!ok) {
    delete c;
    return fail(CSG_ERR_UNSUPPORTED, unsupported_size_message(n));
  }
  c->dp = DevParams{c->p.A, c->p.B, c->p.chi, c->p.P, c->p.ca, c->p.cb, c->p.cc,
                    c->p.kappa, c->p.mobility, c->p.Q, spacing, (1.0 / spacing) / n};
  c->tsize = (precision == CSG_FP64) ? sizeof(double) : sizeof(float);
  const size_t nn = static_cast<size_t>(n) * n;
  int rc;
  auto cleanup = [&](int code) { csg_destroy(c); return code; };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  if ((rc = dmalloc(c, &c->S, nn * sizeof(double)))) return cleanup(rc);
  if ((rc = dmalloc(c, &c->A, nn * c->tsize))) return cleanup(rc);
  if ((rc = dmalloc(c, &c->Bb, nn * c->tsize))) return cleanup(rc);
  if ((rc = dmalloc(c, &c->phi, nn * c->tsize))) return cleanup(rc);
  if ((rc = dmalloc(c, reinterpret_cast<void**>(&c->real), nn * sizeof(double)))) return cleanup(rc);
  const size_t npart = 5 * c->ops.max_row1 + c->ops.max_col + c->ops.max_row3;
  if ((rc = dmalloc(c, reinterpret_cast<void**>(&c->part), npart * sizeof(double)))) return cleanup(rc);
  if ((rc = dmalloc(c, reinterpret_cast<void**>(&c->ctr), 2 * sizeof(long long)))) return cleanup(rc);
  if ((rc = dmalloc(c, reinterpret_cast<void**>(&c->scratch_row), 8 * sizeof(double)))) return cleanup(rc);
  // twiddles exp(-2 pi i m / N) and the FFT stage tables
  const std::vector<cplx<double>> tw64 = make_twiddles<double>(n);
  if ((rc = dmalloc(c, &c->tw64, tw64.size() * sizeof(cplx<double>)))) return cleanup(rc);
  e = hipMemcpy(c->tw64, tw64.data(), tw64.size() * sizeof(cplx<double>), hipMemcpyHostToDevice);
  if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
  if (precision == CSG_FP64) {
    c->tw = c->tw64;
  } else {
    const std::vector<cplx<float>> tw32 = make_twiddles<float>(n);
    if ((rc = dmalloc(c, &c->tw, tw32.size() * sizeof(cplx<float>)))) return cleanup(rc);
    e = hipMemcpy(c->tw, tw32.data(), tw32.size() * sizeof(cplx<float>), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(fail(CSG_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
  }
  c->nt = nontemporal_for(c->bytes, false);
  *out = c;
  return CSG_OK;
}

int csg_destroy(csg_ctx* c) {
  if (!c) return CSG_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)drop_step_graphs(c);
  drop_adjoint_graphs(c);
  void* bufs[] = {c->S, c->S2, c->A, c->Bb, c->phi, c->real, c->part, c->ctr, c->hist,
                  c->scratch_row, c->rate, c->tw64, c->G, c->Xr, c->adj_acc};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (void* b : c->ckpt)
    if (b) (void)hipFree(b);
  for (void* b : c->seg)
    if (b) (void)hipFree(b);
  if (c->tw && c->tw != c->tw64) (void)hipFree(c->tw);
  if (c->flag_pin) (void)hipHostFree(c->flag_pin);
  if (c->hstage) (void)hipHostFree(c->hstage);
  for (hipEvent_t ev : c->flag_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (int b = 0; b < 2; ++b) {
    if (c->up_pin[b]) (void)hipHostFree(c->up_pin[b]);
    if (c->up_ev[b]) (void)hipEventDestroy(c->up_ev[b]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CSG_OK;
}

int64_t csg_device_bytes(const csg_ctx* c) { return c ? static_cast<int64_t>(c->bytes) : 0; }

int csg_set_params(csg_ctx* c, const csg_params* p) {
  if (int rc = check_ctx(c)) return rc;
  if (!p) return fail(CSG_ERR_ARG, "params is NULL");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->p = *p;
  c->dp = DevParams{p->A, p->B, p->chi, p->P, p->ca, p->cb, p->cc, p->kappa, p->mobility, p->Q,
                    c->spacing, (1.0 / c->spacing) / c->n};
  // captured graphs hold parameters in their kernel arguments: each records
  // the parameters it was captured with and is rebuilt when used with others
  c->adj_steps = -1;   // a checkpointed forward run belongs to the old parameters
  return CSG_OK;
}

int csg_set_fields(csg_ctx* c, const double* c0, const double* phi) {
  if (int rc = check_ctx(c)) return rc;
  if (!c0 || !phi) return fail(CSG_ERR_ARG, "c0 and phi must be non-NULL");
  HIPCHK(hipSetDevice(c->device));
  const size_t nn = static_cast<size_t>(c->n) * c->n;
  if (c->precision == CSG_FP64) {
    if (int rc = upload(c, static_cast<double*>(c->phi), phi, nn)) return rc;
  } else {
    if (int rc = upload(c, static_cast<float*>(c->phi), phi, nn)) return rc;
  }
  if (int rc = upload_real(c, c0)) return rc;
  if (int rc = set_state_from_real(c, c->S)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->fields_set = true;
  return CSG_OK;
}

int csg_run(csg_ctx* c, double dt, int64_t steps, int64_t hist_every, double* history) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->fields_set) return fail(CSG_ERR_STATE, "fields not set (csg_set_fields)");
  if (steps < 0 || hist_every < 1 || !std::isfinite(dt)) return fail(CSG_ERR_ARG, "bad steps/hist_every/dt");
  HIPCHK(hipSetDevice(c->device));
  // the device keeps rows s = 0, hist_every, 2 hist_every, ... only
  const int64_t rows = steps / hist_every + 1;
  if (int rc = ensure_hist(c, static_cast<size_t>(rows))) return rc;
  if (int rc = ensure_hstage(c, history ? static_cast<size_t>(rows) : 0)) return rc;
  if (int rc = reset_ctr(c)) return rc;
  int64_t done = 0;
  bool stopped = false;
  // Long runs replay a captured block of STEPS_PER_GRAPH iterations; the
  // remainder and the last, history-only iteration are launched directly.
  // (Replays are no faster than direct launches on MI355X, and the first
  // kernel of a directly launched run starts sooner; the graph saves host
  // work on long runs.)
  if (steps >= 2 * STEPS_PER_GRAPH) {
    if (int rc = build_graph(c, dt, hist_every)) return rc;
    // Long runs check the device's first-non-finite step between blocks of
    // CHECK_GRAPHS replays, one block behind the one being enqueued, so the
    // queue never drains and a run that went non-finite stops within two
    // blocks instead of computing every remaining step.
    const bool check = steps >= 2 * CHECK_GRAPHS * STEPS_PER_GRAPH;
    if (check) {
      if (int rc = ensure_flag_slots(c)) return rc;
      c->flag_pin[0] = c->flag_pin[1] = -1;
    }
    int64_t block = 0;
    while (steps - done >= STEPS_PER_GRAPH) {
      HIPCHK(hipGraphLaunch(c->gexec, c->stream));
      done += STEPS_PER_GRAPH;
      if (check && (done / STEPS_PER_GRAPH) % CHECK_GRAPHS == 0) {
        const int slot = static_cast<int>(block & 1);
        HIPCHK(hipMemcpyAsync(c->flag_pin + slot, c->ctr + 1, sizeof(long long), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipEventRecord(c->flag_ev[slot], c->stream));
        if (block > 0) {
          HIPCHK(hipEventSynchronize(c->flag_ev[slot ^ 1]));
          if (c->flag_pin[slot ^ 1] >= 0) {
            stopped = true;
            break;
          }
        }
        ++block;
      }
    }
  }
  if (!stopped) {
    for (; done < steps; ++done) enqueue_iteration(c, dt, true, hist_every);
    enqueue_iteration(c, dt, false, hist_every);
  }
  HIPCHK(hipGetLastError());
  long long* ctr = c->hstage;
  double* h = reinterpret_cast<double*>(c->hstage + 2);
  HIPCHK(hipMemcpyAsync(ctr, c->ctr, 2 * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
  if (history)
    HIPCHK(hipMemcpyAsync(h, c->hist, static_cast<size_t>(rows) * 6 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (history) {
    for (int64_t r = 0; r < rows; ++r)
      for (int col = 0; col < 6; ++col) history[r + col * rows] = h[static_cast<size_t>(r) * 6 + col];
  }
  c->nonfinite = ctr[1];
  if (ctr[1] >= 0)
    return fail(CSG_ERR_NONFINITE, "Non-finite composition at step " + std::to_string(ctr[1]));
  return CSG_OK;
}

int csg_step(csg_ctx* c, double dt, int64_t nsteps) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->fields_set) return fail(CSG_ERR_STATE, "fields not set (csg_set_fields)");
  if (nsteps < 0 || !std::isfinite(dt)) return fail(CSG_ERR_ARG, "bad nsteps/dt");
  HIPCHK(hipSetDevice(c->device));
  for (int64_t s = 0; s < nsteps; ++s) {
    enqueue_p123(c, c->S);
    enqueue_p4(c, c->S, dt, nullptr);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return CSG_OK;
}

int csg_get_field(csg_ctx* c, double* out) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->fields_set) return fail(CSG_ERR_STATE, "fields not set (csg_set_fields)");
  if (!out) return fail(CSG_ERR_ARG, "out is NULL");
  HIPCHK(hipSetDevice(c->device));
  PassArgs a = base_args(c);
  a.in0 = c->S;
  a.out0 = c->real;
  c->ops.inv_state(a, c->tw64, c->stream);
  HIPCHK(hipGetLastError());
  if (int rc = download(c, out, c->real, static_cast<size_t>(c->n) * c->n)) return rc;
  return CSG_OK;
}

int64_t csg_nonfinite_step(const csg_ctx* c) { return c ? c->nonfinite : -1; }

}  // extern "C"

namespace {

int prepare_single(csg_ctx* c, const double* cin) {
  if (int rc = check_ctx(c)) return rc;
  if (!cin) return fail(CSG_ERR_ARG, "c is NULL");
  if (!c->fields_set) return fail(CSG_ERR_STATE, "phi not set (csg_set_fields)");
  HIPCHK(hipSetDevice(c->device));
  if (!c->S2) {
    if (int rc = dmalloc(c, &c->S2, static_cast<size_t>(c->n) * c->n * sizeof(double))) return rc;
  }
  if (int rc = upload_real(c, cin)) return rc;
  return set_state_from_real(c, c->S2);
}

// History-style reduction of P1..P3 on S2 into scratch_row.
int single_reduce(csg_ctx* c, double out[6]) {
  enqueue_p123(c, c->S2);
  if (int rc = reset_ctr(c)) return rc;
  enqueue_finalize(c, c->scratch_row, 0.0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->scratch_row, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return CSG_OK;
}

}  // namespace

extern "C" {

int csg_energy(csg_ctx* c, const double* cin, double* F) {
  if (int rc = prepare_single(c, cin)) return rc;
  double row[6];
  if (int rc = single_reduce(c, row)) return rc;
  if (F) *F = row[2];
  return CSG_OK;
}

int csg_chemical_potential(csg_ctx* c, const double* cin, double* mu) {
  if (int rc = prepare_single(c, cin)) return rc;
  if (!mu) return fail(CSG_ERR_ARG, "mu is NULL");
  PassArgs a = base_args(c);
  a.in0 = c->S2;
  a.out0 = c->A;
  a.part = part1(c);
  c->ops.p1(a, c->tw, c->stream);
  a = base_args(c);
  a.in0 = c->S2;
  a.in1 = c->A;
  a.out0 = c->Bb;
  c->ops.mu(a, c->tw, c->stream);
  a = base_args(c);
  a.in0 = c->Bb;
  a.out0 = c->real;
  c->ops.inv_bulk(a, c->tw, c->stream);
  HIPCHK(hipGetLastError());
  return download(c, mu, c->real, static_cast<size_t>(c->n) * c->n);
}

int csg_transport_rate(csg_ctx* c, const double* cin, double* rate_re, double* rate_im,
                       double* dissipation) {
  if (int rc = prepare_single(c, cin)) return rc;
  const int n = c->n;
  const size_t half = static_cast<size_t>(n / 2 + 1) * n;
  if (!c->rate) {
    if (int rc = dmalloc(c, reinterpret_cast<void**>(&c->rate), half * 2 * sizeof(double))) return rc;
  }
  double row[6];
  if (int rc = single_reduce(c, row)) return rc;   // leaves Jx/Jy in A/Bb
  if (dissipation) *dissipation = row[5];
  if (!rate_re && !rate_im) return CSG_OK;
  PassArgs a = base_args(c);
  a.in0 = c->A;
  a.in1 = c->Bb;
  a.out0 = c->rate;
  c->ops.rate(a, c->tw, c->stream);
  HIPCHK(hipGetLastError());
  std::vector<double> h(half * 2);
  HIPCHK(hipMemcpyAsync(h.data(), c->rate, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // h: line kx (0..N/2) of N values over ky.  Full spectrum (column-major,
  // element (kx, ky) at kx + ky*N) by Hermitian symmetry for kx > N/2.
  for (int ky = 0; ky < n; ++ky) {
    for (int kx = 0; kx < n; ++kx) {
      double re, im;
      if (kx <= n / 2) {
        re = h[2 * (static_cast<size_t>(kx) * n + ky)];
        im = h[2 * (static_cast<size_t>(kx) * n + ky) + 1];
      } else {
        const int mx = n - kx, my = (n - ky) % n;
        re = h[2 * (static_cast<size_t>(mx) * n + my)];
        im = -h[2 * (static_cast<size_t>(mx) * n + my) + 1];
      }
      const size_t o = static_cast<size_t>(kx) + static_cast<size_t>(ky) * n;
      if (rate_re) rate_re[o] = re;
      if (rate_im) rate_im[o] = im;
    }
  }
  return CSG_OK;
}

}  // extern "C"

namespace {

// Pointwise single-call evaluations over the N*N real field.
template <int MODE>
int pointwise(csg_ctx* c, const double* cin, double* o0, double* o1, double* o2) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->fields_set) return fail(CSG_ERR_STATE, "phi not set (csg_set_fields)");
  HIPCHK(hipSetDevice(c->device));
  const long long nn = static_cast<long long>(c->n) * c->n;
  const size_t bytes = static_cast<size_t>(nn) * sizeof(double);
  double* outs[3] = {o0, o1, o2};
  const int nout = (MODE == PT_GEOMETRY) ? 3 : 1;
  // device staging: outputs in A/Bb/S2 regions sized >= N*N doubles
  if (!c->S2) {
    if (int rc = dmalloc(c, &c->S2, bytes)) return rc;
  }
  double* d_out[3] = {c->real, static_cast<double*>(c->S2), static_cast<double*>(c->S)};
  void* d_tmp = nullptr;
  if (MODE == PT_GEOMETRY) {
    // S holds the resident state: use a temporary for the third output
    HIPCHK(hipMalloc(&d_tmp, bytes));
    d_out[2] = static_cast<double*>(d_tmp);
  } else {
    HIPCHK(hipMemcpyAsync(c->real, cin, bytes, hipMemcpyHostToDevice, c->stream));
    d_out[0] = static_cast<double*>(c->S2);
  }
  const dim3 grid(static_cast<unsigned>((nn + 255) / 256));
  if (c->precision == CSG_FP64)
    hipLaunchKernelGGL((pointwise_kernel<double, MODE>), grid, dim3(256), 0, c->stream, c->real,
                       static_cast<const double*>(c->phi), d_out[0], d_out[1], d_out[2], nn, c->dp);
  else
    hipLaunchKernelGGL((pointwise_kernel<float, MODE>), grid, dim3(256), 0, c->stream, c->real,
                       static_cast<const float*>(c->phi), d_out[0], d_out[1], d_out[2], nn, c->dp);
  hipError_t e = hipGetLastError();
  for (int k = 0; k < nout && e == hipSuccess; ++k)
    if (outs[k]) e = hipMemcpyAsync(outs[k], d_out[k], bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_tmp) (void)hipFree(d_tmp);
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("pointwise: ") + hipGetErrorString(e));
  return CSG_OK;
}

}  // namespace

extern "C" {

int csg_geometry(csg_ctx* c, double* h, double* g, double* M) {
  return pointwise<PT_GEOMETRY>(c, nullptr, h, g, M);
}

int csg_bulk_energy(csg_ctx* c, const double* cin, double* f) {
  if (!cin || !f) return fail(CSG_ERR_ARG, "c and f must be non-NULL");
  return pointwise<PT_BULK_ENERGY>(c, cin, f, nullptr, nullptr);
}

int csg_bulk_derivative(csg_ctx* c, const double* cin, double* fp) {
  if (!cin || !fp) return fail(CSG_ERR_ARG, "c and fp must be non-NULL");
  return pointwise<PT_BULK_DERIVATIVE>(c, cin, fp, nullptr, nullptr);
}

int csg_time_passes(csg_ctx* c, double dt, int64_t nsteps, double* ms) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->fields_set) return fail(CSG_ERR_STATE, "fields not set (csg_set_fields)");
  if (nsteps < 1 || !ms) return fail(CSG_ERR_ARG, "nsteps >= 1 and ms required");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = ensure_hist(c, static_cast<size_t>(nsteps) + 1)) return rc;
  if (int rc = reset_ctr(c)) return rc;
  // loop iterations in the order of csg_run, with timing events (no system
  // fence) recorded between the kernels on the context's stream
  // all iterations are enqueued before the first wait, so no host submission
  // gap falls inside an interval
  std::vector<hipEvent_t> ev(static_cast<size_t>(4 * nsteps + 3));
  int made = 0;
  hipError_t e = hipSuccess;
  for (auto& x : ev) {
    e = hipEventCreateWithFlags(&x, hipEventDisableSystemFence);
    if (e != hipSuccess) break;
    ++made;
  }
  auto destroy = [&]() { for (int i = 0; i < made; ++i) (void)hipEventDestroy(ev[i]); };
  if (e != hipSuccess) { destroy(); return fail(CSG_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e)); }
  size_t q = 0;
  auto rec = [&]() { if (e == hipSuccess) e = hipEventRecord(ev[q++], c->stream); };
  rec();
  for (int64_t s = 0; s < nsteps; ++s) {
    PassArgs a = base_args(c);
    a.in0 = c->S; a.out0 = c->A; a.part = part1(c);
    c->ops.p1(a, c->tw, c->stream);
    rec();
    a = base_args(c);
    a.in0 = c->S; a.in1 = c->A; a.out0 = c->A; a.out1 = c->Bb; a.part = part2(c);
    c->ops.p2(a, c->tw, c->stream);
    rec();
    a = base_args(c);
    a.in0 = c->A; a.in1 = c->Bb; a.out0 = c->A; a.out1 = c->Bb; a.part = part3(c);
    c->ops.p3(a, c->tw, c->stream);
    rec();
    enqueue_p4(c, c->S, dt, c->hist);
    rec();
  }
  // standalone history reduction (a run's last row), on a fresh counter
  if (e == hipSuccess) e = hipMemsetAsync(c->ctr, 0, sizeof(long long), c->stream);
  rec();
  enqueue_finalize(c, c->hist, dt);
  rec();
  if (e == hipSuccess) e = hipEventSynchronize(ev[q - 1]);
  double acc[5] = {0, 0, 0, 0, 0};
  for (int64_t s = 0; s < nsteps && e == hipSuccess; ++s)
    for (int k = 0; k < 4 && e == hipSuccess; ++k) {
      float t = 0.f;
      e = hipEventElapsedTime(&t, ev[4 * s + k], ev[4 * s + k + 1]);
      acc[k] += t;
    }
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[q - 2], ev[q - 1]);
  destroy();
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("csg_time_passes: ") + hipGetErrorString(e));
  for (int k = 0; k < 4; ++k) ms[k] = acc[k] / static_cast<double>(nsteps);
  ms[4] = t;
  return CSG_OK;
}

}  // extern "C"

namespace {

// Streaming copy of n 16-byte words: a grid-stride loop with four words per
// thread in flight, every index checked against n; NT: non-temporal loads and
// stores, for buffers that stream through HBM.
template <bool NT>
__global__ void __launch_bounds__(256) copy_words(const double2* __restrict__ src, double2* __restrict__ dst,
                                                  size_t n) {
  typedef double w2 __attribute__((ext_vector_type(2)));
  const w2* s = reinterpret_cast<const w2*>(src);
  w2* d = reinterpret_cast<w2*>(dst);
  const size_t T = static_cast<size_t>(gridDim.x) * 256;
  size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  for (; i + 3 * T < n; i += 4 * T) {
    w2 w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = NT ? __builtin_nontemporal_load(s + i + u * T) : s[i + u * T];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (NT)
        __builtin_nontemporal_store(w[u], d + i + u * T);
      else
        d[i + u * T] = w[u];
    }
  }
  for (; i < n; i += T) d[i] = s[i];
}

// Launch shape measured on MI355X (grid x unroll x hint sweep): buffers beyond
// the Infinity Cache copy fastest non-temporally with 64 workgroups per CU
// (1 GiB: 5.3 TB/s), cache-resident ones plainly with 16 per CU (32 MiB: 7.2 TB/s).
constexpr int64_t COPY_MAX_BYTES = int64_t(16) << 30;
constexpr int64_t COPY_STREAM_BYTES = int64_t(256) << 20;
}  // namespace

extern "C" {

int csg_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double* gbs) {
  if (!gbs || bytes < 16 || bytes % 16 != 0 || bytes > COPY_MAX_BYTES || reps < 1)
    return fail(CSG_ERR_ARG, "copy: bytes must be a multiple of 16 in [16, 16 GiB], reps >= 1, gbs non-NULL");
  HIPCHK(hipSetDevice(device));
  void *src = nullptr, *dst = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const size_t nb = static_cast<size_t>(bytes), n = nb / 16;
  const bool stream = bytes > COPY_STREAM_BYTES;
  const unsigned grid = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, stream ? 16384 : 4096));
  float ms = 0.0f;
  hipError_t e = hipMalloc(&src, nb);
  if (e == hipSuccess) e = hipMalloc(&dst, nb);
  if (e == hipSuccess) e = hipStreamCreate(&s);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(src, 0, nb, s);
  // three untimed copies, then reps timed ones
  for (int r = 0; e == hipSuccess && r < 3 + reps; ++r) {
    if (r == 3) e = hipEventRecord(e0, s);
    if (e != hipSuccess) break;
    if (stream)
      hipLaunchKernelGGL(copy_words<true>, dim3(grid), dim3(256), 0, s, static_cast<const double2*>(src),
                         static_cast<double2*>(dst), n);
    else
      hipLaunchKernelGGL(copy_words<false>, dim3(grid), dim3(256), 0, s, static_cast<const double2*>(src),
                         static_cast<double2*>(dst), n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (s) (void)hipStreamDestroy(s);
  if (dst) (void)hipFree(dst);
  if (src) (void)hipFree(src);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(CSG_ERR_HIP, std::string("csg_copy_bandwidth: ") + hipGetErrorString(e));
  }
  if (!(ms > 0.0f)) return fail(CSG_ERR_HIP, "csg_copy_bandwidth: the timed copies measured no elapsed time");
  *gbs = 2.0 * static_cast<double>(nb) * reps / (static_cast<double>(ms) * 1e-3) / 1e9;
  return CSG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint: checkpointed forward run and backward sweep (FP64 contexts).
namespace {

void free_back(csg_ctx* c, std::vector<void*>& v, size_t keep, size_t bytes) {
  while (v.size() > keep) {
    (void)hipFree(v.back());
    c->bytes -= bytes;
    v.pop_back();
  }
}

// Exactly `count` buffers of `bytes` in v.  On an allocation failure the
// buffers this call allocated are freed again, so a caller can retry with
// fewer (keeping every state -> checkpoints).  Callers drop the captured
// adjoint graphs before the count can shrink.
int ensure_bufs(csg_ctx* c, std::vector<void*>& v, size_t count, size_t bytes) {
  const size_t had = v.size();
  if (had > count) {
    free_back(c, v, count, bytes);
    return CSG_OK;
  }
  while (v.size() < count) {
    void* p = nullptr;
    if (int rc = dmalloc(c, &p, bytes)) {
      free_back(c, v, had, bytes);
      return rc;
    }
    v.push_back(p);
  }
  return CSG_OK;
}

// The backward sweep's work buffers: adjoint field, real-space scratch,
// parameter-term partials.
int ensure_adjoint_work(csg_ctx* c) {
  const size_t bytes = static_cast<size_t>(c->n) * c->n * sizeof(double);
  if (!c->G)
    if (int rc = dmalloc(c, &c->G, bytes)) return rc;
  if (!c->Xr)
    if (int rc = dmalloc(c, &c->Xr, bytes)) return rc;
  const size_t nacc = 2 * static_cast<size_t>(c->ops.n_adj_col) + c->ops.n_adj_row;
  if (!c->adj_acc)
    if (int rc = dmalloc(c, reinterpret_cast<void**>(&c->adj_acc), nacc * sizeof(double))) return rc;
  return CSG_OK;
}

// Segment buffers for checkpoint interval K (none when every state is kept).
int ensure_segments(csg_ctx* c, int64_t K) {
  const size_t bytes = static_cast<size_t>(c->n) * c->n * sizeof(double);
  if (K > 1) {
    // the recompute and backward graphs bind the segment buffers
    if (c->seg.size() != static_cast<size_t>(K) + 1) drop_adjoint_graphs(c);
    return ensure_bufs(c, c->seg, static_cast<size_t>(K) + 1, bytes);
  }
  // every state kept: no segments; a change of K has dropped the graphs that bound them
  free_back(c, c->seg, 0, bytes);
  return CSG_OK;
}

// One step without history: Sout = step(Sin) (in place when equal).
void enqueue_plain_step(csg_ctx* c, const void* Sin, void* Sout, double dt) {
  enqueue_p123(c, Sin);
  enqueue_p4(c, Sout, dt, nullptr, Sin == Sout ? nullptr : Sin);
}

void drop_adjoint_graphs(csg_ctx* c) {
  for (hipGraphExec_t* g : {&c->adj_fwd, &c->adj_rec, &c->adj_bwd}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
  for (std::vector<hipGraphExec_t>* v : {&c->adj_all_fwd, &c->adj_all_bwd}) {
    for (hipGraphExec_t g : *v)
      if (g) (void)hipGraphExecDestroy(g);
    v->clear();
  }
  c->adj_all_T = 0;
  c->adj_gK = 0;
}

// Capture what `body` enqueues on the context's stream into an executable graph.
template <class F>
int capture_graph(csg_ctx* c, hipGraphExec_t* out, F&& body) {
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  body();
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return fail(CSG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    *out = nullptr;
    return fail(CSG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return CSG_OK;
}

AdjArgs adjoint_args(const csg_ctx* c) {
  AdjArgs a{};
  a.G = static_cast<cplx<double>*>(c->G);
  a.X0 = static_cast<cplx<double>*>(c->A);
  a.X1 = static_cast<cplx<double>*>(c->Bb);
  a.Xr = static_cast<cplx<double>*>(c->Xr);
  a.phi = static_cast<const double*>(c->phi);
  a.dt = c->adj_dt;
  a.p = c->dp;
  return a;
}

// Backward step through the forward step S0 -> S1: gbar (c->G) becomes
// dL/dS0, the parameter terms accumulate in c->adj_acc.
void enqueue_backward_step(csg_ctx* c, const void* S0, const void* S1) {
  AdjArgs a = adjoint_args(c);
  const int na = c->ops.n_adj_col;
  a.S0 = static_cast<const cplx<double>*>(S0);
  a.S1 = static_cast<const cplx<double>*>(S1);
  a.acc = c->adj_acc;
  c->ops.adj1(a, c->tw, c->stream);
  PassArgs pa = base_args(c);
  pa.in0 = c->A; pa.in1 = c->Bb; pa.out0 = c->A; pa.out1 = c->Bb; pa.part = part3(c);
  c->ops.p3(pa, c->tw, c->stream);
  a.acc = c->adj_acc + na;
  c->ops.adj3(a, c->tw, c->stream);
  a.acc = c->adj_acc + 2 * na;
  c->ops.adj4(a, c->tw, c->stream);
}

// Backward sweep over steps m-1 .. 0 of a segment whose states are seg[0..m].
void enqueue_backward_segment(csg_ctx* c, int64_t m) {
  for (int64_t i = m - 1; i >= 0; --i) enqueue_backward_step(c, c->seg[i], c->seg[i + 1]);
}

// Recompute seg[1..m] from seg[0].
void enqueue_recompute_segment(csg_ctx* c, int64_t m) {
  for (int64_t i = 0; i < m; ++i) enqueue_plain_step(c, c->seg[i], c->seg[i + 1], c->adj_dt);
}

}  // namespace

extern "C" {

int csg_adjoint_forward(csg_ctx* c, double dt, int64_t steps, int64_t ckpt_every) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->fields_set) return fail(CSG_ERR_STATE, "fields not set (csg_set_fields)");
  if (!c->ops.adj1) return fail(CSG_ERR_UNSUPPORTED, "the adjoint needs an FP64 context");
  if (steps < 1 || ckpt_every < 1 || !std::isfinite(dt)) return fail(CSG_ERR_ARG, "bad steps/ckpt_every/dt");
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = static_cast<size_t>(c->n) * c->n * sizeof(double);
  const int64_t K = ckpt_every, nfull = steps / K;
  const size_t nck = static_cast<size_t>(nfull) + 1;
  // captured graphs bind the checkpoint buffers: drop them before any change
  if (c->adj_gK != K || c->adj_gdt != dt || !same_params(c->adj_gp, c->dp) || (K == 1 && c->adj_all_T != steps) ||
      c->ckpt.size() != nck)
    drop_adjoint_graphs(c);
  c->adj_steps = -1;   // no valid forward run until this one completes
  // everything the backward sweep needs is allocated here, so a run that does
  // not fit fails now (and can be retried with checkpoints) rather than there
  if (int rc = ensure_adjoint_work(c)) return rc;
  if (int rc = ensure_segments(c, K)) return rc;
  if (int rc = ensure_bufs(c, c->ckpt, nck, bytes)) return rc;
  c->adj_dt = dt;
  if (K == 1) {
    // every state kept: each step writes its result straight into the next
    // checkpoint (out-of-place P4), so the backward sweep needs no recompute
    const int64_t nb = (steps + STEPS_PER_GRAPH - 1) / STEPS_PER_GRAPH;
    if (c->adj_all_fwd.empty()) {
      for (int64_t bk = 0; bk < nb; ++bk) {
        hipGraphExec_t g = nullptr;
        const int64_t s0 = bk * STEPS_PER_GRAPH, s1 = std::min<int64_t>(s0 + STEPS_PER_GRAPH, steps);
        if (int rc = capture_graph(c, &g, [&]() {
              for (int64_t s = s0; s < s1; ++s) enqueue_plain_step(c, c->ckpt[s], c->ckpt[s + 1], dt);
            }))
          return rc;
        c->adj_all_fwd.push_back(g);
      }
      c->adj_all_T = steps;
      c->adj_gK = 1;
      c->adj_gdt = dt;
      c->adj_gp = c->dp;
    }
    HIPCHK(hipMemcpyAsync(c->ckpt[0], c->S, bytes, hipMemcpyDeviceToDevice, c->stream));
    for (hipGraphExec_t g : c->adj_all_fwd) HIPCHK(hipGraphLaunch(g, c->stream));
    HIPCHK(hipMemcpyAsync(c->S, c->ckpt[steps], bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->adj_steps = steps;
    c->adj_every = 1;
    return CSG_OK;
  }
  // full segments replay one captured graph of K steps when there are several
  if (nfull >= 2 && !c->adj_fwd) {
    if (int rc = capture_graph(c, &c->adj_fwd, [&]() {
          for (int64_t s = 0; s < K; ++s) enqueue_plain_step(c, c->S, c->S, dt);
        }))
      return rc;
    c->adj_gK = K;
    c->adj_gdt = dt;
    c->adj_gp = c->dp;
  }
  HIPCHK(hipMemcpyAsync(c->ckpt[0], c->S, bytes, hipMemcpyDeviceToDevice, c->stream));
  for (int64_t k = 0; k < nfull; ++k) {
    if (c->adj_fwd) {
      HIPCHK(hipGraphLaunch(c->adj_fwd, c->stream));
    } else {
      for (int64_t s = 0; s < K; ++s) enqueue_plain_step(c, c->S, c->S, dt);
    }
    HIPCHK(hipMemcpyAsync(c->ckpt[k + 1], c->S, bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  for (int64_t s = nfull * K; s < steps; ++s) enqueue_plain_step(c, c->S, c->S, dt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->adj_steps = steps;
  c->adj_every = ckpt_every;
  return CSG_OK;
}

int csg_adjoint_backward(csg_ctx* c, const double* dL_dcT, double* dL_dc0, double* dL_dparams) {
  if (int rc = check_ctx(c)) return rc;
  if (c->adj_steps < 1) return fail(CSG_ERR_STATE, "no checkpointed forward run (csg_adjoint_forward)");
  if (!dL_dcT) return fail(CSG_ERR_ARG, "dL_dcT is NULL");
  HIPCHK(hipSetDevice(c->device));
  const int64_t T = c->adj_steps, K = c->adj_every;
  const double dt = c->adj_dt;
  const size_t bytes = static_cast<size_t>(c->n) * c->n * sizeof(double);
  if (int rc = ensure_adjoint_work(c)) return rc;   // allocated by the forward run already
  if (int rc = ensure_segments(c, K)) return rc;
  const int na = c->ops.n_adj_col, nr = c->ops.n_adj_row;
  const size_t nacc = 2 * static_cast<size_t>(na) + nr;
  HIPCHK(hipMemsetAsync(c->adj_acc, 0, nacc * sizeof(double), c->stream));
  // gbar_T into the state's packed line layout
  if (int rc = upload_real(c, dL_dcT)) return rc;
  if (int rc = set_state_from_real(c, c->G)) return rc;
  const int64_t nseg = (T + K - 1) / K, nfull = T / K;
  (void)dt;
  if (K == 1) {
    // every state kept: step i's adjoint reads checkpoints i and i + 1.  The
    // block graphs bind those buffers, dt and the parameters; any change of
    // them, of the schedule or of T drops them (drop_adjoint_graphs), so
    // graphs that exist here were captured for this run.
    const int64_t nb = (T + STEPS_PER_GRAPH - 1) / STEPS_PER_GRAPH;
    if (!c->adj_all_bwd.empty() && (c->adj_all_T != T || c->adj_gdt != c->adj_dt || !same_params(c->adj_gp, c->dp)))
      drop_adjoint_graphs(c);
    if (c->adj_all_bwd.empty()) {
      c->adj_all_T = T;
      c->adj_gK = 1;
      c->adj_gdt = c->adj_dt;
      c->adj_gp = c->dp;
      for (int64_t bk = 0; bk < nb; ++bk) {
        hipGraphExec_t g = nullptr;
        const int64_t s0 = bk * STEPS_PER_GRAPH, s1 = std::min<int64_t>(s0 + STEPS_PER_GRAPH, T);
        if (int rc = capture_graph(c, &g, [&]() {
              for (int64_t i = s1 - 1; i >= s0; --i) enqueue_backward_step(c, c->ckpt[i], c->ckpt[i + 1]);
            }))
          return rc;
        c->adj_all_bwd.push_back(g);
      }
    }
    for (int64_t bk = nb - 1; bk >= 0; --bk) HIPCHK(hipGraphLaunch(c->adj_all_bwd[bk], c->stream));
  } else {
  // full segments replay captured recompute and backward graphs when there are several
  if (c->adj_gK && (c->adj_gK != K || c->adj_gdt != c->adj_dt || !same_params(c->adj_gp, c->dp)))
    drop_adjoint_graphs(c);
  if (nfull >= 2 && (!c->adj_rec || !c->adj_bwd)) {
    if (!c->adj_rec) {
      if (int rc = capture_graph(c, &c->adj_rec, [&]() { enqueue_recompute_segment(c, K); })) return rc;
    }
    if (!c->adj_bwd) {
      if (int rc = capture_graph(c, &c->adj_bwd, [&]() { enqueue_backward_segment(c, K); })) return rc;
    }
    c->adj_gK = K;
    c->adj_gdt = c->adj_dt;
    c->adj_gp = c->dp;
  }
  for (int64_t sg = nseg - 1; sg >= 0; --sg) {
    const int64_t s0 = sg * K, s1 = std::min<int64_t>(s0 + K, T), m = s1 - s0;
    // recompute the segment's states from its checkpoint, then sweep back
    HIPCHK(hipMemcpyAsync(c->seg[0], c->ckpt[sg], bytes, hipMemcpyDeviceToDevice, c->stream));
    if (m == K && c->adj_rec && c->adj_bwd) {
      HIPCHK(hipGraphLaunch(c->adj_rec, c->stream));
      HIPCHK(hipGraphLaunch(c->adj_bwd, c->stream));
    } else {
      enqueue_recompute_segment(c, m);
      enqueue_backward_segment(c, m);
    }
  }
  }
  HIPCHK(hipGetLastError());
  std::vector<double> acc(nacc);
  HIPCHK(hipMemcpyAsync(acc.data(), c->adj_acc, nacc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (dL_dc0) {
    PassArgs pa = base_args(c);
    pa.in0 = c->G;
    pa.out0 = c->real;
    c->ops.inv_state(pa, c->tw64, c->stream);
    if (int rc = download(c, dL_dc0, c->real, bytes / sizeof(double))) return rc;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dL_dparams) {
    double kap = 0.0, srg = 0.0;
    for (int i = 0; i < 2 * na; ++i) kap += acc[i];
    for (int i = 0; i < nr; ++i) srg += acc[2 * na + i];
    dL_dparams[0] = kap;
    dL_dparams[1] = -c->p.chi * srg;
    dL_dparams[2] = -c->p.P * srg;
  }
  return CSG_OK;
}

}  // extern "C"

// ---- kernel schedules and the runtime tuner ----------------------------------

namespace {

constexpr int TUNE_REPS = 4;   // timed graph replays (32 loop iterations each) per candidate

// Mean device time per loop iteration of the captured loop with schedules
// sel[0..3]; the history counter is reset so the rows stay in the buffer.
int time_schedule(csg_ctx* c, double dt, const int sel[4], hipEvent_t e0, hipEvent_t e1, double* us) {
  select_schedule(c->ops, sel);
  if (int rc = drop_step_graphs(c)) return rc;
  if (int rc = build_graph(c, dt, 1)) return rc;
  if (int rc = reset_ctr(c)) return rc;
  HIPCHK(hipGraphLaunch(c->gexec, c->stream));   // warm-up
  if (int rc = reset_ctr(c)) return rc;
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < TUNE_REPS; ++r) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *us = 1e3 * ms / (TUNE_REPS * STEPS_PER_GRAPH);
  return CSG_OK;
}

}  // namespace

extern "C" {

int csg_schedule_info(const csg_ctx* c, int32_t pass, int32_t* count, int32_t* sched) {
  if (!c) return fail(CSG_ERR_ARG, "null context");
  if (pass < 0 || pass > 3 || !count) return fail(CSG_ERR_ARG, "pass must be 0..3 and count non-NULL");
  const PassCands& pc = c->ops.cand[pass];
  *count = pc.n;
  if (sched)
    for (int i = 0; i < pc.n; ++i)
      for (int q = 0; q < SCHED_FIELDS; ++q) sched[SCHED_FIELDS * i + q] = pc.sched[i][q];
  return CSG_OK;
}

int csg_get_schedule(const csg_ctx* c, int32_t* sel) {
  if (!c || !sel) return fail(CSG_ERR_ARG, "null context or sel");
  for (int q = 0; q < 4; ++q) sel[q] = c->ops.sel[q];
  return CSG_OK;
}

int csg_set_schedule(csg_ctx* c, const int32_t* sel) {
  if (!c || !sel) return fail(CSG_ERR_ARG, "null context or sel");
  const int s[4] = {sel[0], sel[1], sel[2], sel[3]};
  bool same = true;
  for (int q = 0; q < 4; ++q) same = same && s[q] == c->ops.sel[q];
  if (same) return CSG_OK;
  if (!select_schedule(c->ops, s)) return fail(CSG_ERR_ARG, "schedule index out of range (csg_schedule_info)");
  HIPCHK(hipSetDevice(c->device));
  // the captured loops hold the previous kernels
  if (int rc = drop_step_graphs(c)) return rc;
  drop_adjoint_graphs(c);
  return CSG_OK;
}

int csg_autotune(csg_ctx* c, double dt, int32_t rounds, int32_t* sel_out, double* us_per_step) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->fields_set) return fail(CSG_ERR_STATE, "fields not set (csg_set_fields)");
  if (rounds < 1 || !std::isfinite(dt)) return fail(CSG_ERR_ARG, "rounds >= 1 and finite dt required");
  HIPCHK(hipSetDevice(c->device));
  const size_t nn = static_cast<size_t>(c->n) * c->n;
  if (!c->S2) {
    if (int rc = dmalloc(c, &c->S2, nn * sizeof(double))) return rc;
  }
  // the resident field is kept and restored afterwards
  HIPCHK(hipMemcpyAsync(c->S2, c->S, nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (int rc = ensure_hist(c, static_cast<size_t>(STEPS_PER_GRAPH) * TUNE_REPS + 1)) return rc;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int best[4] = {c->ops.sel[0], c->ops.sel[1], c->ops.sel[2], c->ops.sel[3]};
  double best_us = 0.0;
  int rc = time_schedule(c, dt, best, e0, e1, &best_us);
  // coordinate descent over the four passes
  for (int r = 0; r < rounds && rc == CSG_OK; ++r) {
    for (int q = 0; q < 4 && rc == CSG_OK; ++q) {
      for (int i = 0; i < c->ops.cand[q].n && rc == CSG_OK; ++i) {
        if (i == best[q]) continue;
        int trial[4] = {best[0], best[1], best[2], best[3]};
        trial[q] = i;
        double us = 0.0;
        rc = time_schedule(c, dt, trial, e0, e1, &us);
        if (rc == CSG_OK && us < best_us) {
          best_us = us;
          best[q] = i;
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  select_schedule(c->ops, best);
  (void)drop_step_graphs(c);
  drop_adjoint_graphs(c);
  HIPCHK(hipMemcpyAsync(c->S, c->S2, nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rc) return rc;
  if (sel_out)
    for (int q = 0; q < 4; ++q) sel_out[q] = best[q];
  if (us_per_step) *us_per_step = best_us;
  return CSG_OK;
}

}  // extern "C"
