/*
 * csg.h -- C ABI of the MI355X (gfx950) core-shell Cahn-Hilliard step library.
 *
 * Replaces the time-step loop of core_shell.jl (run_simulation, :84-103) and
 * the functions it calls (step :72-76, transport_rate :62-70,
 * chemical_potential :57-60, energy :78-82, bulk_derivative :51-55,
 * bulk_energy :45-49, geometry :39-43, make_grid :17-25) with HIP kernels on
 * one GPU.  The reference exposes these as plain Julia functions; a Julia
 * caller binds the entry points below with `ccall` (the host package shipped
 * beside this library mirrors the same API in Python through ctypes).
 *
 * Conventions
 *   - Every entry point returns a csg_status (0 = CSG_OK).  No C++ exception
 *     crosses the ABI.  csg_last_error() returns a thread-local message.
 *   - Real fields are N x N float64 in Julia column-major order: element
 *     (i, j) (0-based, i = first index, along which k_x varies) at i + j*N.
 *     Host buffers are copied; the caller keeps ownership.
 *   - Calls are synchronous on return.  One context per host thread; a
 *     context is not re-entrant.
 *   - Supported grids: N = 2^m with 16 <= N <= 4096, or N = 3*2^m with
 *     48 <= N <= 768 (radix-3 first stage); spacing > 0.
 */
#ifndef CSG_H
#define CSG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum csg_status {
  CSG_OK = 0,
  CSG_ERR_ARG = 1,          /* invalid argument (size, pointer, precision) */
  CSG_ERR_HIP = 2,          /* HIP runtime error (message in csg_last_error) */
  CSG_ERR_NONFINITE = 3,    /* non-finite composition: see csg_nonfinite_step */
  CSG_ERR_UNSUPPORTED = 4,  /* grid size not compiled in */
  CSG_ERR_STATE = 5         /* fields not set */
} csg_status;

typedef enum csg_precision {
  CSG_FP64 = 0,   /* FP64 throughout */
  CSG_MIXED = 1   /* FP32 FFTs and pointwise work; FP64 state, k=0 mode, mean,
                     reductions and stabilisation denominator */
} csg_precision;

/* Mirrors `Parameters` (core_shell.jl:4-15), field for field, same order. */
typedef struct csg_params {
  double A, B, chi, P, ca, cb, cc, kappa, mobility, Q;
} csg_params;

typedef struct csg_ctx csg_ctx;

/* The defaults of core_shell.jl:5-14. */
void csg_default_params(csg_params* p);

/* Create a context for an N x N grid (make_grid, core_shell.jl:17-25).
 * precision: csg_precision; device: HIP device ordinal. */
int csg_create(csg_ctx** out, int32_t n, double spacing, const csg_params* p,
               int32_t precision, int32_t device);
int csg_destroy(csg_ctx* ctx);

/* Upload c0 and phi (host, column-major N x N; copied).  geometry(phi, p)
 * (core_shell.jl:39-43) is evaluated on the device inside the passes. */
int csg_set_fields(csg_ctx* ctx, const double* c0, const double* phi);

/* The loop of run_simulation (core_shell.jl:90-95): for s = 0..steps record
 * the history row of c_s, stop on non-finite c_s, advance unless s == steps.
 * history: caller-allocated, (rows x 6) column-major float64 with
 * rows = steps / hist_every + 1 (rows s = 0, hist_every, 2*hist_every, ...);
 * may be NULL.  Columns: [t, mean c, F, min c, max c, dissipation].
 * Returns CSG_ERR_NONFINITE if any c_s is non-finite; csg_nonfinite_step
 * then gives the first such s, history rows up to it are valid, and the
 * resident field is unspecified (a run of >= 4096 steps checks the device
 * flag every 2048 steps and stops within 4096 steps of s; a shorter run
 * completes its steps first). */
int csg_run(csg_ctx* ctx, double dt, int64_t steps, int64_t hist_every, double* history);

/* nsteps applications of step (core_shell.jl:72-76) to the resident field. */
int csg_step(csg_ctx* ctx, double dt, int64_t nsteps);

/* Download the resident composition field (column-major N x N). */
int csg_get_field(csg_ctx* ctx, double* c_out);

/* First step with non-finite composition in the last csg_run, or -1. */
int64_t csg_nonfinite_step(const csg_ctx* ctx);

/* chemical_potential(c, shape, grid, p) (core_shell.jl:57-60).  c, mu: host
 * column-major N x N.  Does not touch the resident field. */
int csg_chemical_potential(csg_ctx* ctx, const double* c, double* mu);

/* transport_rate(c, shape, grid, p) (core_shell.jl:62-70): the full N x N
 * k-space rate in FFTW convention (unnormalised forward transform), split into
 * real and imaginary column-major arrays, and the dissipation scalar.  Any of
 * the output pointers may be NULL. */
int csg_transport_rate(csg_ctx* ctx, const double* c, double* rate_re, double* rate_im,
                       double* dissipation);

/* energy(c, shape, grid, p) (core_shell.jl:78-82). */
int csg_energy(csg_ctx* ctx, const double* c, double* F);

/* geometry(phi, p) (core_shell.jl:39-43) of the uploaded phi: h, g, M
 * (column-major N x N; any may be NULL). */
int csg_geometry(csg_ctx* ctx, double* h, double* g, double* M);

/* bulk_energy(c, shape, p) (core_shell.jl:45-49), pointwise f(c, phi). */
int csg_bulk_energy(csg_ctx* ctx, const double* c, double* f);

/* bulk_derivative(c, shape, p) (core_shell.jl:51-55), pointwise f'(c, phi). */
int csg_bulk_derivative(csg_ctx* ctx, const double* c, double* fp);

/* initial_fields(n; spacing, radius, width, seed) (core_shell.jl:27-37),
 * host only: phi = 0.5(1 - tanh((r - radius)/width)) about index n/2, and
 * c = 0.5 + interior-mean-free noise 0.01(2 rand - 1) inside phi >= 0.5,
 * with rand = Julia's rand(MersenneTwister(seed), n, n) bit for bit
 * (Julia >= 1.11 seeding, dSFMT-19937).  Any n >= 1. */
int csg_initial_fields(int32_t n, double spacing, double radius, double width, int64_t seed,
                       double* c_out, double* phi_out);

/* Julia's rand(MersenneTwister(seed), count) in column-major order, host only. */
int csg_julia_rand(int64_t seed, int64_t count, double* out);

/* Mean device time per launch of each pass kernel over nsteps loop
 * iterations, from HIP events (no system fence) recorded between the kernels
 * of each iteration, launched in csg_run's order: ms[0..3] = P1, P2, P3,
 * P4 (P4 with the fused history-row reduction), ms[4] = one launch of the
 * standalone reduction kernel used for a run's last row.  Advances the
 * resident field by nsteps steps. */
int csg_time_passes(csg_ctx* ctx, double dt, int64_t nsteps, double* ms);

/* Measured copy bandwidth of the device: a streaming copy kernel (16-byte
 * words) over two fresh buffers of `bytes` bytes, 3 untimed and `reps` timed
 * launches between HIP events; *gbs = 2 * bytes * reps / time, in GB/s (read
 * plus write).  A diagnostic with no counterpart in the reference: the
 * measured ceiling that pass bandwidths are read against, beside the
 * published peak.  bytes: a multiple of 16 in [16, 16 GiB].  Allocates and
 * frees its buffers; synchronous. */
int csg_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double* gbs);

/* Adjoint of the time-step loop, FP64 contexts only.  No counterpart in the
 * reference (which has no derivatives); it differentiates step
 * (core_shell.jl:72-76) iterated from the resident field.
 *
 * csg_adjoint_forward: advance the resident field by `steps` steps of `dt`,
 * keeping checkpoints every `ckpt_every` steps; the field c(T) stays
 * resident.  ckpt_every = 1 keeps every state ((steps + 1) N^2 doubles of
 * device memory), so the backward sweep recomputes nothing; larger
 * intervals (sqrt(steps) minimises memory) trade memory for recomputation. */
int csg_adjoint_forward(csg_ctx* ctx, double dt, int64_t steps, int64_t ckpt_every);

/* csg_adjoint_backward: given dL/dc(T) (host, column-major N x N), return
 * dL/dc(0) (may be NULL) and dL_dparams[0..2] = the gradient with respect to
 * (kappa, P, chi) through the dynamics (may be NULL); the loss's own direct
 * dependence on the parameters is the caller's.  Recomputes each segment's
 * states from its checkpoint (none when every state was kept); the resident
 * field is not changed.  The result does not depend on ckpt_every. */
int csg_adjoint_backward(csg_ctx* ctx, const double* dL_dcT, double* dL_dc0, double* dL_dparams);

/* Bytes of device memory held by the context. */
int64_t csg_device_bytes(const csg_ctx* ctx);

/* Replace the context's Parameters (core_shell.jl:4-15) without reallocating
 * anything; the resident field is kept.  A checkpointed forward run
 * (csg_adjoint_forward) is discarded. */
int csg_set_params(csg_ctx* ctx, const csg_params* p);

/* ---- kernel schedules ------------------------------------------------------
 * No counterpart in the reference (FFTW plans its transforms itself).  Each
 * of the four pass kernels of a step (pass 0..3 = P1 row, P2 column, P3 row,
 * P4 column) is compiled in up to 7 candidate schedules, fastest first as
 * measured on MI355X; candidate 0 is the default.  All candidates compute the
 * same step (they differ in the FFT radix plan, so results agree to rounding,
 * not bit for bit).
 *
 * csg_schedule_info: count of candidates of `pass`; `sched` (may be NULL,
 * else 7 * count int32) receives each candidate's [elements per thread, rows
 * or lines per workgroup, tile-load group, prefetch (when the second input
 * is loaded: 0 where used, 1 with the first loads, 2 once they have arrived),
 * FFT exchange layout,
 * twiddle source (1 = per-stage tables, 2 = half table in LDS beyond 16 KiB),
 * P3 split (1 = one gradient component per workgroup, 2 = the same with the
 * two components of a row tile on neighbouring workgroups, 0 = both in one;
 * 0 for the other passes)]. */
int csg_schedule_info(const csg_ctx* ctx, int32_t pass, int32_t* count, int32_t* sched);
/* Selected candidate per pass (sel[0..3]). */
int csg_get_schedule(const csg_ctx* ctx, int32_t* sel);
int csg_set_schedule(csg_ctx* ctx, const int32_t* sel);
/* Runtime tuner: times the captured loop at `dt` for the candidates (rounds
 * of coordinate descent over the passes, starting from the current
 * selection), keeps the fastest combination and returns it in sel_out (may be
 * NULL) with its device time per loop iteration in microseconds (may be
 * NULL).  Needs csg_set_fields; the resident field is restored. */
int csg_autotune(csg_ctx* ctx, double dt, int32_t rounds, int32_t* sel_out, double* us_per_step);

/* ---- ensembles ------------------------------------------------------------
 * No counterpart in the reference, which runs one particle per call: an
 * ensemble is `count` independent instances of run_simulation
 * (core_shell.jl:84-103), each with its own grid size, spacing, parameters,
 * c0 and phi, advanced together.  Instances of equal size share batched
 * kernel launches (one instance per grid row of workgroups); the sizes run as
 * parallel branches of one captured graph per 32 loop iterations.  Same
 * supported sizes and precisions as csg_create. */
typedef struct csg_batch csg_batch;

/* n[i], spacing[i] (NULL: all 1.0), params[i] (NULL: all defaults). */
int csg_batch_create(csg_batch** out, int32_t count, const int32_t* n, const double* spacing,
                     const csg_params* params, int32_t precision, int32_t device);
int csg_batch_destroy(csg_batch* b);

/* Upload instance i's c0 and phi (host, column-major n[i] x n[i]; copied). */
int csg_batch_set_fields(csg_batch* b, int32_t i, const double* c0, const double* phi);

/* csg_run for every instance: history[i] (may be NULL, as may `history`)
 * receives instance i's (rows x 6) column-major history as in csg_run.  An
 * instance whose composition turns non-finite keeps its own record
 * (csg_batch_nonfinite_step) and does not stop the others; the call then
 * returns CSG_ERR_NONFINITE after all instances have run. */
int csg_batch_run(csg_batch* b, double dt, int64_t steps, int64_t hist_every, double* const* history);

/* Download instance i's composition field (column-major n[i] x n[i]). */
int csg_batch_get_field(csg_batch* b, int32_t i, double* c_out);

/* Persistent mode (enable = 1): csg_batch_run executes every pass of every
 * member for a block of 32 loop iterations in ONE launch of a persistent
 * kernel (captured in a graph), with per-member completion counters in place
 * of kernel boundaries, instead of one batched launch per size class and
 * pass.  Grid sizes 128, 256, 384 and 512 only (CSG_ERR_UNSUPPORTED
 * otherwise), at most 4 of them; the kernel uses its own fixed schedules (the
 * members' csg_batch_set_schedule choices apply to the batched mode only), so
 * results agree with the batched mode to rounding.  enable = 0 returns to the
 * batched mode. */
int csg_batch_set_persistent(csg_batch* b, int32_t enable);

/* First non-finite step of instance i in the last csg_batch_run, or -1. */
int64_t csg_batch_nonfinite_step(const csg_batch* b, int32_t i);

/* Bytes of device memory held by the ensemble. */
int64_t csg_batch_device_bytes(const csg_batch* b);

/* Schedules of the members of grid size n (as csg_get_schedule /
 * csg_set_schedule), and the runtime tuner applied to each size class alone
 * (the resident fields are restored). */
int csg_batch_get_schedule(const csg_batch* b, int32_t n, int32_t* sel);
int csg_batch_set_schedule(csg_batch* b, int32_t n, const int32_t* sel);
int csg_batch_autotune(csg_batch* b, double dt, int32_t rounds);

/* ---- allocation-free stepping ----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code between graph captures, tight serving loops, harnesses that
 * forbid hipMalloc): nsteps applications of step (core_shell.jl:72-76) to a
 * composition field already in device memory, with the same kernels, default
 * schedule and twiddles as csg_create + csg_set_fields + csg_step +
 * csg_get_field, so the result is identical bit for bit.
 *
 * c_dev: device, column-major N x N float64, advanced in place.  phi_dev:
 * device, column-major N x N float64.  p: NULL for the defaults.  stream: a
 * hipStream_t, or NULL for the null stream; the call returns after the work
 * on it has finished.
 *
 * Everything else the kernels use (packed state, work arrays, partials,
 * twiddles) lives in a static device array of CSG_NOALLOC_SCRATCH_BYTES that
 * each GPU the library runs on holds until the process ends; the call
 * allocates nothing.  Grids above CSG_NOALLOC_MAX_N are rejected with
 * CSG_ERR_ARG.  Calls on one device share that array: the library serialises
 * them with a host lock within a process; work a caller enqueues on the same
 * device concurrently is not affected (the array is private to the library). */
#define CSG_NOALLOC_MAX_N 1024
#define CSG_NOALLOC_SCRATCH_BYTES 26009344
int csg_step_noalloc(int32_t n, double spacing, const csg_params* p, int32_t precision, int32_t device,
                     double dt, int64_t nsteps, double* c_dev, const double* phi_dev, void* stream);

const char* csg_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CSG_H */
