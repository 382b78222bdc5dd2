"""Benchmark: lattice-site updates/s of the core-shell time loop at 1024^2.

Workload (one "step" = one iteration of run_simulation's loop, core_shell.jl:
90-95, history row included): N = 1024, R = 320, w = 2, dt = 0.01, seed 5749,
default Parameters, fields resident in HBM before the timed region.  `value`
is the FP64 form, the reference's precision.  The mixed form (FP32 FFTs and
pointwise work; FP64 state, k = 0 mode, mean and reductions) runs beside it on
the same inputs and is reported with its speed-up and its error against the
FP64 GPU run at a fixed 1000-step horizon (independent of --steps).  Before
each timed run the library's runtime tuner (csg_autotune) picks the kernel
schedules on the resident field (--no-tune keeps the defaults).

Also reported:
  roofline      dominant pass of the FP64 form: algorithmic bytes per launch
                (per-site bytes x N^2) / its mean in-loop HIP-event duration
                (timing "hip_event_in_loop"); beside it the rocprofv3
                kernel-trace average and HBM traffic committed under
                profiles/ for the same pass, form and schedule (null when the
                schedule run here has no committed profile); measured_copy
                holds the device copy ceilings beside the 8 TB/s spec peak
  cpu_baseline  the FP64 numpy oracle (reference-equivalent: 17 complex
                N^2 FFTs per iteration), 1 thread, on a bounded sample of the
                same workload, with the host's CPU model and thread counts
  ensemble      the uneven 256-particle ensemble (mixed and FP64 forms)
  large_field   4096^2 Janus-type field, 5000 steps, per-pass HBM GB/s, mixed error
  gradient      gradient evaluations/s of the 512^2 adjoint workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per site and pass (P1..P4); see the pass table in the
# kernel header: state S is always FP64, bulk arrays and phi are 8 or 4 bytes.
PASS_BYTES = {
    "fp64": (24.0, 32.0, 40.0, 32.0),
    "mixed": (16.0, 20.0, 20.0, 24.0),
}
PASS_NAMES = ("P1_row", "P2_col", "P3_row", "P4_col")


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--radius", type=float, default=320.0)
    ap.add_argument("--width", type=float, default=2.0)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default kernel schedules (no csg_autotune before the timed runs)")
    ap.add_argument("--large-steps", type=int, default=5000, help="steps of the 4096^2 field (0: skip)")
    ap.add_argument("--ens-count", type=int, default=256, help="ensemble members (0: skip the ensemble)")
    ap.add_argument("--ens-steps", type=int, default=2000)
    ap.add_argument("--grad-steps", type=int, default=2000,
                    help="steps of the 512^2 adjoint workload timed for gradient evaluations/s "
                         "(0 skips it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (3 times), the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--counters", type=int, default=0, metavar="STEPS",
                    help="only run STEPS eager loop iterations (no graph capture) of each form at "
                         "the --n field and at 4096^2 with the default schedules: the workload of "
                         "the rocprofv3 --pmc passes, whose counter collection does not follow "
                         "kernels launched from captured graphs")
    return ap.parse_args()


def counter_pass(pkg, a) -> None:
    """Eager loop iterations (csg_step: one launch per pass, no graph) of both
    forms on the --n workload and on the 4096^2 large field."""
    for n, radius, p in ((a.n, a.radius, pkg.Parameters()), (4096, 1300.0, janus_params(pkg))):
        c0, phi = pkg.initial_fields(n, radius=radius, width=a.width, seed=5749)
        for form in ("fp64", "mixed"):
            s = pkg.Session(n, 1.0, p, form)
            s.set_fields(c0, phi)
            s.step(a.dt, a.counters)
            print(json.dumps({"counters_workload": f"{n}x{n}", "form": form, "steps": a.counters,
                              "schedule": list(s.schedule)}), flush=True)
            s.close()


def calibrate_copy(mib: int) -> None:
    """hipMemcpy device-to-device of a known size (runtime copy kernel)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, n) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, n, 3) != 0:   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": mib << 20, "repeats": 3}), flush=True)


def copy_ceiling(pkg) -> dict:
    """Measured device copy ceilings (read + write bytes / time, GB/s): the
    library's streaming copy kernel and the runtime's hipMemcpyAsync, each on
    a 1 GiB buffer pair, which streams through HBM, and on a 32 MiB pair,
    which stays inside the 256 MB Infinity Cache like the 1024^2 working set."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    def memcpy_gbs(mib, reps):
        n = ctypes.c_size_t(mib << 20)
        a, b, e0, e1 = vp(), vp(), vp(), vp()
        ok(hip.hipMalloc(ctypes.byref(a), n), "hipMalloc")
        ok(hip.hipMalloc(ctypes.byref(b), n), "hipMalloc")
        ok(hip.hipMemset(a, 1, n), "hipMemset")
        ok(hip.hipEventCreate(ctypes.byref(e0)), "hipEventCreate")
        ok(hip.hipEventCreate(ctypes.byref(e1)), "hipEventCreate")
        for _ in range(3):
            ok(hip.hipMemcpyAsync(b, a, n, 3, vp()), "hipMemcpyAsync")   # hipMemcpyDeviceToDevice
        ok(hip.hipEventRecord(e0, vp()), "hipEventRecord")
        for _ in range(reps):
            ok(hip.hipMemcpyAsync(b, a, n, 3, vp()), "hipMemcpyAsync")
        ok(hip.hipEventRecord(e1, vp()), "hipEventRecord")
        ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = ctypes.c_float()
        ok(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1), "hipEventElapsedTime")
        for e in (e0, e1):
            hip.hipEventDestroy(e)
        for q in (a, b):
            hip.hipFree(q)
        return 2.0 * (mib << 20) * reps / (ms.value * 1e-3) / 1e9

    return {"copy_kernel": {"hbm_1GiB": pkg.copy_bandwidth(1 << 30, 20),
                            "cache_32MiB": pkg.copy_bandwidth(32 << 20, 400)},
            "hipMemcpyAsync": {"hbm_1GiB": memcpy_gbs(1024, 20), "cache_32MiB": memcpy_gbs(32, 400)},
            "unit": "GB/s"}


def _sched_str(sched) -> str:
    return ", ".join(str(int(x)) for x in sched)


def _traffic(kernel: str, workload: str, sched: str):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE runs,
    if a summary for this kernel, workload and schedule has been committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        if d.get("schedule") == sched:
            return d.get("bytes_per_launch")
    return None


def _rocprof_us(kernel: str, workload: str, form: str, sched: str):
    """rocprofv3 kernel-trace average duration (us) of this pass, form and
    schedule from the committed summary, if there is one."""
    path = os.path.join(ROOT, "profiles", f"kernel_trace_{workload}.json")
    if os.path.exists(path):
        with open(path) as f:
            for e in json.load(f).get("kernels", []):
                if e.get("pass") == kernel and e.get("form") == form and e.get("schedule") == sched:
                    return e.get("avg_us")
    return None


_HIP = []


def _device_sync():
    """hipDeviceSynchronize: the barrier on both sides of a timed region (the
    runtime library is loaded once, outside the timed regions)."""
    if not _HIP:
        import ctypes
        _HIP.append(ctypes.CDLL("libamdhip64.so"))
    if _HIP[0].hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def host_cpu() -> dict:
    """The host's CPU as this process sees it: usable threads, model name and
    threads per core (lscpu, else /proc/cpuinfo)."""
    import subprocess
    info = {"nproc": len(os.sched_getaffinity(0)), "model": None, "threads_per_core": None}
    try:
        r = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=20)
        for ln in r.stdout.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Thread(s) per core":
                info["threads_per_core"] = int(v.strip())
    except (OSError, ValueError, subprocess.TimeoutExpired):
        pass
    if info["model"] is None:
        try:
            with open("/proc/cpuinfo") as f:
                for ln in f:
                    if ln.startswith("model name"):
                        info["model"] = ln.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def run_form(pkg, form, c0, phi, a):
    s = pkg.Session(a.n, 1.0, pkg.Parameters(), form)
    s.set_fields(c0, phi)
    tune = None
    if not a.no_tune:
        t0 = time.perf_counter()
        sel, us = s.autotune(a.dt, 2)            # runtime tuner, resident field kept
        tune = dict(schedule=list(sel), tuned_us_per_step=us, seconds=time.perf_counter() - t0)
    if a.warmup:
        s.run(a.dt, a.warmup)                 # also captures the step graph
    s.set_fields(c0, phi)
    _device_sync()
    t0 = time.perf_counter()
    hist = s.run(a.dt, a.steps)               # synchronous: returns after the stream drains
    _device_sync()
    t1 = time.perf_counter()
    c = s.field()
    # five repeats of at least 200 steps (continuing from the field) for a
    # median and spread
    reps = []
    k = max(200, a.steps // 10)
    for _ in range(5):
        r0 = time.perf_counter()
        s.run(a.dt, k)
        reps.append(1e3 * (time.perf_counter() - r0) / (k + 1))
    ms_pass = s.time_passes(a.dt, 50)
    sched = list(s.schedule)
    info = s.schedule_info()
    resolved = [_sched_str(info[q][sched[q]][k] for k in ("E", "W", "G", "PF", "X", "T", "S")) for q in range(4)]
    s.close()
    return dict(seconds=t1 - t0, hist=hist, c=c, ms_pass=ms_pass, schedule=sched, resolved=resolved, tune=tune,
                reps=reps, rep_steps=k)


def gradient_rate(pkg, a):
    """Gradient evaluations/s of the adjoint workload: 512^2, R=160, w=2,
    a.grad_steps steps, loss 1/2||c(T)-c*||^2 + 1e-3 F.  Timed with the
    library's default checkpointing (every state kept when the run fits the
    device-memory budget: no recompute) and with one checkpoint per
    ~sqrt(steps) steps (recompute of every segment)."""
    n, steps = 512, a.grad_steps
    c0, phi = pkg.initial_fields(n, radius=160.0, width=2.0, seed=5749)
    p0 = pkg.Parameters()
    pt = pkg.Parameters(kappa=p0.kappa * 1.1, P=p0.P * 1.1, chi=p0.chi * 1.1)
    s = pkg.Session(n, 1.0, pt, "fp64")
    s.set_fields(c0, phi)
    s.step(a.dt, steps)
    tgt = s.field()
    s.close()

    def rate(k):
        pkg.adjoint.gradient(c0, phi, tgt, dt=a.dt, steps=steps, lam=1e-3, ckpt_every=k)      # warm-up
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            g = pkg.adjoint.gradient(c0, phi, tgt, dt=a.dt, steps=steps, lam=1e-3, ckpt_every=k)
        return (time.perf_counter() - t0) / reps, g

    kd = pkg.adjoint.default_ckpt_every(n, steps)
    ks = max(1, round(steps ** 0.5))
    secs, g = rate(kd)
    secs_s, _ = rate(ks)
    kept = "every state kept" if kd == 1 else f"checkpoints every {kd}"
    return {"workload": f"adjoint {n}x{n}, R=160, w=2, {steps} steps, {kept} "
                        f"({(steps // kd + 1) * n * n * 8 / 1e9:.1f} GB)",
            "ckpt_every": kd, "gradient_evals_per_s": 1.0 / secs, "seconds_per_gradient": secs, "loss": g.loss,
            "sqrt_checkpoints": {"ckpt_every": ks, "gradient_evals_per_s": 1.0 / secs_s,
                                 "seconds_per_gradient": secs_s}}


# Janus-type two-well set of the large-field workload: B raised 2x and chi
# shifted by +0.3 from core_shell.jl's defaults, with the stabilisation Q
# raised 2x: at Q = 0.5 the scheme turns non-finite at dt = 0.01 with B = 8
# (step 964 at 4096^2).  The reference has no such set; this is the definition
# used here.
def janus_params(pkg):
    d = pkg.Parameters()
    return pkg.Parameters(B=2.0 * d.B, chi=d.chi + 0.3, Q=2.0 * d.Q)


def large_field(pkg, a):
    """The large single field: 4096^2, R = 1300, w = 2, dt, a.large_steps
    steps, Janus-type parameters; FP64 and mixed forms on the same inputs, HBM
    GB/s per pass from event-timed passes, mixed error against FP64."""
    n = 4096
    c0, phi = pkg.initial_fields(n, radius=1300.0, width=2.0, seed=5749)
    p = janus_params(pkg)
    res = {}
    for form in ("fp64", "mixed"):
        s = pkg.Session(n, 1.0, p, form)
        s.set_fields(c0, phi)
        if not a.no_tune:
            s.autotune(a.dt, 1)
        s.run(a.dt, 64)                            # captures the step graph
        s.set_fields(c0, phi)
        t0 = time.perf_counter()
        hist = s.run(a.dt, a.large_steps)
        secs = time.perf_counter() - t0
        res[form] = dict(secs=secs, hist=hist, c=s.field(), ms=s.time_passes(a.dt, 20), sched=s.schedule)
        s.close()
    out = {"workload": f"large field {n}x{n}, R=1300, w=2, Janus-type set B={p.B:g}, chi={p.chi:g}, Q={p.Q:g}, "
                       f"{a.large_steps} steps, dt={a.dt:g}, history every step",
           "deviation": "Q is doubled as well (the stated set changes only B and chi): with Q=0.5 the "
                        "stabilised scheme turns non-finite at dt=0.01 (step 964 at 4096^2) in both forms"}
    sites = n * n
    for form, r in res.items():
        ms = r["ms"]
        out[form] = {"site_updates_per_s": sites * (a.large_steps + 1) / r["secs"],
                     "ms_per_step": 1e3 * r["secs"] / (a.large_steps + 1),
                     "steps_per_s": (a.large_steps + 1) / r["secs"],
                     "pass_gbs": {PASS_NAMES[k]: PASS_BYTES[form][k] * sites / (ms[k] * 1e-3) / 1e9 for k in range(4)},
                     "pass_ms": {PASS_NAMES[k]: float(ms[k]) for k in range(4)},
                     "step_frac_hbm": sum(PASS_BYTES[form]) * sites * (a.large_steps + 1) / r["secs"] / 1e9 / HBM_PEAK_GBS,
                     "schedule": list(r["sched"]),
                     "mass_drift": float(np.max(np.abs(r["hist"][:, 1] - r["hist"][0, 1]))),
                     "energy_final": float(r["hist"][-1, 2])}
    out["mixed"]["max_dc_vs_fp64"] = float(np.max(np.abs(res["mixed"]["c"] - res["fp64"]["c"])))
    out["mixed"]["rel_energy_err_vs_fp64"] = float(np.max(np.abs(res["mixed"]["hist"][:, 2] - res["fp64"]["hist"][:, 2])
                                                          / np.abs(res["fp64"]["hist"][:, 2])))
    return out


def ensemble_rate(pkg, a):
    """The uneven ensemble: a.ens_count members, N drawn from {128, 256, 384,
    512}, R = 0.3 N, w = 2, P and chi within +-30% of the defaults, seed 5749 +
    index, a.ens_steps steps at dt, history every step kept on the device.
    Mixed form timed (FP64 beside it); the error of every member's mixed run
    against its FP64 run is reported."""
    ms = pkg.ensemble_members(a.ens_count)
    fields = [pkg.member_fields(m) for m in ms]
    sizes = sorted({m.n for m in ms})
    sites = sum(m.n * m.n for m in ms)
    res = {}
    for form in ("mixed", "fp64"):
        ens = pkg.Ensemble([m.n for m in ms], params=[m.p for m in ms], precision=form)
        for i, (c0, phi) in enumerate(fields):
            ens.set_fields(i, c0, phi)
        tuned = None
        if not a.no_tune:
            t0 = time.perf_counter()
            tuned = {str(n): list(sel) for n, sel in ens.autotune(a.dt, 1).items()}   # per size class
            tuned["seconds"] = time.perf_counter() - t0
        ens.run(a.dt, 64, histories=False)           # warm-up: graph capture, caches
        for i, (c0, phi) in enumerate(fields):
            ens.set_fields(i, c0, phi)
        t0 = time.perf_counter()
        hs = ens.run(a.dt, a.ens_steps)
        secs = time.perf_counter() - t0
        res[form] = dict(secs=secs, hist=hs, c=[ens.field(i) for i in range(len(ms))],
                         bytes=ens.device_bytes(), tuned=tuned)
        ens.close()
    iters = a.ens_steps + 1
    out = {"workload": f"ensemble of {len(ms)} particles, N in {sizes} "
                       f"({', '.join(f'{n}: {sum(m.n == n for m in ms)}' for n in sizes)}), R=0.3N, w=2, "
                       f"P, chi within +-30%, {a.ens_steps} steps, dt={a.dt:g}, history every step",
           "sites": sites}
    for form, r in res.items():
        byts = sites * sum(PASS_BYTES[form]) * iters   # packed line layout: no N/2+1 column
        gbs = byts / r["secs"] / 1e9
        out[form] = {"site_updates_per_s": sites * iters / r["secs"], "ms_per_step": 1e3 * r["secs"] / iters,
                     "steps_per_s": iters / r["secs"],
                     "algorithmic_gbs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS, "device_bytes": r["bytes"],
                     "autotuned_schedules": r["tuned"]}
    dc = [float(np.max(np.abs(x - y))) for x, y in zip(res["mixed"]["c"], res["fp64"]["c"])]
    de = [float(np.max(np.abs(x[:, 2] - y[:, 2]) / np.abs(y[:, 2])))
          for x, y in zip(res["mixed"]["hist"], res["fp64"]["hist"])]
    out["mixed"]["max_dc_vs_fp64"] = max(dc)
    out["mixed"]["median_dc_vs_fp64"] = float(np.median(dc))
    out["mixed"]["rel_energy_err_vs_fp64"] = max(de)
    out["mixed"]["mass_drift"] = max(float(np.max(np.abs(h[:, 1] - h[0, 1]))) for h in res["mixed"]["hist"])
    return out


def horizon_errors(pkg, form, c0, phi, a, steps=1000):
    """`form` against the FP64 GPU run on the same inputs after `steps` steps
    (t = 10 at dt = 0.01, the horizon of the FP64 parity gate), whatever
    --steps is: max|dc|, the largest relative energy error over the history,
    and the form's mass drift."""
    out = []
    for f in ("fp64", form):
        s = pkg.Session(a.n, 1.0, pkg.Parameters(), f)
        s.set_fields(c0, phi)
        h = s.run(a.dt, steps)
        out.append((s.field(), h))
        s.close()
    (c64, h64), (cf, hf) = out
    return {"steps": steps, "max_dc_vs_fp64": float(np.max(np.abs(cf - c64))),
            "rel_energy_err_vs_fp64": float(np.max(np.abs(hf[:, 2] - h64[:, 2]) / np.abs(h64[:, 2]))),
            "mass_drift": float(np.max(np.abs(hf[:, 1] - hf[0, 1])))}


def cpu_all_cores_baseline(c0, phi, a, threads, c_ref):
    """The oracle loop restated in C with OpenMP (oracle/cs_cpu.c, the same 17
    complex 2D transforms per loop iteration as core_shell.jl), timed on
    `threads` threads over a.cpu_steps steps and on one thread over a third of
    them, and checked against the numpy oracle's field (c_ref) after
    a.cpu_steps steps.  A missing C compiler or library does not stop the
    benchmark: the leg then reports the error."""
    try:
        from oracle import cpu_restatement as cr
        c0, phi = np.asarray(c0), np.asarray(phi)
        t0 = time.perf_counter()
        c, _ = cr.run_simulation(c0, phi, dt=a.dt, steps=a.cpu_steps, threads=threads)
        secs = time.perf_counter() - t0
        k1 = max(1, a.cpu_steps // 3)
        t0 = time.perf_counter()
        cr.run_simulation(c0, phi, dt=a.dt, steps=k1, threads=1)
        secs1 = time.perf_counter() - t0
    except Exception as e:   # noqa: BLE001 - reported, not fatal
        return dict(error=f"{type(e).__name__}: {e}")
    return dict(value=a.n * a.n * (a.cpu_steps + 1) / secs, unit="site-updates/s", cores=threads, kind="port",
                single_thread_value=a.n * a.n * (k1 + 1) / secs1,
                max_dc_vs_numpy_oracle=float(np.max(np.abs(c - c_ref))),
                sample=f"the oracle loop restated in C with OpenMP (oracle/cs_cpu.c, radix-2 transforms), "
                       f"{a.cpu_steps + 1} loop iterations of the {a.n}^2 workload on {threads} threads in "
                       f"{secs:.1f} s; {k1 + 1} iterations on 1 thread in {secs1:.1f} s")


def _cpu_threads() -> int:
    """Threads of the all-cores leg: the CPU share of this process (at most
    16, the share of one GPU on the benchmark machines), or OMP_NUM_THREADS."""
    n = min(16, len(os.sched_getaffinity(0)))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(c0, phi, a, c_gpu_steps=None):
    from oracle import core_shell_oracle as orc
    c0 = np.asarray(c0)
    phi = np.asarray(phi)
    t0 = time.perf_counter()
    c, _ = orc.run_simulation(c0, phi, dt=a.dt, steps=a.cpu_steps)
    t1 = time.perf_counter()
    secs = t1 - t0
    # run_simulation evaluates steps+1 loop iterations (the last without advancing)
    rate = a.n * a.n * (a.cpu_steps + 1) / secs
    return c, dict(value=rate, unit="site-updates/s", cores=1, kind="port", host=host_cpu(),
                   sample=f"oracle run_simulation (numpy FP64, 17 complex {a.n}^2 FFTs per iteration), "
                          f"{a.cpu_steps + 1} loop iterations of the {a.n}^2 workload in {secs:.1f} s")


def gpu_clocks():
    """The GPU's current clock levels as rocm-smi reports them (None if the
    tool is not available, or under rocprofv3, whose preload would make the
    child initialise the GPU before rocm-smi's interpreter starts)."""
    import subprocess
    if any(k.startswith("ROCPROF") for k in os.environ):
        return None
    try:
        r = subprocess.run(["rocm-smi", "--showclocks"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=30)
    except (OSError, subprocess.TimeoutExpired):
        return None
    keep = [ln.split(":", 1)[-1].strip() if "GPU[" in ln else ln.strip()
            for ln in r.stdout.splitlines() if "clock level" in ln.lower()]
    return keep or None


def main():
    a = _args()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.counters:
        import __graft_entry__
        counter_pass(__graft_entry__.load_package(), a)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the library has no "
              f"multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    c0, phi = pkg.initial_fields(a.n, radius=a.radius, width=a.width, seed=5749)
    forms = [f for f in a.forms.split(",") if f]
    forms = sorted(forms, key=lambda f: f != "fp64")          # FP64 first
    res = {f: run_form(pkg, f, c0, phi, a) for f in forms}
    sites = a.n * a.n
    workload = f"{a.n}x{a.n}"
    out = {}
    ref = res.get("fp64")
    for f, r in res.items():
        ups = sites * (a.steps + 1) / r["seconds"]
        ms = r["ms_pass"]
        per = [PASS_BYTES[f][k] * sites / (ms[k] * 1e-3) / 1e9 for k in range(4)]
        dom = int(np.argmax(ms[:4]))
        step_bytes = sum(PASS_BYTES[f]) * sites
        o = dict(value=ups, ms_per_step=1e3 * r["seconds"] / (a.steps + 1), steps_per_s=(a.steps + 1) / r["seconds"],
                 pass_ms=dict(zip(PASS_NAMES + ("finalize",), [float(x) for x in ms])),
                 pass_gbs=dict(zip(PASS_NAMES, per)),
                 step_gbs=step_bytes * (a.steps + 1) / r["seconds"] / 1e9,
                 dominant=PASS_NAMES[dom], dominant_gbs=per[dom], dominant_ms=float(ms[dom]),
                 dominant_bytes=PASS_BYTES[f][dom] * sites, dominant_schedule=r["resolved"][dom],
                 energy_final=float(r["hist"][-1, 2]),
                 mass_drift=float(np.max(np.abs(r["hist"][:, 1] - np.mean(c0)))),
                 schedule=r["schedule"], schedule_resolved=r["resolved"], autotune=r["tune"],
                 ms_per_step_repeats=[float(x) for x in r["reps"]], steps_per_repeat=r["rep_steps"],
                 ms_per_step_median_of_repeats=float(np.median(r["reps"])))
        if ref is not None and f != "fp64":
            o["speedup_vs_fp64"] = ups / (sites * (a.steps + 1) / ref["seconds"])
            o["max_dc_vs_fp64_at_timed_steps"] = float(np.max(np.abs(r["c"] - ref["c"])))
            o["rel_energy_err_vs_fp64_at_timed_steps"] = float(np.max(np.abs(r["hist"][:, 2] - ref["hist"][:, 2])
                                                                      / np.abs(ref["hist"][:, 2])))
            o["horizon"] = horizon_errors(pkg, f, c0, phi, a)
        out[f] = o
    pick = "fp64" if "fp64" in out else forms[0]
    head = out[pick]
    dom = head["dominant"]
    sched = head["dominant_schedule"]
    prof_us = _rocprof_us(dom, workload, pick, sched)
    roof = dict(bound="hbm", achieved=head["dominant_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=head["dominant_gbs"] / HBM_PEAK_GBS, kernel=dom, schedule=sched, timing="hip_event_in_loop",
                bytes_per_launch=head["dominant_bytes"], ms_per_launch=head["dominant_ms"],
                traffic=_traffic(dom, f"{workload}_{pick}", sched),
                rocprof_ms_per_launch=None if prof_us is None else prof_us * 1e-3,
                frac_rocprof=None if prof_us is None else head["dominant_bytes"] / (prof_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                measured_copy=copy_ceiling(pkg))
    line = {
        "metric": "site-updates/s at 1024² (FP64 vs mixed, max|Δc|≤1e-5); % HBM roofline",
        "value": head["value"], "unit": "site-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if pick == "fp64" else "f32+f64",
        "data": "synthetic: initial_fields (Julia MersenneTwister(5749) noise), no checkpoint",
        "config": {"workload": f"plain method {workload}, R={a.radius:g}, w={a.width:g}, "
                               f"dt={a.dt:g}, {a.steps} steps, history every step",
                   "form": pick, "grid": a.n},
        "roofline": roof,
        "forms": out,
        "gpu_clocks": gpu_clocks(),
    }
    if a.ens_count:
        line["ensemble"] = ensemble_rate(pkg, a)
    if a.large_steps:
        line["large_field"] = large_field(pkg, a)
    if a.grad_steps:
        line["gradient"] = gradient_rate(pkg, a)
    if not a.no_cpu:
        c_cpu, base = cpu_baseline(c0, phi, a)
        # the same number of GPU steps for a parity figure at full size
        s = pkg.Session(a.n, 1.0, pkg.Parameters(), "fp64")
        s.set_fields(c0, phi)
        s.run(a.dt, a.cpu_steps)
        base["max_dc_gpu_fp64_vs_oracle"] = float(np.max(np.abs(s.field() - c_cpu)))
        s.close()
        line["cpu_baseline"] = base
        line["cpu_baseline_all_cores"] = cpu_all_cores_baseline(c0, phi, a, _cpu_threads(), c_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
