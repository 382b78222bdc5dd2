"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle, the
reference's known answers (check-results.txt) and the checks.jl identities.

FP64 tolerances: single evaluations 1e-12 relative to the field scale; time
loops max|dc| <= 1e-10 after 1000 steps (core_shell.jl run at 64^2 and 256^2).
"""
import json
import os
import sys

import numpy as np
import pytest

from oracle import core_shell_oracle as o

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "check_results.json")))
P = o.Parameters()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fields(pkg, n, radius=None, width=4.0, spacing=1.0, seed=5749):
    radius = 0.25 * n * spacing if radius is None else radius
    c, phi = pkg.initial_fields(n, spacing=spacing, radius=radius, width=width, seed=seed)
    return np.asarray(c), np.asarray(phi)


SIZES = [16, 32, 64, 128, 256, 512, 1024]


@pytest.mark.parametrize("n", SIZES)
def test_state_roundtrip(pkg, n):
    say(f"roundtrip n={n}")
    c, phi = _fields(pkg, n)
    s = pkg.Session(n)
    s.set_fields(c, phi)
    assert np.max(np.abs(s.field() - c)) < 1e-15


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_upload_roundtrip_2048(pkg, precision):
    """Uploads of 32 MB fields (16 pinned chunks each, the chunks reused across
    the phi and c uploads and across calls): the state round-trips, and phi
    arrives whole (h(phi) over every site, FP32 in the mixed form)."""
    say(f"upload roundtrip 2048^2 {precision}")
    n = 2048
    rng = np.random.default_rng(11)
    s = pkg.Session(n, precision=precision)
    tol_h = 1e-14 if precision == "fp64" else 1e-6
    for k in range(2):
        c = np.asfortranarray(0.5 + 0.4 * rng.standard_normal((n, n)))
        phi = np.asfortranarray(rng.random((n, n)))
        s.set_fields(c, phi)
        assert np.max(np.abs(s.field() - c)) < 1e-14
        shape = pkg.geometry(phi, precision=precision)
        assert np.max(np.abs(shape.h - o.geometry(phi, P)["h"])) < tol_h


@pytest.mark.parametrize("n", SIZES)
def test_energy_matches_oracle(pkg, n):
    say(f"energy n={n}")
    c, phi = _fields(pkg, n)
    shape = pkg.geometry(phi, pkg.Parameters())
    F = pkg.energy(c, shape, pkg.make_grid(n, 1.0), pkg.Parameters())
    Fr = o.energy(c, o.geometry(phi, P), o.make_grid(n, 1.0), P)
    assert abs(F / Fr - 1) < 1e-13


def test_energy_known_answers(pkg):
    say("energy known answers")
    for n, r, key in ((64, 16.0, "runs_64"), (256, 70.0, "run_256")):
        c, phi = _fields(pkg, n, radius=r)
        F = pkg.energy(c, pkg.geometry(phi), pkg.make_grid(n, 1.0))
        assert abs(F / GOLD[key]["energy_t0"] - 1) < 1e-14


@pytest.mark.parametrize("n", [16, 64, 256])
def test_pointwise_matches_oracle(pkg, n):
    say(f"pointwise n={n}")
    c, phi = _fields(pkg, n)
    shape = pkg.geometry(phi)
    ref = o.geometry(phi, P)
    for k in ("h", "g", "M"):   # values in [0, 1]; the device contracts to FMA
        assert np.max(np.abs(getattr(shape, k) - ref[k])) < 1e-14
    assert np.max(np.abs(pkg.bulk_energy(c, shape) - o.bulk_energy(c, ref, P))) < 1e-14
    assert np.max(np.abs(pkg.bulk_derivative(c, shape) - o.bulk_derivative(c, ref, P))) < 1e-14


@pytest.mark.parametrize("n,spacing", [(16, 1.0), (64, 1.0), (64, 0.5), (256, 1.0), (1024, 1.0)])
def test_chemical_potential_matches_oracle(pkg, n, spacing):
    say(f"chemical_potential n={n} spacing={spacing}")
    c, phi = _fields(pkg, n, spacing=spacing)
    mu = pkg.chemical_potential(c, pkg.geometry(phi), pkg.make_grid(n, spacing))
    mr = o.chemical_potential(c, o.geometry(phi, P), o.make_grid(n, spacing), P)
    assert np.max(np.abs(mu - mr)) < 1e-12 * max(1.0, np.max(np.abs(mr)))


def test_function_api_alternating_particles(pkg):
    """Function-API calls that alternate between two particles of one size
    (one cached device context) each use their own phi, also after a
    run_simulation on a third particle replaced the resident fields, and a
    Shape keeps the phi it was made from when the caller's array changes."""
    say("function API, alternating particles at 64^2")
    n = 64
    c, phi1 = _fields(pkg, n, radius=16.0)
    _, phi2 = _fields(pkg, n, radius=22.0, width=2.0)
    _, phi3 = _fields(pkg, n, radius=10.0)
    g = pkg.make_grid(n, 1.0)
    mine = phi1.copy(order="F")
    s1, s2 = pkg.geometry(mine), pkg.geometry(phi2)
    mine[:] = 0.0
    for shape, phi in ((s1, phi1), (s2, phi2), (s1, phi1), (None, phi3), (s2, phi2), (s1, phi1)):
        if shape is None:
            pkg.run_simulation(c, phi, dt=0.01, steps=3)
            continue
        mu = pkg.chemical_potential(c, shape, g)
        mr = o.chemical_potential(c, o.geometry(phi, P), o.make_grid(n, 1.0), P)
        assert np.max(np.abs(mu - mr)) < 1e-12 * max(1.0, np.max(np.abs(mr)))
        assert abs(pkg.energy(c, shape, g) - o.energy(c, o.geometry(phi, P), o.make_grid(n, 1.0), P)) \
            < 1e-12 * abs(o.energy(c, o.geometry(phi, P), o.make_grid(n, 1.0), P))


@pytest.mark.parametrize("n,spacing", [(16, 1.0), (64, 1.0), (64, 2.0), (256, 1.0), (1024, 1.0)])
def test_transport_rate_matches_oracle(pkg, n, spacing):
    say(f"transport_rate n={n} spacing={spacing}")
    c, phi = _fields(pkg, n, spacing=spacing)
    rate, d = pkg.transport_rate(c, pkg.geometry(phi), pkg.make_grid(n, spacing))
    rr, dr = o.transport_rate(c, o.geometry(phi, P), o.make_grid(n, spacing), P)
    assert np.max(np.abs(rate - rr)) < 1e-12 * np.max(np.abs(rr))
    assert abs(d / dr - 1) < 1e-12


@pytest.mark.parametrize("n", SIZES)
def test_one_step_matches_oracle(pkg, n):
    say(f"one step n={n}")
    c, phi = _fields(pkg, n)
    c1 = pkg.step(c, pkg.geometry(phi), pkg.make_grid(n, 1.0), pkg.Parameters(), 0.01)
    r1 = o.step(c, o.geometry(phi, P), o.make_grid(n, 1.0), P, 0.01)
    assert np.max(np.abs(c1 - r1)) < 1e-14


@pytest.mark.parametrize("n", [48, 96, 192, 384, 768])
def test_radix3_grids_match_oracle(pkg, n):
    """N = 3 * 2^m grids (radix-3 first stage): energy, one step, 100 steps."""
    say(f"radix-3 grid n={n}")
    c, phi = _fields(pkg, n)
    shape, grid = pkg.geometry(phi), pkg.make_grid(n, 1.0)
    ro, rg = o.geometry(phi, P), o.make_grid(n, 1.0)
    assert abs(pkg.energy(c, shape, grid) / o.energy(c, ro, rg, P) - 1) < 1e-13
    c1 = pkg.step(c, shape, grid, pkg.Parameters(), 0.01)
    assert np.max(np.abs(c1 - o.step(c, ro, rg, P, 0.01))) < 1e-14
    cg, hg = pkg.run_simulation(c, phi, steps=100)
    cr, hr = o.run_simulation_fast(c, phi, steps=100)
    assert np.max(np.abs(cg - cr)) < 1e-12
    assert np.max(np.abs(hg - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12


def test_nondefault_parameters(pkg):
    say("non-default parameters")
    p = pkg.Parameters(A=1.5, B=6.0, chi=-0.3, P=2.0, ca=0.4, cb=0.1, cc=0.9, kappa=1.3, mobility=0.7, Q=0.9)
    po = o.Parameters(A=1.5, B=6.0, chi=-0.3, P=2.0, ca=0.4, cb=0.1, cc=0.9, kappa=1.3, mobility=0.7, Q=0.9)
    n = 64
    c, phi = _fields(pkg, n)
    cg, hg = pkg.run_simulation(c, phi, dt=0.02, steps=50, spacing=0.8, p=p)
    cr, hr = o.run_simulation_fast(c, phi, dt=0.02, steps=50, spacing=0.8, p=po)
    assert np.max(np.abs(cg - cr)) < 1e-12
    assert np.max(np.abs(hg - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12


def test_known_answer_loop_64(pkg):
    say("known answers 64^2, dt=0.01/0.005/0.0025")
    finals = []
    for dt in ("0.01", "0.005", "0.0025"):
        c0, phi = _fields(pkg, 64, radius=16.0)
        final, hist = pkg.run_simulation(c0, phi, dt=float(dt), steps=round(10 / float(dt)))
        ref = GOLD["runs_64"]["by_dt"][dt]
        assert hist.shape == (round(10 / float(dt)) + 1, 6)
        assert abs(hist[0, 2] / GOLD["runs_64"]["energy_t0"] - 1) < 1e-14
        assert abs(hist[-1, 2] / ref["energy_final"] - 1) < 1e-12
        assert np.max(np.abs(hist[:, 1] - c0.mean())) < 1e-12          # checks.jl:19
        assert np.max(np.diff(hist[:, 2])) < 1e-8                        # checks.jl:20
        assert abs(final.min() - ref["range"][0]) < 1e-10
        assert abs(final.max() - ref["range"][1]) < 1e-10
        finals.append(final)
    r1 = np.sqrt(np.mean((finals[0] - finals[1]) ** 2))
    r2 = np.sqrt(np.mean((finals[1] - finals[2]) ** 2))
    assert abs(r1 / GOLD["runs_64"]["rms_0.01_vs_0.005"] - 1) < 1e-6
    assert abs(r2 / GOLD["runs_64"]["rms_0.005_vs_0.0025"] - 1) < 1e-6


def test_loop_64_field_matches_oracle(pkg):
    say("64^2 1000 steps vs oracle field")
    c0, phi = _fields(pkg, 64, radius=16.0)
    final, hist = pkg.run_simulation(c0, phi, dt=0.01, steps=1000)
    fr, hr = o.run_simulation_fast(c0, phi, dt=0.01, steps=1000)
    assert np.max(np.abs(final - fr)) < 1e-10
    assert np.max(np.abs(hist - hr) / np.maximum(1.0, np.abs(hr))) < 1e-10


def test_known_answer_loop_256(pkg):
    say("known answers 256^2, 1000 steps, and oracle field")
    c0, phi = _fields(pkg, 256, radius=70.0)
    final, hist = pkg.run_simulation(c0, phi, dt=0.01, steps=1000)
    ref = GOLD["run_256"]
    assert abs(hist[0, 2] / ref["energy_t0"] - 1) < 1e-14
    assert abs(hist[-1, 2] / ref["energy_final"] - 1) < 1e-12
    assert np.max(np.abs(hist[:, 1] - c0.mean())) < 1e-13
    assert np.max(np.diff(hist[:, 2])) < 1e-8
    assert abs(final.min() - ref["range"][0]) < 1e-10
    assert abs(final.max() - ref["range"][1]) < 1e-10
    fr, hr = o.run_simulation_fast(c0, phi, dt=0.01, steps=1000)
    assert np.max(np.abs(final - fr)) < 1e-10


def test_stated_field_256_r80_w2(pkg):
    """The plain-method workload as stated (256^2, R=80, w=2, dt=0.01, 1000
    steps, history every 10 steps), FP64, against the oracle field; the
    checks.jl:19-20 assertions on the history."""
    say("256^2 R=80 w=2 1000 steps fp64 vs oracle")
    c0, phi = _fields(pkg, 256, radius=80.0, width=2.0)
    final, hist = pkg.run_simulation(c0, phi, dt=0.01, steps=1000, history_every=10)
    fr, hr = o.run_simulation_fast(c0, phi, dt=0.01, steps=1000)
    assert hist.shape == (101, 6)
    assert np.max(np.abs(final - fr)) < 1e-10
    assert np.max(np.abs(hist - hr[::10]) / np.maximum(1.0, np.abs(hr[::10]))) < 1e-10
    assert np.max(np.abs(hist[:, 1] - c0.mean())) < 1e-13
    assert np.max(np.diff(hist[:, 2])) < 1e-8


@pytest.mark.parametrize("n,steps", [(512, 2000), (1024, 200)])
def test_mixed_mass_drift_fp64_level(pkg, n, steps):
    """Mixed form: the mean composition changes only by FP64 rounding (the
    increment's k = 0 mode is removed in FP64), as checks.jl:19 asks of the
    FP64 path, with a 10x tighter bound; energy non-increasing (checks.jl:20)."""
    say(f"mixed mass drift {n}^2, {steps} steps")
    c0, phi = _fields(pkg, n, radius=0.3 * n, width=2.0)
    _, hist = pkg.run_simulation(c0, phi, dt=0.01, steps=steps, precision="mixed")
    drift = np.max(np.abs(hist[:, 1] - hist[0, 1]))
    assert drift <= 1e-13, drift
    assert abs(hist[0, 1] - c0.mean()) < 1e-14
    assert np.max(np.diff(hist[:, 2])) < 1e-8


def test_mixed_horizon_1024(pkg):
    """Mixed against FP64 on the 1024^2 plain-method workload (R=320, w=2)
    after 1000 steps.  The model amplifies any perturbation by ~1e3 over this
    horizon (an FP64 c0 perturbation of 1e-8 grows to ~9e-6), so the bound is
    2e-5 on max|dc| (measured 9.5e-6 to 1.13e-5 across kernel layouts) and
    1e-6 on the relative energy."""
    say("mixed vs fp64 1024^2, 1000 steps")
    c0, phi = _fields(pkg, 1024, radius=320.0, width=2.0)
    f64, h64 = pkg.run_simulation(c0, phi, dt=0.01, steps=1000, history_every=100)
    fmx, hmx = pkg.run_simulation(c0, phi, dt=0.01, steps=1000, history_every=100, precision="mixed")
    dc = np.max(np.abs(fmx - f64))
    assert dc <= 2e-5, dc
    assert abs(hmx[-1, 2] / h64[-1, 2] - 1) <= 1e-6
    assert np.max(np.abs(hmx[:, 1] - h64[0, 1])) <= 1e-13


def test_checks_jl_identities_on_gpu(pkg):
    say("checks.jl identities against the GPU backend")
    c, phi = _fields(pkg, 64, radius=16.0)
    shape, grid = pkg.geometry(phi), pkg.make_grid(64, 1.0)
    v = np.random.default_rng(8).standard_normal(c.shape)
    eps = 1e-6
    num = (pkg.energy(c + eps * v, shape, grid) - pkg.energy(c - eps * v, shape, grid)) / (2 * eps)
    mu = pkg.chemical_potential(c, shape, grid)
    assert np.isclose(num, np.sum(mu * v), rtol=1e-6, atol=1e-6)                       # checks.jl:11
    rate, diss = pkg.transport_rate(c, shape, grid)
    dcdt = np.real(np.fft.ifft2(rate))
    assert np.isclose(np.sum(mu * dcdt), -diss, rtol=1e-10)                             # checks.jl:13
    assert abs(np.sum(dcdt)) < 1e-10                                                    # checks.jl:14


def test_history_every_and_run_shape(pkg):
    say("history_every")
    c0, phi = _fields(pkg, 64)
    _, h1 = pkg.run_simulation(c0, phi, steps=100)
    _, h10 = pkg.run_simulation(c0, phi, steps=100, history_every=10)
    assert h10.shape == (11, 6)
    assert np.array_equal(h10, h1[::10])
    _, h0 = pkg.run_simulation(c0, phi, steps=0)
    assert h0.shape == (1, 6) and h0[0, 0] == 0.0


def test_graph_replays_match_eager(pkg):
    """Runs of >= 64 steps replay captured blocks of 32 iterations and launch
    the remainder directly; every run length agrees bit for bit with eager
    per-step launches, replay after replay, and its history rows with those
    of a longer run."""
    say("graph vs eager")
    c0, phi = _fields(pkg, 128)
    s = pkg.Session(128)
    s.set_fields(c0, phi)
    s.step(0.01, 200)                       # eager launches
    eager = s.field()
    for _ in range(3):
        cg, _ = pkg.run_simulation(c0, phi, steps=200)   # 6 replays + 8 direct
        assert np.array_equal(cg, eager)
    _, h_all = pkg.run_simulation(c0, phi, steps=200)
    for k in (0, 1, 5, 20, 31, 33, 63, 64, 95):
        say(f"graph vs eager: {k} steps")
        s.set_fields(c0, phi)
        s.step(0.01, k)
        ek = s.field()
        s.set_fields(c0, phi)
        hk = s.run(0.01, k)
        assert np.array_equal(s.field(), ek), k
        assert np.array_equal(hk, h_all[: k + 1]), k
    s.close()


def test_nonfinite_at_step0(pkg):
    say("non-finite c0")
    c0, phi = _fields(pkg, 64)
    c0[3, 5] = np.nan
    with pytest.raises(pkg.NonFiniteCompositionError) as e:
        pkg.run_simulation(c0, phi, steps=10)
    assert e.value.step == 0 and str(e.value) == "Non-finite composition at step 0"


def test_nonfinite_blowup_step_matches_oracle(pkg, tmp_path):
    say("non-finite blow-up")
    c0, phi = _fields(pkg, 32)
    c0[10, 10] = 3.0e3     # the cubic term overflows within a few steps
    with pytest.raises(o.NonFiniteError) as er:
        o.run_simulation(c0, phi, dt=0.01, steps=200)
    with pytest.raises(pkg.NonFiniteCompositionError) as eg:
        pkg.run_simulation(c0, phi, dt=0.01, steps=200, directory=str(tmp_path / "out"))
    assert str(eg.value) == str(er.value)
    assert not (tmp_path / "out").exists()   # no files on error (core_shell.jl:93-101)


def test_nonfinite_long_run_stops_early(pkg):
    """A long captured run that turns non-finite stops within two check blocks
    (2 x 64 graph replays of 32 steps) of the first non-finite step, as the
    reference's loop stops at once (core_shell.jl:93): one million steps at
    32^2 take ~10 s in full, the stopped run well under a second."""
    import time
    say("non-finite blow-up in a 1e6-step run")
    c0, phi = _fields(pkg, 32)
    c0[10, 10] = 3.0e3
    with pytest.raises(pkg.NonFiniteCompositionError) as short:
        pkg.run_simulation(c0, phi, dt=0.01, steps=200)
    t0 = time.perf_counter()
    with pytest.raises(pkg.NonFiniteCompositionError) as long_run:
        pkg.run_simulation(c0, phi, dt=0.01, steps=1_000_000, history_every=1000)
    secs = time.perf_counter() - t0
    assert long_run.value.step == short.value.step
    assert secs < 3.0, secs


def test_results_directory(pkg, tmp_path):
    say("results-*/ writers")
    c0, phi = _fields(pkg, 16, radius=4.0)
    d = tmp_path / "results-16"
    c, hist = pkg.run_simulation(c0, phi, steps=5, directory=str(d))
    comp = np.loadtxt(d / "composition.csv", delimiter=",")
    part = np.loadtxt(d / "particle.csv", delimiter=",")
    h = np.loadtxt(d / "history.csv", delimiter=",")
    assert np.array_equal(comp, c) and np.array_equal(part, phi) and np.array_equal(h, hist)
    assert h.shape == (6, 6)


def test_set_params_matches_fresh_context(pkg):
    """csg_set_params: the captured loop picks up the new parameters (it was
    built with the old ones) and the run equals a fresh context's bit for bit."""
    say("set_params vs fresh context")
    n = 64
    c, phi = _fields(pkg, n)
    p2 = pkg.Parameters(B=5.0, chi=0.7, kappa=1.5)
    s = pkg.Session(n)
    s.set_fields(c, phi)
    s.run(0.01, 100)                     # captures the loop with the defaults
    s.set_params(p2)
    s.set_fields(c, phi)
    h = s.run(0.01, 100)
    f = pkg.Session(n, 1.0, p2)
    f.set_fields(c, phi)
    hf = f.run(0.01, 100)
    assert np.array_equal(s.field(), f.field()) and np.array_equal(h, hf)
    cr, _ = o.run_simulation_fast(c, phi, steps=100, p=o.Parameters(B=5.0, chi=0.7, kappa=1.5))
    assert np.max(np.abs(s.field() - cr)) < 1e-12
    s.close()
    f.close()


def test_copy_bandwidth_is_plausible_and_checks_arguments(pkg):
    """csg_copy_bandwidth: a 64 MiB streaming copy reports a positive rate
    below a generous bound (read + write, GB/s); sizes that are not a
    multiple of 16 bytes, or beyond 16 GiB, and zero repeats are rejected."""
    say("copy bandwidth")
    gbs = pkg.copy_bandwidth(64 << 20, 10)
    assert 100.0 < gbs < 50000.0
    for nbytes, reps in ((24, 1), (0, 1), ((16 << 30) + 16, 1), (1 << 20, 0)):
        with pytest.raises(pkg.CsgError):
            pkg.copy_bandwidth(nbytes, reps)
