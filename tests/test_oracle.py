"""Pins the FP64 oracle (CPU only).

Known answers: check-results.txt:2-7 (via tests/golden/check_results.json).
Identities: checks.jl:7-14.  Closed forms: single-mode linear growth around
c = 1/2 inside the particle (f''(1/2) = -B for h = 1, cubic remainder), the
frozen Nyquist mode (k_d = 0), phi = 0 and uniform fixed points, Parseval
gradient energy including the Nyquist contribution.  numpy.fft against a
brute-force DFT.
"""
import json
import os

import numpy as np
import pytest

from oracle import core_shell_oracle as o
from oracle import julia_rng

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "check_results.json")))
P = o.Parameters()


def test_numpy_fft_matches_bruteforce_dft():
    rng = np.random.default_rng(0)
    for n in (4, 6, 8, 12, 16):
        x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        m = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(m, m) / n)
        ref = W @ x @ W.T
        assert np.max(np.abs(o.fft(x) - ref)) < 1e-12 * n * n
        assert np.max(np.abs(o.ifft(ref) - x)) < 1e-13 * n


def test_fftfreq_matches_definition():
    assert np.array_equal(o.fftfreq(4, 1.0), [0.0, 0.25, -0.5, -0.25])
    assert np.array_equal(o.fftfreq(5, 1.0), [0.0, 0.2, 0.4, -0.4, -0.2])
    g = o.make_grid(8, 1.0)
    assert g["kx"][4, 0] == 0.0 and g["k2"][4, 0] == np.pi ** 2   # Nyquist: zeroed / kept


def test_known_answer_energy_t0_64():
    c, phi = o.initial_fields(64, radius=16.0)
    F0 = o.energy(c, o.geometry(phi, P), o.make_grid(64, 1.0), P)
    assert abs(F0 / GOLD["runs_64"]["energy_t0"] - 1) < 2e-15


def test_known_answer_energy_t0_256():
    c, phi = o.initial_fields(256)
    F0 = o.energy(c, o.geometry(phi, P), o.make_grid(256, 1.0), P)
    assert abs(F0 / GOLD["run_256"]["energy_t0"] - 1) < 2e-15


def test_rng_stream_is_prefix_consistent():
    a = julia_rng.MersenneTwister(5749).rand(16, 16).ravel(order="F")    # cache path
    b = julia_rng.MersenneTwister(5749).rand(64, 64).ravel(order="F")    # bulk + cache path
    assert np.array_equal(a, b[:256])
    assert np.all((b >= 0) & (b < 1))


def test_rng_seed_hash_limbs():
    # a seed above 2^32 hashes two limbs; different seeds give different streams
    a = julia_rng.MersenneTwister(5749).rand(4)
    b = julia_rng.MersenneTwister(5749 + (1 << 32)).rand(4)
    c = julia_rng.MersenneTwister(-5750).rand(4)
    assert not np.array_equal(a, b) and not np.array_equal(a, c)


def _run_64(dt):
    c, phi = o.initial_fields(64, radius=16.0)
    return c, o.run_simulation(c, phi, dt=dt, steps=round(10 / dt))


def test_known_answer_loop_64_dt001():
    c0, (final, hist) = _run_64(0.01)
    ref = GOLD["runs_64"]["by_dt"]["0.01"]
    drift = np.max(np.abs(hist[:, 1] - c0.mean()))
    assert drift < 1e-12
    assert abs(hist[-1, 2] / ref["energy_final"] - 1) < 1e-13
    assert abs(final.min() - ref["range"][0]) < 1e-12 and abs(final.max() - ref["range"][1]) < 1e-12
    assert np.max(np.diff(hist[:, 2])) < 1e-8
    assert hist.shape == (1001, 6) and hist[-1, 0] == 1000 * 0.01


@pytest.mark.slow
def test_known_answer_dt_refinement_64():
    finals = []
    for dt in ("0.01", "0.005", "0.0025"):
        c0, phi = o.initial_fields(64, radius=16.0)
        final, hist = o.run_simulation_fast(c0, phi, dt=float(dt), steps=round(10 / float(dt)))
        ref = GOLD["runs_64"]["by_dt"][dt]
        assert abs(hist[-1, 2] / ref["energy_final"] - 1) < 1e-13
        assert np.max(np.abs(hist[:, 1] - c0.mean())) < 1e-12
        finals.append(final)
    r1 = np.sqrt(np.mean((finals[0] - finals[1]) ** 2))
    r2 = np.sqrt(np.mean((finals[1] - finals[2]) ** 2))
    assert abs(r1 / GOLD["runs_64"]["rms_0.01_vs_0.005"] - 1) < 1e-9
    assert abs(r2 / GOLD["runs_64"]["rms_0.005_vs_0.0025"] - 1) < 1e-9


@pytest.mark.slow
def test_known_answer_loop_256():
    c0, phi = o.initial_fields(256)
    final, hist = o.run_simulation_fast(c0, phi, dt=0.01, steps=1000)
    ref = GOLD["run_256"]
    assert abs(hist[0, 2] / ref["energy_t0"] - 1) < 2e-15
    assert abs(hist[-1, 2] / ref["energy_final"] - 1) < 1e-13
    assert abs(final.min() - ref["range"][0]) < 1e-12 and abs(final.max() - ref["range"][1]) < 1e-12
    assert np.max(np.abs(hist[:, 1] - c0.mean())) < 1e-12
    assert np.max(np.diff(hist[:, 2])) < 1e-8


def test_checks_jl_identities():
    c, phi = o.initial_fields(64, radius=16.0)
    grid, shape = o.make_grid(64, 1.0), o.geometry(phi, P)
    v = np.random.default_rng(8).standard_normal(c.shape)
    eps = 1e-6
    num = (o.energy(c + eps * v, shape, grid, P) - o.energy(c - eps * v, shape, grid, P)) / (2 * eps)
    ana = np.sum(o.chemical_potential(c, shape, grid, P) * v)
    assert np.isclose(num, ana, rtol=1e-6, atol=1e-6)
    rate, diss = o.transport_rate(c, shape, grid, P)
    dcdt = np.real(o.ifft(rate))
    assert np.isclose(np.sum(o.chemical_potential(c, shape, grid, P) * dcdt), -diss, rtol=1e-10)
    assert abs(np.sum(dcdt)) < 1e-10


def test_planted_mistake_fails_identity():
    """A dropped mobility factor in the flux breaks the dissipation identity."""
    c, phi = o.initial_fields(32, radius=8.0)
    grid, shape = o.make_grid(32, 1.0), o.geometry(phi, P)
    bad = dict(shape, M=np.ones_like(shape["M"]))
    rate, _ = o.transport_rate(c, bad, grid, P)
    _, diss = o.transport_rate(c, shape, grid, P)
    lhs = np.sum(o.chemical_potential(c, shape, grid, P) * np.real(o.ifft(rate)))
    assert not np.isclose(lhs, -diss, rtol=1e-3)


def _growth_factor(k, dt, p):
    k2 = k * k
    return 1 - dt * k2 * (-p.B + 2 * p.kappa * k2) / (1 + 2 * p.Q * p.kappa * k2 * k2 * dt)


@pytest.mark.parametrize("mx,my", [(3, 0), (0, 5), (2, 7)])
def test_single_mode_growth_closed_form(mx, my):
    n, dt, delta, sp = 32, 0.01, 1e-6, 1.0
    i = np.arange(n).reshape(n, 1)
    j = np.arange(n).reshape(1, n)
    c = 0.5 + delta * np.cos(2 * np.pi * (mx * i + my * j) / n)
    phi = np.ones((n, n))
    grid, shape = o.make_grid(n, sp), o.geometry(phi, P)
    c1 = o.step(c, shape, grid, P, dt)
    k = 2 * np.pi * np.hypot(mx, my) / (n * sp)
    ratio = (c1 - 0.5)[np.abs(c - 0.5) > 0.5 * delta] / (c - 0.5)[np.abs(c - 0.5) > 0.5 * delta]
    assert np.allclose(ratio, _growth_factor(k, dt, P), rtol=1e-9)


def test_nyquist_mode_is_frozen():
    n = 16
    i = np.arange(n).reshape(n, 1)
    c = 0.5 + 1e-3 * (-1.0) ** i * np.ones((1, n))
    phi = np.ones((n, n))
    c1 = o.step(c, o.geometry(phi, P), o.make_grid(n, 1.0), P, 0.01)
    assert np.max(np.abs(c1 - c)) < 1e-15


def test_fixed_points():
    n = 16
    c = np.random.default_rng(1).uniform(0.2, 0.8, (n, n))
    phi0 = np.zeros((n, n))   # M = 0: nothing moves
    c1 = o.step(c, o.geometry(phi0, P), o.make_grid(n, 1.0), P, 0.01)
    assert np.max(np.abs(c1 - c)) < 1e-15
    cu = np.full((n, n), 0.3)  # uniform c in a uniform particle
    c2 = o.step(cu, o.geometry(np.ones((n, n)), P), o.make_grid(n, 1.0), P, 0.01)
    assert np.max(np.abs(c2 - cu)) < 1e-15


def test_energy_closed_form_with_nyquist():
    n, d = 16, 1e-2
    i = np.arange(n).reshape(n, 1)
    phi = np.ones((n, n))
    for m, kk in ((3, 2 * np.pi * 3 / n), (n // 2, np.pi)):
        c = 0.5 + d * np.cos(2 * np.pi * m * i / n) * np.ones((1, n))
        F = o.energy(c, o.geometry(phi, P), o.make_grid(n, 1.0), P)
        bulk = np.sum(P.B * c ** 2 * (c - 1) ** 2)
        mean_sq = 0.5 if m != n // 2 else 1.0   # <cos^2>: 1/2, or 1 for (-1)^i
        grad = P.kappa * kk ** 2 * d * d * mean_sq * n * n
        assert np.isclose(F, bulk + grad, rtol=1e-12)
