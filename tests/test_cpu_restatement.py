"""Pins the C/OpenMP restatement of the time loop (oracle/cs_cpu.c), the
all-cores CPU baseline of bench.py (CPU only).

Against the numpy oracle (itself pinned by tests/test_oracle.py) on the same
inputs, against the reference's known answers of check-results.txt:2 (64^2,
1000 steps), across thread counts, and on the non-finite stop.
"""
import json
import os

import numpy as np
import pytest

from oracle import core_shell_oracle as o
from oracle import cpu_restatement as cr

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "check_results.json")))


@pytest.mark.parametrize("n,spacing,steps", [(16, 1.0, 30), (64, 1.0, 20), (128, 0.5, 10)])
def test_matches_numpy_oracle(n, spacing, steps):
    c0, phi = o.initial_fields(n, spacing=spacing, radius=0.25 * n * spacing, width=4.0)
    c, h = cr.run_simulation(c0, phi, dt=0.01, steps=steps, spacing=spacing, threads=2)
    cr_, hr = o.run_simulation(c0, phi, dt=0.01, steps=steps, spacing=spacing)
    assert np.max(np.abs(c - cr_)) < 1e-13
    assert np.max(np.abs(h - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12


def test_nondefault_parameters():
    p = o.Parameters(A=1.5, B=6.0, chi=-0.3, P=2.0, ca=0.4, cb=0.1, cc=0.9, kappa=1.3, mobility=0.7, Q=0.9)
    c0, phi = o.initial_fields(32, radius=8.0)
    c, h = cr.run_simulation(c0, phi, dt=0.02, steps=25, spacing=0.8, p=p, threads=2)
    cr_, hr = o.run_simulation(c0, phi, dt=0.02, steps=25, spacing=0.8, p=p)
    assert np.max(np.abs(c - cr_)) < 1e-13
    assert np.max(np.abs(h - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12


def test_known_answers_64():
    """check-results.txt:2: F(t=0), F(t=10), mass drift and final range of the
    64^2 run at dt = 0.01 (1000 steps)."""
    g = GOLD["runs_64"]
    ref = g["by_dt"]["0.01"]
    c0, phi = o.initial_fields(64, radius=g["radius"], width=g["width"], seed=g["seed"])
    c, h = cr.run_simulation(c0, phi, dt=0.01, steps=1000, threads=4)
    assert abs(h[0, 2] / g["energy_t0"] - 1) < 1e-14
    assert abs(h[-1, 2] / ref["energy_final"] - 1) < 1e-12
    assert np.max(np.abs(h[:, 1] - np.mean(c0))) < 1e-12            # checks.jl:19
    assert np.max(np.diff(h[:, 2])) < 1e-8                           # checks.jl:20
    assert abs(c.min() - ref["range"][0]) < 1e-10 and abs(c.max() - ref["range"][1]) < 1e-10


def test_thread_count_changes_only_rounding():
    c0, phi = o.initial_fields(64, radius=16.0)
    c1, h1 = cr.run_simulation(c0, phi, steps=15, threads=1)
    c3, h3 = cr.run_simulation(c0, phi, steps=15, threads=3)
    assert np.max(np.abs(c1 - c3)) < 1e-14
    assert np.max(np.abs(h1 - h3) / np.maximum(1.0, np.abs(h1))) < 1e-13


def test_nonfinite_step_matches_oracle():
    c0, phi = o.initial_fields(32, radius=8.0)
    c0[10, 10] = 3.0e3     # the cubic term overflows within a few steps
    with pytest.raises(o.NonFiniteError) as er:
        o.run_simulation(c0, phi, dt=0.01, steps=200)
    with pytest.raises(cr.NonFiniteError) as ec:
        cr.run_simulation(c0, phi, dt=0.01, steps=200, threads=2)
    assert str(ec.value) == str(er.value)


def test_rejects_grids_that_are_not_powers_of_two():
    c0, phi = o.initial_fields(48, radius=12.0)
    with pytest.raises(RuntimeError):
        cr.run_simulation(c0, phi, steps=1)
