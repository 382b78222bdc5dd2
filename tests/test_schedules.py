"""Kernel schedules: every compiled candidate of every step pass computes the
same step (oracle parity), the runtime tuner keeps the resident field and
selects valid candidates, ensembles follow per-size selections, and the tuned
choice cache round-trips."""
import sys

import numpy as np
import pytest

from oracle import core_shell_oracle as o


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fields(pkg, n):
    c, phi = pkg.initial_fields(n, radius=0.3 * n, width=2.0)
    return np.asarray(c), np.asarray(phi)


# ---- CPU -------------------------------------------------------------------

def test_schedule_cache_roundtrip(pkg, tmp_path, monkeypatch):
    monkeypatch.setenv("CSG_SCHEDULE_CACHE", str(tmp_path / "s.json"))
    assert pkg.schedules.load(1024, "fp64") is None
    pkg.schedules.save(1024, "fp64", (1, 0, 2, 0), 55.5)
    pkg.schedules.save(512, "mixed", [0, 1, 0, 1], 20.0)
    assert pkg.schedules.load(1024, "fp64") == (1, 0, 2, 0)
    assert pkg.schedules.load(512, "mixed") == (0, 1, 0, 1)
    assert pkg.schedules.load(512, "fp64") is None


def test_schedule_cache_rejects_a_changed_candidate_table(pkg, tmp_path, monkeypatch):
    """An entry saved with the candidate table is ignored once the same
    indices select different schedules (a rebuilt library)."""
    monkeypatch.setenv("CSG_SCHEDULE_CACHE", str(tmp_path / "s.json"))
    a = {"E": 8, "W": 4, "G": 1, "PF": 0, "X": 0, "T": 1}
    b = {"E": 4, "W": 2, "G": 1, "PF": 0, "X": 2, "T": 1}
    info = [[a, b], [a], [a], [a]]
    pkg.schedules.save(1024, "fp64", (1, 0, 0, 0), 50.0, info)
    assert pkg.schedules.load(1024, "fp64", info) == (1, 0, 0, 0)
    assert pkg.schedules.load(1024, "fp64", [[b, a], [a], [a], [a]]) is None   # reordered table
    assert pkg.schedules.load(1024, "fp64", [[a], [a], [a], [a]]) is None      # index out of range
    assert pkg.schedules.load(1024, "fp64") == (1, 0, 0, 0)                     # no table: trusted


# ---- GPU -------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("n", [128, 192, 256, 384, 512, 768, 1024])
def test_every_candidate_matches_oracle_fp64(pkg, n):
    c, phi = _fields(pkg, n)
    steps = 10
    cr, hr = o.run_simulation_fast(c, phi, steps=steps)
    s = pkg.Session(n)
    info = s.schedule_info()
    assert [len(x) for x in info] and all(1 <= len(x) <= 7 for x in info)
    for q in range(4):
        for i, sc in enumerate(info[q]):
            sel = [0, 0, 0, 0]
            sel[q] = i
            say(f"n={n} fp64 pass {q} candidate {i} {sc}")
            s.set_schedule(sel)
            assert s.schedule == tuple(sel)
            s.set_fields(c, phi)
            h = s.run(0.01, steps)
            assert np.max(np.abs(s.field() - cr)) < 1e-12
            assert np.max(np.abs(h - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [128, 256, 384, 512, 1024])
def test_every_candidate_mixed_within_tolerance(pkg, n):
    """Mixed form: FP32 bulk work, so each candidate is held to the mixed-form
    tolerance against the FP64 oracle over 10 steps."""
    c, phi = _fields(pkg, n)
    cr, _ = o.run_simulation_fast(c, phi, steps=10)
    s = pkg.Session(n, precision="mixed")
    info = s.schedule_info()
    for q in range(4):
        for i in range(len(info[q])):
            sel = [0, 0, 0, 0]
            sel[q] = i
            say(f"n={n} mixed pass {q} candidate {i}")
            s.set_schedule(sel)
            s.set_fields(c, phi)
            s.run(0.01, 10)
            assert np.max(np.abs(s.field() - cr)) < 5e-6
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n, precision, tol", [(2048, "fp64", 1e-12), (4096, "fp64", 1e-12),
                                               (2048, "mixed", 5e-6), (4096, "mixed", 5e-6)])
def test_large_candidates_agree_with_default(pkg, n, precision, tol):
    """At 2048^2 and 4096^2 (too large for a quick oracle run of every
    candidate) each compiled candidate must reproduce the default schedule's
    field after 3 steps; the default itself is held to the oracle by the
    large-field tests.  FP64 to 1e-12; mixed to its own tolerance, since the
    FP32 bulk rounds differently per radix plan and exchange order."""
    c, phi = _fields(pkg, n)
    s = pkg.Session(n, precision=precision)
    s.set_fields(c, phi)
    s.run(0.01, 3)
    ref = s.field()
    info = s.schedule_info()
    for q in range(4):
        for i in range(1, len(info[q])):
            sel = [0, 0, 0, 0]
            sel[q] = i
            say(f"n={n} {precision} pass {q} candidate {i} {info[q][i]}")
            s.set_schedule(sel)
            s.set_fields(c, phi)
            s.run(0.01, 3)
            assert np.max(np.abs(s.field() - ref)) < tol
    s.close()


@pytest.mark.gpu
def test_set_schedule_rejects_bad_index(pkg):
    s = pkg.Session(64)
    with pytest.raises(pkg.CsgError):
        s.set_schedule((0, 0, 7, 0))
    with pytest.raises(pkg.CsgError):
        s.set_schedule((-1, 0, 0, 0))
    assert s.schedule == (0, 0, 0, 0)
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_keeps_field_and_selects_valid(pkg, precision):
    n = 512
    c, phi = _fields(pkg, n)
    say(f"autotune n={n} {precision}")
    s = pkg.Session(n, precision=precision)
    s.set_fields(c, phi)
    s.run(0.01, 5)
    before = s.field()
    sel, us = s.autotune(0.01, 1)
    info = s.schedule_info()
    assert all(0 <= sel[q] < len(info[q]) for q in range(4)) and us > 0
    assert s.schedule == tuple(sel)
    assert np.array_equal(s.field(), before)
    # the run after tuning continues from the kept field (5 + 10 steps in all)
    s.run(0.01, 10)
    cr, _ = o.run_simulation_fast(c, phi, steps=15)
    tol = 1e-12 if precision == "fp64" else 5e-6
    assert np.max(np.abs(s.field() - cr)) < tol
    s.close()


@pytest.mark.gpu
def test_cached_schedule_is_applied(pkg, tmp_path, monkeypatch):
    monkeypatch.setenv("CSG_SCHEDULE_CACHE", str(tmp_path / "s.json"))
    s = pkg.Session(256)
    last = tuple(len(x) - 1 for x in s.schedule_info())
    s.close()
    pkg.schedules.save(256, "fp64", last, 1.0)
    s = pkg.Session(256, schedule="cached")
    assert s.schedule == last
    s.close()
    c, phi = _fields(pkg, 256)
    c1, h1 = pkg.run_simulation(c, phi, steps=20, schedule="cached")
    c0, h0 = pkg.run_simulation(c, phi, steps=20)
    assert np.max(np.abs(c1 - c0)) < 1e-13


@pytest.mark.gpu
def test_ensemble_schedules_follow_single_sessions(pkg):
    """Members of a class with a non-default selection equal single sessions
    with the same selection bit for bit; the ensemble tuner keeps the fields."""
    say("ensemble schedules")
    d = pkg.Parameters()
    ms = [pkg.Member(n=128, radius=38.4, seed=3), pkg.Member(n=192, radius=57.6, seed=4),
          pkg.Member(n=128, radius=38.4, seed=5)]
    fields = [pkg.member_fields(m) for m in ms]
    ens = pkg.Ensemble([m.n for m in ms], params=[d] * 3, precision="fp64")
    for i, (c0, phi) in enumerate(fields):
        ens.set_fields(i, c0, phi)
    sel = (1, 1, 1, 1)
    ens.set_schedule(128, sel)
    assert ens.schedule(128) == sel and ens.schedule(192) == (0, 0, 0, 0)
    ens.run(0.01, 70)
    for i, m in enumerate(ms):
        s = pkg.Session(m.n, schedule=sel if m.n == 128 else None)
        s.set_fields(*fields[i])
        s.run(0.01, 70)
        assert np.array_equal(ens.field(i), s.field()), i
        s.close()
    before = [ens.field(i) for i in range(3)]
    tuned = ens.autotune(0.01, 1)
    assert set(tuned) == {128, 192}
    for i in range(3):
        assert np.array_equal(ens.field(i), before[i])
    ens.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nontemporal_outputs_are_bit_identical(pkg, monkeypatch, precision):
    """The cache hints (non-temporal stores, column loads and phi loads, chosen from the
    device footprint, forced here with CSG_NONTEMPORAL=0..7) change where data
    is cached, never values: fields and histories agree bit for bit, single
    fields and ensembles alike."""
    n, steps = 256, 40
    c, phi = _fields(pkg, n)
    out = {}
    for v in ("0", "1", "2", "3", "4", "7"):
        monkeypatch.setenv("CSG_NONTEMPORAL", v)
        s = pkg.Session(n, precision=precision)
        s.set_fields(c, phi)
        h = s.run(0.01, steps)
        ens = pkg.Ensemble([n, 128], precision=precision)
        ens.set_fields(0, c, phi)
        c1, phi1 = _fields(pkg, 128)
        ens.set_fields(1, c1, phi1)
        ens.run(0.01, steps)
        out[v] = (s.field(), h, ens.field(0), ens.field(1))
        s.close()
        ens.close()
    for v in ("1", "2", "3", "4", "7"):
        for a, b in zip(out["0"], out[v]):
            assert np.array_equal(a, b), v
