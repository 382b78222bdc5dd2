"""Ensembles (csg_batch_*): every member equals the single-field path on the
same inputs bit for bit (same kernels, one instance per workgroup row), and the
FP64 ensemble matches the oracle's run_simulation; argument checks on the CPU.
"""
import ctypes
import dataclasses
import sys

import numpy as np
import pytest

from oracle import core_shell_oracle as o


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _members(pkg):
    d = pkg.Parameters()
    spec = [(64, 5749, 1.0, 1.0), (48, 11, 1.2, 0.8), (96, 12, 0.7, 1.3), (64, 13, 1.3, 0.7),
            (128, 14, 0.9, 1.1), (16, 15, 1.0, 1.0), (48, 16, 1.1, 0.9)]
    out = []
    for n, seed, fp, fchi in spec:
        out.append(pkg.Member(n=n, radius=0.3 * n, width=2.0, seed=seed,
                              p=dataclasses.replace(d, P=d.P * fp, chi=d.chi * fchi)))
    return out


# ---- CPU -------------------------------------------------------------------

def test_batch_arguments_rejected_without_gpu(pkg):
    from csg_amd import _capi
    L = pkg.lib()
    h = _capi._VP()
    ns = (ctypes.c_int32 * 3)(64, 100, 128)
    assert L.csg_batch_create(h, 3, ns, None, None, 0, 0) == _capi.CSG_ERR_UNSUPPORTED
    assert "not supported" in L.csg_last_error().decode()
    ok = (ctypes.c_int32 * 2)(64, 128)
    assert L.csg_batch_create(h, 0, ok, None, None, 0, 0) == _capi.CSG_ERR_ARG
    assert L.csg_batch_create(h, 2, None, None, None, 0, 0) == _capi.CSG_ERR_ARG
    assert L.csg_batch_create(None, 2, ok, None, None, 0, 0) == _capi.CSG_ERR_ARG
    assert L.csg_batch_create(h, 2, ok, None, None, 5, 0) == _capi.CSG_ERR_ARG
    sp = np.array([1.0, -1.0])
    assert L.csg_batch_create(h, 2, ok, _capi.ptr(sp), None, 0, 0) == _capi.CSG_ERR_ARG
    assert L.csg_batch_run(None, 0.01, 10, 1, None) == _capi.CSG_ERR_ARG
    assert L.csg_batch_nonfinite_step(None, 0) == -1
    assert L.csg_batch_destroy(None) == 0


def test_ensemble_members_workload(pkg):
    ms = pkg.ensemble_members(256)
    assert len(ms) == 256 and ms == pkg.ensemble_members(256)   # deterministic
    d = pkg.Parameters()
    sizes = [m.n for m in ms]
    assert set(sizes) == {128, 256, 384, 512}
    assert min(sizes.count(s) for s in (128, 256, 384, 512)) > 40
    for i, m in enumerate(ms):
        assert m.radius == 0.3 * m.n and m.width == 2.0 and m.seed == 5749 + i
        assert 0.7 * d.P <= m.p.P <= 1.3 * d.P and 0.7 * d.chi <= m.p.chi <= 1.3 * d.chi
        assert dataclasses.replace(m.p, P=d.P, chi=d.chi) == d


# ---- GPU -------------------------------------------------------------------

def _single(pkg, m, fields, steps, precision, dt=0.01):
    c0, phi = fields
    s = pkg.Session(m.n, m.spacing, m.p, precision)
    s.set_fields(c0, phi)
    h = s.run(dt, steps)
    c = s.field()
    s.close()
    return c, h


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("steps", [100, 7])
def test_ensemble_members_equal_single_runs(pkg, precision, steps):
    """Uneven sizes (16..128, radix 2 and 3), per-member P/chi and seeds;
    100 steps use the captured graph plus an eager tail, 7 steps only eager."""
    ms = _members(pkg)
    say(f"ensemble vs single, {precision}, {steps} steps, sizes {[m.n for m in ms]}")
    fields = [pkg.member_fields(m) for m in ms]
    ens = pkg.Ensemble([m.n for m in ms], params=[m.p for m in ms], precision=precision)
    for i, (c0, phi) in enumerate(fields):
        ens.set_fields(i, c0, phi)
    hs = ens.run(0.01, steps)
    for i, m in enumerate(ms):
        c1, h1 = _single(pkg, m, fields[i], steps, precision)
        assert np.array_equal(ens.field(i), c1), (i, m.n)
        assert np.array_equal(hs[i], h1), (i, m.n)
    # a second run continues from the resident fields, like csg_run
    hs2 = ens.run(0.01, 3)
    assert np.array_equal(hs2[0][0, 1:], hs[0][-1, 1:])
    ens.close()


@pytest.mark.gpu
def test_ensemble_fp64_matches_oracle(pkg):
    ms = _members(pkg)[:4]
    say("ensemble fp64 vs oracle, 200 steps")
    fields = [pkg.member_fields(m) for m in ms]
    res = pkg.run_ensemble([f[0] for f in fields], [f[1] for f in fields], steps=200,
                           params=[m.p for m in ms], precision="fp64")
    for (c, h), m, (c0, phi) in zip(res, ms, fields):
        po = o.Parameters(**dataclasses.asdict(m.p))
        cr, hr = o.run_simulation_fast(np.asarray(c0), np.asarray(phi), steps=200, p=po)
        assert np.max(np.abs(c - cr)) < 1e-12, m.n
        assert np.max(np.abs(h - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12, m.n


@pytest.mark.gpu
def test_ensemble_history_every_and_spacing(pkg):
    say("ensemble history_every=5, spacing 0.5/2.0")
    d = pkg.Parameters()
    ns, sps = [32, 64], [0.5, 2.0]
    fields = [pkg.initial_fields(n, spacing=sp, radius=0.3 * n * sp, width=2.0 * sp) for n, sp in zip(ns, sps)]
    ens = pkg.Ensemble(ns, spacings=sps, params=[d, d], precision="fp64")
    for i, (c0, phi) in enumerate(fields):
        ens.set_fields(i, c0, phi)
    hs = ens.run(0.01, 70, history_every=5)
    for i, (n, sp) in enumerate(zip(ns, sps)):
        s = pkg.Session(n, sp, d, "fp64")
        s.set_fields(*fields[i])
        h = s.run(0.01, 70, history_every=5)
        assert hs[i].shape == (15, 6) and np.array_equal(hs[i], h)
        assert np.array_equal(ens.field(i), s.field())
        s.close()
    ens.close()


@pytest.mark.gpu
def test_ensemble_nonfinite_member_does_not_stop_others(pkg):
    say("ensemble with a non-finite member")
    ms = _members(pkg)[:3]
    fields = [pkg.member_fields(m) for m in ms]
    bad = np.array(fields[1][0], order="F", copy=True)
    bad[3, 5] = np.nan
    ens = pkg.Ensemble([m.n for m in ms], params=[m.p for m in ms], precision="fp64")
    for i, (c0, phi) in enumerate(fields):
        ens.set_fields(i, bad if i == 1 else c0, phi)
    with pytest.raises(pkg.EnsembleNonFiniteError) as ei:
        ens.run(0.01, 40)
    assert ei.value.steps == [-1, 0, -1] and ei.value.members == [1]
    for i in (0, 2):
        c1, _ = _single(pkg, ms[i], fields[i], 40, "fp64")
        assert np.array_equal(ens.field(i), c1)
    ens.close()


def _c3_members(pkg):
    """Two members of each ensemble size (128, 256, 384, 512): R = 0.3 N,
    w = 2, P and chi within +-30 % of the defaults, distinct seeds."""
    d = pkg.Parameters()
    spec = [(128, 0.72, 1.25), (128, 1.28, 0.75), (256, 0.85, 1.1), (256, 1.15, 0.9),
            (384, 0.7, 1.3), (384, 1.3, 0.7), (512, 0.95, 1.2), (512, 1.2, 0.8)]
    return [pkg.Member(n=n, radius=0.3 * n, width=2.0, seed=7001 + i,
                       p=dataclasses.replace(d, P=d.P * fp, chi=d.chi * fc))
            for i, (n, fp, fc) in enumerate(spec)]


@pytest.mark.gpu
def test_ensemble_c3_sizes_match_oracle(pkg):
    """Ensemble of all four C3 sizes: FP64 members against the oracle's
    run_simulation (50 steps, <= 1e-12), mixed members against their FP64
    twins, and the mixed mean composition conserved at FP64 level."""
    ms = _c3_members(pkg)
    steps = 50
    say(f"C3 sizes ensemble vs oracle, {steps} steps, sizes {[m.n for m in ms]}")
    fields = [pkg.member_fields(m) for m in ms]
    out = {}
    for prec in ("fp64", "mixed"):
        out[prec] = pkg.run_ensemble([f[0] for f in fields], [f[1] for f in fields], steps=steps,
                                     params=[m.p for m in ms], precision=prec)
    for i, m in enumerate(ms):
        c0, phi = fields[i]
        po = o.Parameters(**dataclasses.asdict(m.p))
        cr, hr = o.run_simulation_fast(np.asarray(c0), np.asarray(phi), steps=steps, p=po)
        c64, h64 = out["fp64"][i]
        cmx, hmx = out["mixed"][i]
        assert np.max(np.abs(c64 - cr)) < 1e-12, (i, m.n)
        assert np.max(np.abs(h64 - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12, (i, m.n)
        assert np.max(np.abs(cmx - c64)) < 1e-5, (i, m.n)
        assert np.max(np.abs(hmx[:, 2] - h64[:, 2]) / np.abs(h64[:, 2])) < 1e-5, (i, m.n)
        assert np.max(np.abs(hmx[:, 1] - hmx[0, 1])) <= 1e-13, (i, m.n)


@pytest.mark.gpu
def test_ensemble_c3_full_workload_properties(pkg):
    """The full uneven ensemble (256 members, N in {128, 256, 384, 512},
    2000 steps, mixed form): every member finite, its mean composition kept
    to 1e-13 and its energy non-increasing (checks.jl:19-20 per member)."""
    ms = pkg.ensemble_members(256)
    say("C3 full ensemble, 256 members, 2000 steps, mixed")
    fields = [pkg.member_fields(m) for m in ms]
    ens = pkg.Ensemble([m.n for m in ms], params=[m.p for m in ms], precision="mixed")
    for i, (c0, phi) in enumerate(fields):
        ens.set_fields(i, c0, phi)
    hs = ens.run(0.01, 2000, history_every=10)
    for i, (m, h) in enumerate(zip(ms, hs)):
        c0 = fields[i][0]
        assert h.shape == (201, 6)
        assert np.all(np.isfinite(h)), i
        assert abs(h[0, 1] - np.mean(c0)) < 1e-14, i
        assert np.max(np.abs(h[:, 1] - h[0, 1])) <= 1e-13, (i, m.n)
        assert np.max(np.diff(h[:, 2])) < 1e-8, (i, m.n)
    for i in (0, 100, 255):
        c = ens.field(i)
        assert np.all(np.isfinite(c)) and c.shape == (ms[i].n, ms[i].n)
    ens.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_persistent_ensemble_matches_oracle_and_batched(pkg, precision):
    """The persistent kernel (every pass of every member in one launch per
    32 steps, per-member counters) on all four C3 sizes: FP64 members against
    the oracle to 1e-12 after 70 steps (two captured blocks plus an eager
    tail), both precisions against the batched mode, mean composition kept."""
    ms = _c3_members(pkg)
    steps = 70
    say(f"persistent ensemble {precision}, {steps} steps, sizes {[m.n for m in ms]}")
    fields = [pkg.member_fields(m) for m in ms]
    out = {}
    for persistent in (False, True):
        out[persistent] = pkg.run_ensemble([f[0] for f in fields], [f[1] for f in fields], steps=steps,
                                           params=[m.p for m in ms], precision=precision,
                                           persistent=persistent)
    tol = 1e-12 if precision == "fp64" else 1e-6
    for i, m in enumerate(ms):
        (cp, hp), (cb, hb) = out[True][i], out[False][i]
        assert hp.shape == hb.shape == (steps + 1, 6)
        assert np.max(np.abs(cp - cb)) < tol, (i, m.n)
        assert np.max(np.abs(hp[:, 2] - hb[:, 2]) / np.abs(hb[:, 2])) < tol, (i, m.n)
        assert np.max(np.abs(hp[:, 1] - hp[0, 1])) <= 1e-13, (i, m.n)
        if precision == "fp64":
            po = o.Parameters(**dataclasses.asdict(m.p))
            cr, hr = o.run_simulation_fast(np.asarray(fields[i][0]), np.asarray(fields[i][1]), steps=steps, p=po)
            assert np.max(np.abs(cp - cr)) < 1e-12, (i, m.n)
            assert np.max(np.abs(hp - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12, (i, m.n)


@pytest.mark.gpu
def test_persistent_ensemble_history_every_and_second_run(pkg):
    """history_every and a second run continuing from the resident fields,
    persistent mode against batched mode (FP64)."""
    ms = _c3_members(pkg)[::2]
    say("persistent ensemble history_every=7, two runs")
    fields = [pkg.member_fields(m) for m in ms]
    res = []
    for persistent in (False, True):
        ens = pkg.Ensemble([m.n for m in ms], params=[m.p for m in ms], precision="fp64", persistent=persistent)
        for i, (c0, phi) in enumerate(fields):
            ens.set_fields(i, c0, phi)
        h1 = ens.run(0.01, 40, history_every=7)
        h2 = ens.run(0.01, 5)
        res.append((h1, h2, [ens.field(i) for i in range(len(ms))]))
        ens.close()
    for i in range(len(ms)):
        assert res[1][0][i].shape == (6, 6)
        assert np.max(np.abs(res[1][0][i] - res[0][0][i]) / np.maximum(1.0, np.abs(res[0][0][i]))) < 1e-12
        assert np.max(np.abs(res[1][1][i] - res[0][1][i]) / np.maximum(1.0, np.abs(res[0][1][i]))) < 1e-12
        assert np.max(np.abs(res[1][2][i] - res[0][2][i])) < 1e-12


def test_persistent_mode_rejects_other_sizes_without_gpu(pkg):
    from csg_amd import _capi
    assert pkg.lib().csg_batch_set_persistent(None, 1) == _capi.CSG_ERR_ARG
