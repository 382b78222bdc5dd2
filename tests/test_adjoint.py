"""Adjoint of the time loop.

CPU: the oracle's adjoint (oracle/adjoint_oracle.py) against central finite
differences of the oracle's own forward loss.
GPU: the HIP adjoint (csg_adjoint_forward / csg_adjoint_backward) against the
oracle adjoint (same inputs, tight tolerance; 64^2 x 200 steps and the
workload's 512^2 grid over 40 steps) and against central finite
differences of the GPU forward loss, including the 512^2, 2000-step,
45-step-checkpoint workload on 32 random directions.
Loss: L = 1/2 ||c(T) - c*||^2 + lam F(c(T)), c* = forward run with kappa, P,
chi raised 10 %.
"""
import dataclasses
import sys

import numpy as np
import pytest

from oracle import adjoint_oracle as ad
from oracle import core_shell_oracle as o

P = o.Parameters()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup_oracle(n, radius, steps, dt, lam):
    c0, phi = o.initial_fields(n, radius=radius, width=2.0)
    pt = dataclasses.replace(P, kappa=P.kappa * 1.1, P=P.P * 1.1, chi=P.chi * 1.1)
    tgt = ad.forward(c0, o.geometry(phi, pt), o.make_grid(n, 1.0), pt, dt, steps)[-1]
    return c0, phi, ad.Loss(tgt, lam=lam)


def test_oracle_adjoint_matches_finite_differences():
    n, steps, dt, eps = 32, 40, 0.01, 1e-6
    c0, phi, loss = _setup_oracle(n, 8.0, steps, dt, 0.1)
    L, g, dth = ad.gradient(c0, phi, P, dt, steps, loss)
    rng = np.random.default_rng(0)
    for _ in range(3):
        v = rng.standard_normal(c0.shape)
        fd = (ad.loss_of(c0 + eps * v, phi, P, dt, steps, loss)
              - ad.loss_of(c0 - eps * v, phi, P, dt, steps, loss)) / (2 * eps)
        assert abs(fd - np.sum(g * v)) <= 1e-6 * abs(fd)
    for i, name in enumerate(("kappa", "P", "chi")):
        pp = dataclasses.replace(P, **{name: getattr(P, name) + eps})
        pm = dataclasses.replace(P, **{name: getattr(P, name) - eps})
        fd = (ad.loss_of(c0, phi, pp, dt, steps, loss) - ad.loss_of(c0, phi, pm, dt, steps, loss)) / (2 * eps)
        assert abs(fd - dth[i]) <= 1e-6 * abs(fd), name


def test_oracle_adjoint_planted_sign_error_fails():
    """Flipping the sign of the stabilisation term in the VJP breaks FD agreement."""
    n, steps, dt, eps = 16, 10, 0.05, 1e-6
    c0, phi, loss = _setup_oracle(n, 4.0, steps, dt, 0.0)
    grid, shape = o.make_grid(n, 1.0), o.geometry(phi, P)
    cs = ad.forward(c0, shape, grid, P, dt, steps)
    gbar = cs[-1] - loss.target
    bad = dataclasses.replace(P, Q=-P.Q)
    for s in range(steps - 1, -1, -1):
        gbar, _ = ad.step_vjp(cs[s], gbar, shape, grid, bad, dt)
    v = np.random.default_rng(1).standard_normal(c0.shape)
    fd = (ad.loss_of(c0 + eps * v, phi, P, dt, steps, loss) - ad.loss_of(c0 - eps * v, phi, P, dt, steps, loss)) / (2 * eps)
    assert abs(fd - np.sum(gbar * v)) > 1e-4 * abs(fd)


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle_adjoint(pkg):
    say("adjoint: GPU vs oracle, 64^2, 200 steps")
    n, steps, dt = 64, 200, 0.01
    c0, phi, loss = _setup_oracle(n, 16.0, steps, dt, 0.1)
    L, g, dth = ad.gradient(c0, phi, P, dt, steps, loss)
    r = pkg.adjoint.gradient(c0, phi, loss.target, dt=dt, steps=steps, lam=0.1, ckpt_every=14)
    assert abs(r.loss / L - 1) < 1e-12
    assert np.max(np.abs(r.dc0 - g)) <= 1e-10 * np.max(np.abs(g))
    for got, ref in zip((r.dkappa, r.dP, r.dchi), dth):
        assert abs(got - ref) <= 1e-9 * abs(ref)


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle_adjoint_512(pkg):
    """The adjoint workload's grid and particle (512^2, R=160, w=2, lam =
    1e-3) against the CPU adjoint of the reference step, over 40 steps with
    7-step checkpoints (segments recomputed from their checkpoints)."""
    say("adjoint: GPU vs oracle, 512^2, 40 steps")
    n, steps, dt = 512, 40, 0.01
    c0, phi, loss = _setup_oracle(n, 160.0, steps, dt, 1e-3)
    L, g, dth = ad.gradient(c0, phi, P, dt, steps, loss)
    r = pkg.adjoint.gradient(c0, phi, loss.target, dt=dt, steps=steps, lam=1e-3, ckpt_every=7)
    assert abs(r.loss / L - 1) < 1e-12
    assert np.max(np.abs(r.dc0 - g)) <= 1e-10 * np.max(np.abs(g))
    for got, ref in zip((r.dkappa, r.dP, r.dchi), dth):
        assert abs(got - ref) <= 1e-9 * abs(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("ckpt", [1, 7, 50])
def test_gpu_adjoint_checkpoint_interval_invariant(pkg, ckpt):
    say(f"adjoint: checkpoint interval {ckpt}")
    n, steps, dt = 32, 50, 0.01
    c0, phi, loss = _setup_oracle(n, 8.0, steps, dt, 0.0)
    ref = pkg.adjoint.gradient(c0, phi, loss.target, dt=dt, steps=steps, ckpt_every=50)
    r = pkg.adjoint.gradient(c0, phi, loss.target, dt=dt, steps=steps, ckpt_every=ckpt)
    assert np.array_equal(r.dc0, ref.dc0)
    assert (r.dkappa, r.dP, r.dchi) == (ref.dkappa, ref.dP, ref.dchi)


def _fd_check(pkg, n, radius, steps, dt, lam, ckpt, ndir, seed=0):
    """Relative errors of the adjoint against finite differences of the GPU
    forward loss, per unit 2-norm direction and per parameter.

    The central difference at eps = 1e-5 is already roundoff-limited on long
    runs (its error grows ~10x at eps = 1e-6: the 1/eps signature of the
    loss's rounding noise, not the eps^2 of truncation), so the gating
    estimate is the fourth-order stencil at h = 1e-4: truncation O(h^4),
    roundoff ~1.5 x noise / h, ten times below the eps = 1e-5 difference.
    The central difference at eps = 1e-6 (the protocol of the 512^2 adjoint
    workload) is also computed: its vector of 32 directional derivatives is
    compared with the adjoint's as a whole.
    Returns (fourth-order direction errors, eps = 1e-5 central direction
    errors, fourth-order parameter errors, (eps = 1e-6 central directional
    derivatives, adjoint directional derivatives))."""
    c0, phi = pkg.initial_fields(n, radius=radius, width=2.0)
    c0, phi = np.asarray(c0), np.asarray(phi)
    pt = pkg.Parameters(kappa=P.kappa * 1.1, P=P.P * 1.1, chi=P.chi * 1.1)
    s = pkg.Session(n, 1.0, pt, "fp64")
    s.set_fields(c0, phi)
    s.step(dt, steps)
    tgt = s.field()
    s.close()
    r = pkg.adjoint.gradient(c0, phi, tgt, dt=dt, steps=steps, lam=lam, ckpt_every=ckpt)
    h = 1e-4

    def loss(c=c0, p=None):   # p: a package Parameters, or the defaults
        kw = {} if p is None else {"p": p}
        return pkg.adjoint.loss(c, phi, tgt, dt=dt, steps=steps, lam=lam, **kw)

    def d4(f):   # (f(-2h) - 8 f(-h) + 8 f(h) - f(2h)) / 12h
        return (f(-2 * h) - 8 * f(-h) + 8 * f(h) - f(2 * h)) / (12 * h)

    rng = np.random.default_rng(seed)
    e4, e2 = [], []
    fd6, ans = [], []
    for _ in range(ndir):
        # unit 2-norm directions keep the steps in the linear regime of the map
        v = rng.standard_normal(c0.shape)
        v /= np.linalg.norm(v)
        an = np.sum(r.dc0 * v)
        fd4 = d4(lambda t: loss(c=c0 + t * v))
        e4.append(abs(fd4 - an) / abs(fd4))
        fd2 = (loss(c=c0 + 1e-5 * v) - loss(c=c0 - 1e-5 * v)) / 2e-5
        e2.append(abs(fd2 - an) / abs(fd2))
        fd6.append((loss(c=c0 + 1e-6 * v) - loss(c=c0 - 1e-6 * v)) / 2e-6)
        ans.append(an)
    perr = []
    for name, got in (("kappa", r.dkappa), ("P", r.dP), ("chi", r.dchi)):
        fd = d4(lambda t: loss(p=pkg.Parameters(**{name: getattr(P, name) + t})))
        perr.append(abs(fd - got) / abs(fd))
    fd6, ans = np.array(fd6), np.array(ans)
    say(f"adjoint FD n={n} steps={steps}: directions 4th-order h=1e-4 max {max(e4):.2e} median "
        f"{np.median(e4):.2e}; central eps=1e-5 median {np.median(e2):.2e} max {max(e2):.2e}; "
        f"central eps=1e-6 vector {np.linalg.norm(fd6 - ans) / np.linalg.norm(ans):.2e} median "
        f"{np.median(np.abs(fd6 - ans) / np.abs(ans)):.2e}; params {max(perr):.2e}")
    return e4, e2, perr, (fd6, ans)


@pytest.mark.gpu
def test_gpu_adjoint_matches_finite_differences_128(pkg):
    e4, e2, perr, _ = _fd_check(pkg, 128, 40.0, 500, 0.01, 1e-3, 22, 8)
    assert max(e4) <= 1e-6 and max(e2) <= 1e-6 and max(perr) <= 1e-6


@pytest.mark.gpu
def test_gpu_adjoint_512_workload(pkg):
    """512^2, R=160, 2000 steps, sqrt(N)=45-step checkpoints, 32 directions.

    The workload's own protocol, central differences at eps = 1e-6 on 32
    random unit directions: the 32 directional derivatives agree with the
    adjoint's to <= 1e-6 relative as a vector (||FD - AD|| / ||AD||; measured
    ~1.7e-7) and per direction on the median.  Single directions do not: the
    loss of this 2000-step run carries rounding noise of ~1e-10 absolute (the
    two perturbed trajectories round differently; summing the loss exactly
    changes nothing), i.e. ~8e-5 on a difference quotient at eps = 1e-6,
    which is 2e-5 relative on the direction with the smallest derivative
    (|AD| = 4 against a median of 114).  The fourth-order estimate at h =
    1e-4 (truncation O(h^4), ten times less noise) bounds every direction
    and every parameter to 1e-6."""
    e4, e2, perr, (fd6, ans) = _fd_check(pkg, 512, 160.0, 2000, 0.01, 1e-3, 45, 32)
    assert np.linalg.norm(fd6 - ans) / np.linalg.norm(ans) <= 1e-6
    assert np.median(np.abs(fd6 - ans) / np.abs(ans)) <= 1e-6
    assert max(e4) <= 1e-6
    assert np.median(e2) <= 1e-6
    assert max(perr) <= 1e-6


@pytest.mark.gpu
def test_mixed_context_rejects_adjoint(pkg):
    from csg_amd import _capi
    s = pkg.Session(32, 1.0, pkg.Parameters(), "mixed")
    c0, phi = pkg.initial_fields(32, radius=8.0)
    s.set_fields(c0, phi)
    assert pkg.lib().csg_adjoint_forward(s.handle, 0.01, 10, 3) == _capi.CSG_ERR_UNSUPPORTED
    s.close()


@pytest.mark.gpu
def test_gpu_kept_states_survive_schedule_change(pkg):
    """With every state kept (ckpt_every = 1), a schedule change between the
    forward and the backward sweep drops the captured graphs but not the
    states: the backward still runs and gives the gradient of the same run
    (its own P3 launches now use the other schedule, so to rounding)."""
    import ctypes
    from csg_amd import _capi
    n, steps, dt = 128, 60, 0.01
    c0, phi = pkg.initial_fields(n, radius=38.4, width=2.0)
    c0, phi = np.asarray(c0), np.asarray(phi)
    rng = np.random.default_rng(3)
    gbar = np.asfortranarray(rng.standard_normal((n, n)))
    out = []
    for change in (False, True):
        s = pkg.Session(n)
        s.set_fields(c0, phi)
        assert pkg.lib().csg_adjoint_forward(s.handle, dt, steps, 1) == 0
        if change:
            s.set_schedule((0, 0, 1, 0))
        dc0 = np.empty((n, n), order="F")
        dp = (ctypes.c_double * 3)()
        _capi.check(pkg.lib().csg_adjoint_backward(s.handle, _capi.ptr(gbar), _capi.ptr(dc0), dp))
        out.append((dc0, np.array(dp[:])))
        s.close()
    assert np.max(np.abs(out[1][0] - out[0][0])) <= 1e-10 * np.max(np.abs(out[0][0]))
    assert np.all(np.abs(out[1][1] - out[0][1]) <= 1e-10 * np.abs(out[0][1]))



@pytest.mark.gpu
def test_keep_every_state_falls_back_to_checkpoints(pkg, monkeypatch):
    """When keeping every state does not fit (forced here with the library's
    test-only device byte budget), the failed forward run frees the
    checkpoints it allocated and gradient() retries with ~sqrt(steps)
    checkpoints within the budget; the gradient equals the one computed with
    that interval directly."""
    n, steps, dt = 128, 400, 0.01
    c0, phi = pkg.initial_fields(n, radius=38.4, width=2.0)
    tgt = np.asfortranarray(np.asarray(c0) + 0.01 * np.sin(np.arange(n * n).reshape(n, n)))
    say("adjoint: keep-every-state fallback under a 24 MiB budget")
    ref = pkg.adjoint.gradient(c0, phi, tgt, dt=dt, steps=steps, ckpt_every=20)
    for k in list(pkg.adjoint._contexts):
        pkg.adjoint._contexts.pop(k).close()
    budget = 24 << 20
    assert (steps + 1) * n * n * 8 > budget            # every state cannot fit
    assert pkg.adjoint.default_ckpt_every(n, steps) == 1
    monkeypatch.setenv("CSG_TEST_DEVICE_BUDGET", str(budget))
    try:
        g = pkg.adjoint.gradient(c0, phi, tgt, dt=dt, steps=steps)
        ctx = pkg.adjoint._contexts[(n, 1.0, 0)]
        assert ctx.device_bytes() <= budget
    finally:
        for k in list(pkg.adjoint._contexts):
            pkg.adjoint._contexts.pop(k).close()
    assert np.array_equal(g.dc0, ref.dc0)
    assert (g.dkappa, g.dP, g.dchi) == (ref.dkappa, ref.dP, ref.dchi)
    assert g.loss == ref.loss
