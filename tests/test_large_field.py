"""The large single field (4096^2, R = 1300, w = 2, Janus-type parameters B = 8,
chi = 0.8, Q = 1): one FP64 step against the oracle, then size-independent properties
of a short run (mass conservation, energy decrease, finite bounds) and the
mixed form against the FP64 form."""
import sys

import numpy as np
import pytest

from oracle import core_shell_oracle as o

pytestmark = pytest.mark.gpu
N = 4096


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def janus(pkg):
    d = pkg.Parameters()
    p = pkg.Parameters(B=2.0 * d.B, chi=d.chi + 0.3, Q=2.0 * d.Q)
    po = o.Parameters(B=2.0 * d.B, chi=d.chi + 0.3, Q=2.0 * d.Q)
    c, phi = pkg.initial_fields(N, radius=1300.0, width=2.0)
    return p, po, np.asarray(c), np.asarray(phi)


def test_large_field_one_step_matches_oracle(pkg, janus):
    p, po, c, phi = janus
    say("4096^2 one FP64 step vs oracle")
    s = pkg.Session(N, 1.0, p, "fp64")
    s.set_fields(c, phi)
    h = s.run(0.01, 1)
    c1 = s.field()
    s.close()
    cr, hr = o.run_simulation_fast(c, phi, steps=1, p=po)
    assert np.max(np.abs(c1 - cr)) < 1e-13
    assert np.max(np.abs(h - hr) / np.maximum(1.0, np.abs(hr))) < 1e-12


def test_large_field_invariants_and_mixed_error(pkg, janus):
    p, _, c, phi = janus
    steps = 200
    out = {}
    for form in ("fp64", "mixed"):
        say(f"4096^2 {form}, {steps} steps")
        s = pkg.Session(N, 1.0, p, form)
        s.set_fields(c, phi)
        h = s.run(0.01, steps)
        out[form] = (s.field(), h)
        s.close()
    cf, hf = out["fp64"]
    assert np.all(np.isfinite(hf))
    assert np.max(np.abs(hf[:, 1] - c.mean())) < 1e-13            # mass (checks.jl:19)
    assert np.max(np.diff(hf[:, 2])) <= 1e-9 * abs(hf[0, 2])       # energy non-increasing (checks.jl:20)
    assert abs(hf[-1, 3] - cf.min()) < 1e-14 and abs(hf[-1, 4] - cf.max()) < 1e-14   # bounds of c_steps
    cm, hm = out["mixed"]
    assert np.max(np.abs(hm[:, 1] - c.mean())) < 1e-11            # FP64 k = 0 mode and mean in the mixed form
    # FP32 bulk arithmetic: 1.3e-5 measured after 200 steps at the interface,
    # where the spinodal dynamics amplify rounding differences
    assert np.max(np.abs(cm - cf)) < 5e-5
    assert np.max(np.abs(hm[:, 2] - hf[:, 2]) / np.abs(hf[:, 2])) < 1e-6


def test_large_field_phi_hint_is_bit_identical(pkg, janus, monkeypatch):
    """Non-temporal phi loads (on by default for this field) change where phi
    is cached, never values: 5 steps with and without the hint agree bit for bit."""
    p, _, c, phi = janus
    out = []
    for v in ("2", "6"):
        say(f"4096^2 FP64, CSG_NONTEMPORAL={v}")
        monkeypatch.setenv("CSG_NONTEMPORAL", v)
        s = pkg.Session(N, 1.0, p, "fp64")
        s.set_fields(c, phi)
        h = s.run(0.01, 5)
        out.append((s.field(), h))
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
