"""Host-side tests (no GPU): the C ABI library loads and exports every symbol
include/csg.h declares; the product's host RNG / initial_fields match the
oracle bit for bit; argument validation; the Julia CSV number format."""
import ctypes
import os
import re
import sys

import numpy as np
import pytest

from oracle import core_shell_oracle as o
from oracle import julia_rng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "csg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    decl = _declared_symbols()
    assert len(decl) >= 19
    for name in decl:
        assert hasattr(L, name), name
    from csg_amd import _capi
    assert set(decl) == set(_capi.EXPORTS)


def test_default_params_match_reference(pkg):
    from csg_amd import _capi
    p = _capi.CsgParams()
    pkg.lib().csg_default_params(ctypes.byref(p))
    ref = o.Parameters()
    for f in ("A", "B", "chi", "P", "ca", "cb", "cc", "kappa", "mobility", "Q"):
        assert getattr(p, f) == getattr(ref, f) == getattr(pkg.Parameters(), f)


@pytest.mark.parametrize("seed,count", [(5749, 4096), (5749, 1000), (5749, 7), (0, 385), (1, 384),
                                        (123456789012, 2049), (-3, 500)])
def test_julia_rand_matches_oracle(pkg, seed, count):
    out = np.empty(count)
    rc = pkg.lib().csg_julia_rand(seed, count, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    ref = julia_rng.MersenneTwister(seed).rand(count)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("n,radius,width,spacing", [(64, 16.0, 4.0, 1.0), (256, 70.0, 4.0, 1.0),
                                                    (17, 5.0, 2.0, 0.5), (128, 38.4, 2.0, 1.0)])
def test_initial_fields_match_oracle(pkg, n, radius, width, spacing):
    c, phi = pkg.initial_fields(n, spacing=spacing, radius=radius, width=width)
    cr, pr = o.initial_fields(n, spacing=spacing, radius=radius, width=width)
    # libm tanh vs numpy tanh: last-ulp differences of values near 1, which
    # the cancellation in 1 - tanh turns into absolute differences <= 1 ulp(1)
    assert np.max(np.abs(phi - pr)) <= 2.3e-16
    assert np.array_equal(phi >= 0.5, pr >= 0.5)
    assert np.max(np.abs(c - cr)) <= 1e-17   # mean subtraction may differ in the last ulp


def test_initial_fields_known_answer_energy_via_oracle(pkg):
    """The product's initial field gives check-results.txt's F(t=0) under the oracle."""
    c, phi = pkg.initial_fields(64, radius=16.0)
    F0 = o.energy(np.asarray(c), o.geometry(np.asarray(phi), o.Parameters()), o.make_grid(64, 1.0),
                  o.Parameters())
    assert abs(F0 / 329.48858453024934 - 1) < 2e-15


@pytest.mark.parametrize("n", [8, 24, 100, 1536, 8192])
def test_unsupported_grid_rejected_without_gpu(pkg, n):
    from csg_amd import _capi
    h = _capi._VP()
    rc = pkg.lib().csg_create(h, n, 1.0, None, 0, 0)
    assert rc == _capi.CSG_ERR_UNSUPPORTED
    assert "not supported" in pkg.lib().csg_last_error().decode()


def test_bad_arguments_rejected_without_gpu(pkg):
    from csg_amd import _capi
    L = pkg.lib()
    h = _capi._VP()
    assert L.csg_create(h, 64, 0.0, None, 0, 0) == _capi.CSG_ERR_ARG
    assert L.csg_create(h, 64, 1.0, None, 7, 0) == _capi.CSG_ERR_ARG
    assert L.csg_create(None, 64, 1.0, None, 0, 0) == _capi.CSG_ERR_ARG
    assert L.csg_run(None, 0.01, 10, 1, None) == _capi.CSG_ERR_ARG
    assert L.csg_nonfinite_step(None) == -1
    assert L.csg_initial_fields(0, 1.0, 1.0, 1.0, 1, None, None) == _capi.CSG_ERR_ARG


@pytest.mark.parametrize("x,s", [(0.5, "0.5"), (1.0, "1.0"), (1e-5, "1.0e-5"), (0.0001, "0.0001"),
                                 (4406.5104552251505, "4406.5104552251505"), (-0.0, "-0.0"),
                                 (0.0, "0.0"), (100000.0, "100000.0"), (1e6, "1.0e6"),
                                 (1234567.0, "1.234567e6"), (123456.7, "123456.7"),
                                 (0.007360143442235456, "0.007360143442235456"),
                                 (-1.25e-7, "-1.25e-7"), (6.661338147750939e-16, "6.661338147750939e-16"),
                                 (float("nan"), "NaN"), (float("inf"), "Inf"), (2.0 ** 70, "1.1805916207174113e21")])
def test_julia_float_format(pkg, x, s):
    assert pkg.julia_float(x) == s


def test_writedlm_layout(pkg, tmp_path):
    a = np.array([[0.5, 1.0], [1e-5, -2.5]], order="F")
    p = tmp_path / "m.csv"
    pkg.writedlm(str(p), a)
    assert p.read_text() == "0.5,1.0\n1.0e-5,-2.5\n"   # row i = first index


def test_fft_engine_host_check(tmp_path):
    """The register Stockham engine's stage code (plans, radix-16/8/4/2
    butterflies, swizzled or padded exchanges, twiddles from the half table
    and from the per-stage tables) against a long-double DFT for every line
    length 8..4096, FP64 and FP32; the twiddle-buffer layout; the LDS slot
    map (injective into the buffer) and tile row strides; and the line-layout
    map with its row-tile element order (a bijection, contiguous per block,
    for line blocks of 1, 2 and 4) -- compiled for the host."""
    import subprocess
    exe = tmp_path / "fft_engine_check"
    src = os.path.join(ROOT, "tests", "cpp", "fft_engine_check.hip")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-std=c++17", src, "-o", str(exe)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout


def test_make_grid_matches_oracle(pkg):
    for n, sp in ((16, 1.0), (64, 0.5), (15, 2.0)):
        g = pkg.make_grid(n, sp)
        r = o.make_grid(n, sp)
        assert np.array_equal(g.kx, r["kx"]) and np.array_equal(g.ky, r["ky"]) and np.array_equal(g.k2, r["k2"])


def test_adjoint_default_checkpointing(pkg):
    """Every state is kept when the run's states fit the memory budget (the
    512^2 x 2000-step workload: 4.2 GB), else ~sqrt(steps) checkpoints."""
    assert pkg.adjoint.default_ckpt_every(512, 2000) == 1
    assert pkg.adjoint.default_ckpt_every(4096, 2000) == 45      # 2001 x 128 MiB > 32 GiB
    budget = pkg.adjoint.KEEP_ALL_BYTES
    n = 1024
    fit = budget // (n * n * 8) - 1          # largest step count whose states fit
    assert pkg.adjoint.default_ckpt_every(n, fit) == 1
    assert pkg.adjoint.default_ckpt_every(n, fit + 1) == max(1, round((fit + 1) ** 0.5))


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with status 2 and says so,
    before any GPU work."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2 and r.stdout == ""
    assert "one GPU only" in r.stderr


def test_function_api_context_cache_is_bounded(pkg, monkeypatch):
    """The function API reuses one context per (n, spacing, precision,
    device), swapping Parameters in place, and keeps at most MAX_SESSIONS
    (host logic; a stand-in Session records what the library would do)."""
    from csg_amd import core_shell as cs
    made, closed = [], []

    class FakeSession:
        def __init__(self, n, spacing, p, precision, device):
            self.key, self.p, self.swaps = (n, spacing, precision, device), p, 0
            made.append(self)

        def set_params(self, p):
            self.p, self.swaps = p, self.swaps + 1

        def close(self):
            closed.append(self)

    monkeypatch.setattr(cs, "Session", FakeSession)
    monkeypatch.setattr(cs, "_sessions", type(cs._sessions)())
    for i in range(50):
        p = pkg.Parameters(P=3.5 + 0.06 * i, chi=0.35 + 0.006 * i)
        s = cs._session(64, 1.0, p, "fp64")
        assert s.p == p
    assert len(made) == 1 and made[0].swaps == 49 and len(cs._sessions) == 1
    for n in (16, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024):
        cs._session(n, 1.0, pkg.Parameters(), "mixed")
    assert len(cs._sessions) == cs.MAX_SESSIONS
    assert len(made) - len(closed) == cs.MAX_SESSIONS
    assert all(s.key[0] in (16, 32, 48, 64, 96) for s in closed)
