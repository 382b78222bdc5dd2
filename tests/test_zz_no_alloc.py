"""Allocation-free stepping (csg_step_noalloc) through the C ABI.

Its result must equal the context path (csg_create + csg_set_fields +
csg_step + csg_get_field) on the same inputs bit for bit: the same kernels,
default schedule and twiddles run, only the work arrays live in the library's
static scratch.  Two calls in a row, the second on a smaller grid, in both
precisions; the largest grid the scratch holds; sizes beyond it rejected.
Device buffers and the stream come from the HIP runtime the library itself
links (libamdhip64 through ctypes); a second HIP runtime in the process, such
as the one torch bundles, would not share its device state.
"""
import ctypes
import dataclasses
import sys

import numpy as np
import pytest


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _context(pkg, n, p, prec, c0, phi, dt, nsteps):
    s = pkg.Session(n, 1.0, p, prec)
    s.set_fields(c0, phi)
    s.step(dt, nsteps)
    c = s.field()
    s.close()
    return c


class _Hip:
    """The few HIP runtime calls the test needs, from the runtime libcsg.so
    links."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")

    def ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    def to_device(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)), "hipMalloc")
        self.ok(self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")
        return p

    def to_host(self, p, n):
        out = np.empty((n, n), order="F")
        self.ok(self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def free(self, p):
        self.ok(self.h.hipFree(p), "hipFree")


def _noalloc(pkg, n, p, prec, c0, phi, dt, nsteps, own_stream):
    from csg_amd import _capi
    hip = _Hip()
    cd, pd = hip.to_device(c0), hip.to_device(phi)
    stream = ctypes.c_void_p()
    if own_stream:
        hip.ok(hip.h.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    pc = p._c()
    rc = pkg.lib().csg_step_noalloc(n, 1.0, ctypes.byref(pc), _capi.CSG_FP64 if prec == "fp64" else _capi.CSG_MIXED,
                                    0, dt, nsteps, cd, pd, stream if own_stream else None)
    _capi.check(rc)
    out = hip.to_host(cd, n)
    if own_stream:
        hip.ok(hip.h.hipStreamDestroy(stream), "hipStreamDestroy")
    hip.free(cd)
    hip.free(pd)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_equals_context_two_calls(pkg, prec):
    d = pkg.Parameters()
    cases = [(512, d, 20, True), (256, dataclasses.replace(d, P=6.1, chi=0.41), 15, False)]
    for n, p, nsteps, ts in cases:
        say(f"noalloc {prec} n={n} steps={nsteps}")
        c0, phi = pkg.initial_fields(n, radius=0.3 * n, width=2.0, seed=5749 + n)
        ref = _context(pkg, n, p, prec, c0, phi, 0.01, nsteps)
        got = _noalloc(pkg, n, p, prec, c0, phi, 0.01, nsteps, ts)
        assert np.array_equal(got, ref), (prec, n, float(np.max(np.abs(got - ref))))


@pytest.mark.gpu
def test_noalloc_largest_grid_and_zero_steps(pkg):
    from csg_amd import _capi
    d = pkg.Parameters()
    n = _capi.CSG_NOALLOC_MAX_N
    say(f"noalloc fp64 n={n} steps=5, then 0 steps")
    c0, phi = pkg.initial_fields(n, radius=0.3 * n, width=2.0)
    assert np.array_equal(_noalloc(pkg, n, d, "fp64", c0, phi, 0.01, 5, True),
                          _context(pkg, n, d, "fp64", c0, phi, 0.01, 5))
    assert np.array_equal(_noalloc(pkg, 64, d, "mixed", c0[:64, :64], phi[:64, :64], 0.01, 0, False),
                          _context(pkg, 64, d, "mixed", c0[:64, :64], phi[:64, :64], 0.01, 0))


def test_noalloc_rejects_sizes_beyond_scratch(pkg):
    """Argument checks run before any device call (CPU)."""
    from csg_amd import _capi
    L = pkg.lib()
    dummy = ctypes.c_void_p(16)
    n = 2 * _capi.CSG_NOALLOC_MAX_N
    assert L.csg_step_noalloc(n, 1.0, None, 0, 0, 0.01, 1, dummy, dummy, None) == _capi.CSG_ERR_ARG
    assert "allocation-free maximum" in L.csg_last_error().decode()
    assert L.csg_step_noalloc(100, 1.0, None, 0, 0, 0.01, 1, dummy, dummy, None) == _capi.CSG_ERR_UNSUPPORTED
    assert L.csg_step_noalloc(64, 1.0, None, 0, 0, 0.01, 1, None, dummy, None) == _capi.CSG_ERR_ARG
    assert L.csg_step_noalloc(64, -1.0, None, 0, 0, 0.01, 1, dummy, dummy, None) == _capi.CSG_ERR_ARG
    assert L.csg_step_noalloc(64, 1.0, None, 2, 0, 0.01, 1, dummy, dummy, None) == _capi.CSG_ERR_ARG
    assert L.csg_step_noalloc(64, 1.0, None, 0, 0, 0.01, -1, dummy, dummy, None) == _capi.CSG_ERR_ARG
