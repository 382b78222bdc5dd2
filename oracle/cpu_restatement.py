"""ctypes binding of ``oracle/cs_cpu.c``: the reference time loop
(``core_shell.jl:84-103``) restated in C with OpenMP, for the CPU baseline
timed on all host cores beside the GPU.

Test infrastructure only, like the rest of ``oracle/``: ``tests/`` and
``bench.py``'s CPU-baseline leg use it; the HIP library never does.  It does
the same work per loop iteration as the reference (17 complex N x N 2D
transforms, radix-2, so N must be a power of two) and is held to the numpy
oracle and to the reference's known answers by ``tests/test_cpu_restatement.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .core_shell_oracle import Parameters

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cs_cpu.c")
LIB = os.path.join(HERE, "libcs_cpu.so")

CS_ERR_NONFINITE = 3

_lib = None


def build() -> str:
    """Compile ``cs_cpu.c`` into ``libcs_cpu.so`` beside it (gcc, OpenMP)."""
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O3", "-std=c11", "-fcx-limited-range", "-fopenmp", "-fPIC", "-shared", "-Wall", SRC, "-o", tmp, "-lm"],
                   check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        D = ctypes.POINTER(ctypes.c_double)
        L.cs_run.restype = ctypes.c_int
        L.cs_run.argtypes = [ctypes.c_int, ctypes.c_double, D, ctypes.c_double, ctypes.c_long, D, D, D, D,
                             ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
        _lib = L
    return _lib


class NonFiniteError(RuntimeError):
    pass


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def run_simulation(c0, phi, dt=0.01, steps=1000, spacing=1.0, p: Parameters = Parameters(), threads=1):
    """``run_simulation`` (core_shell.jl:84-103) on `threads` OpenMP threads:
    (c, history) with history rows [t, mean c, F, min c, max c, dissipation]."""
    c0 = np.asfortranarray(c0, dtype=np.float64)
    phi = np.asfortranarray(phi, dtype=np.float64)
    n = c0.shape[0]
    if c0.shape != (n, n) or phi.shape != (n, n):
        raise ValueError("c0 and phi must be n x n")
    prm = np.array([p.A, p.B, p.chi, p.P, p.ca, p.cb, p.cc, p.kappa, p.mobility, p.Q], dtype=np.float64)
    c = np.empty((n, n), order="F")
    hist = np.zeros((steps + 1, 6))
    bad = ctypes.c_long(-1)
    rc = lib().cs_run(n, float(spacing), _p(prm), float(dt), int(steps), _p(c0), _p(phi), _p(c), _p(hist),
                      int(threads), ctypes.byref(bad))
    if rc == CS_ERR_NONFINITE:
        raise NonFiniteError(f"Non-finite composition at step {bad.value}")
    if rc != 0:
        raise RuntimeError(f"cs_run failed with code {rc}")
    return c, hist
