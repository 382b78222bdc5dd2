"""Ensembles: many independent particles, each a ``run_simulation``
(``core_shell.jl:84-103``) with its own grid size, parameters and seed,
advanced together on one GPU through the ``csg_batch_*`` entry points.

The reference runs one particle per call; an ensemble member gives exactly
the result of ``run_simulation`` with the same inputs and precision (the same
kernels run, one instance per row of workgroups).
"""
from __future__ import annotations

import ctypes
import dataclasses
import weakref
from typing import Optional, Sequence

import numpy as np

from . import _capi
from ._capi import check, lib, ptr
from .core_shell import _PREC, NonFiniteCompositionError, Parameters, _fortran, initial_fields


class EnsembleNonFiniteError(NonFiniteCompositionError):
    """Some members turned non-finite; ``steps[i]`` is member i's first
    non-finite step (-1: finite throughout).  The other members ran to the end."""

    def __init__(self, steps: Sequence[int]):
        bad = [i for i, s in enumerate(steps) if s >= 0]
        super().__init__(int(steps[bad[0]]))
        self.steps = list(steps)
        self.members = bad


@dataclasses.dataclass(frozen=True)
class Member:
    """Inputs of one ensemble member (``initial_fields`` arguments + parameters)."""
    n: int
    radius: float
    width: float = 2.0
    seed: int = 5749
    spacing: float = 1.0
    p: Parameters = Parameters()


def ensemble_members(count: int = 256, sizes: Sequence[int] = (128, 256, 384, 512), *,
                     radius_fraction: float = 0.3, width: float = 2.0, spread: float = 0.3,
                     base_seed: int = 5749, draw_seed: int = 2024) -> list:
    """The uneven ensemble workload: grid sizes drawn uniformly from `sizes`,
    R = radius_fraction * N, P and chi drawn uniformly within +-spread of the
    defaults of core_shell.jl:5-14, noise seed = base_seed + index."""
    rng = np.random.default_rng(draw_seed)
    d = Parameters()
    out = []
    for i in range(count):
        n = int(rng.choice(np.asarray(sizes)))
        P = float(rng.uniform(d.P * (1 - spread), d.P * (1 + spread)))
        chi = float(rng.uniform(d.chi * (1 - spread), d.chi * (1 + spread)))
        out.append(Member(n=n, radius=radius_fraction * n, width=width, seed=base_seed + i,
                          p=dataclasses.replace(d, P=P, chi=chi)))
    return out


def member_fields(m: Member):
    """``initial_fields`` of a member (host, bit-exact Julia noise)."""
    return initial_fields(m.n, spacing=m.spacing, radius=m.radius, width=m.width, seed=m.seed)


class Ensemble:
    """Device state of an ensemble: one ``csg_batch`` handle."""

    def __init__(self, ns: Sequence[int], spacings: Optional[Sequence[float]] = None,
                 params: Optional[Sequence[Parameters]] = None, precision: str = "mixed", device: int = 0,
                 persistent: bool = False):
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        self.ns = [int(n) for n in ns]
        k = len(self.ns)
        if k < 1:
            raise ValueError("an ensemble needs at least one member")
        n_arr = (ctypes.c_int32 * k)(*self.ns)
        sp = None
        if spacings is not None:
            if len(spacings) != k:
                raise ValueError("one spacing per member")
            self._sp = np.asarray(spacings, dtype=np.float64)
            sp = ptr(self._sp)
        pa = None
        if params is not None:
            if len(params) != k:
                raise ValueError("one Parameters per member")
            pa = (_capi.CsgParams * k)(*(p._c() for p in params))
        h = _capi._VP()
        check(lib().csg_batch_create(h, k, n_arr, sp, pa, _PREC[precision], int(device)))
        self.handle = h
        self.precision = precision
        self.persistent = False
        self._fin = weakref.finalize(self, lib().csg_batch_destroy, h)
        if persistent:
            self.set_persistent(True)

    def set_persistent(self, enable: bool = True):
        """Run every pass of every member in one persistent launch per block
        of steps (sizes 128, 256, 384, 512) instead of batched launches."""
        check(lib().csg_batch_set_persistent(self.handle, 1 if enable else 0))
        self.persistent = bool(enable)

    def __len__(self):
        return len(self.ns)

    def close(self):
        self._fin()

    def set_fields(self, i: int, c0, phi):
        c0, phi = _fortran(c0), _fortran(phi)
        n = self.ns[i]
        if c0.shape != (n, n) or phi.shape != (n, n):
            raise ValueError(f"member {i}: expected {n}x{n} arrays")
        check(lib().csg_batch_set_fields(self.handle, int(i), ptr(c0), ptr(phi)))

    def run(self, dt: float, steps: int, history_every: int = 1, histories: bool = True):
        """csg_batch_run; returns the members' (rows x 6) histories (None
        entries when `histories` is False).  Raises EnsembleNonFiniteError
        after the run if any member turned non-finite."""
        k = len(self.ns)
        rows = steps // history_every + 1
        hs = [np.zeros((rows, 6), order="F") for _ in range(k)] if histories else [None] * k
        arr = (_capi._D * k)(*(ptr(h) if h is not None else None for h in hs)) if histories else None
        rc = lib().csg_batch_run(self.handle, float(dt), int(steps), int(history_every), arr)
        if rc == _capi.CSG_ERR_NONFINITE:
            raise EnsembleNonFiniteError([int(lib().csg_batch_nonfinite_step(self.handle, i)) for i in range(k)])
        check(rc)
        return hs

    def field(self, i: int) -> np.ndarray:
        n = self.ns[i]
        c = np.empty((n, n), order="F")
        check(lib().csg_batch_get_field(self.handle, int(i), ptr(c)))
        return c

    def device_bytes(self) -> int:
        return int(lib().csg_batch_device_bytes(self.handle))

    def schedule(self, n: int) -> tuple:
        """Selected kernel candidates of the members of size n (see Session)."""
        sel = (ctypes.c_int32 * 4)()
        check(lib().csg_batch_get_schedule(self.handle, int(n), sel))
        return tuple(sel)

    def set_schedule(self, n: int, sel):
        if len(sel) != 4:
            raise ValueError("one candidate per pass")
        check(lib().csg_batch_set_schedule(self.handle, int(n), (ctypes.c_int32 * 4)(*(int(x) for x in sel))))

    def autotune(self, dt: float = 0.01, rounds: int = 1) -> dict:
        """Tune each size class alone (resident fields kept): {n: selection}."""
        check(lib().csg_batch_autotune(self.handle, float(dt), int(rounds)))
        return {n: self.schedule(n) for n in sorted(set(self.ns))}


def run_ensemble(c0s: Sequence, phis: Sequence, dt: float = 0.01, steps: int = 1000,
                 spacings: Optional[Sequence[float]] = None, params: Optional[Sequence[Parameters]] = None, *,
                 precision: str = "mixed", history_every: int = 1, device: int = 0, persistent: bool = False):
    """``run_simulation`` of every member at once: returns [(c, history), ...]."""
    if len(c0s) != len(phis):
        raise ValueError("one phi per c0")
    ens = Ensemble([np.shape(c)[0] for c in c0s], spacings, params, precision, device, persistent)
    try:
        for i, (c0, phi) in enumerate(zip(c0s, phis)):
            ens.set_fields(i, c0, phi)
        hs = ens.run(dt, steps, history_every)
        return [(ens.field(i), hs[i]) for i in range(len(ens))]
    finally:
        ens.close()
