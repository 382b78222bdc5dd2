"""Host API of the HIP core-shell step library, shaped like ``core_shell.jl``.

Each function keeps the reference's name, argument meaning and error
behaviour (``core_shell.jl:4-103``); the arithmetic runs in ``libcsg.so`` on
the GPU (there is no CPU path).  Arrays are indexed like the Julia matrices,
``a[i, j]`` with ``i`` the first index (along which ``k_x`` varies), and are
handed to the C ABI in column-major order.

Extensions beyond the reference signature, all keyword-only with defaults that
reproduce the reference: ``precision`` ("fp64" or "mixed"), ``history_every``
(1 = every step, as ``core_shell.jl:89-92``) and ``device``.
"""
from __future__ import annotations

import collections
import ctypes
import dataclasses
import os
import weakref

import numpy as np

from . import _capi
from ._capi import check, lib, ptr


@dataclasses.dataclass(frozen=True)
class Parameters:
    """``Parameters`` (core_shell.jl:4-15)."""
    A: float = 2.0
    B: float = 4.0
    chi: float = 0.5
    P: float = 5.0
    ca: float = 0.5
    cb: float = 0.0
    cc: float = 1.0
    kappa: float = 2.0
    mobility: float = 1.0
    Q: float = 0.5

    def _c(self) -> _capi.CsgParams:
        return _capi.CsgParams(*(float(getattr(self, f.name)) for f in dataclasses.fields(self)))


class NonFiniteCompositionError(RuntimeError):
    """Raised as ``error("Non-finite composition at step $s")`` (core_shell.jl:93)."""

    def __init__(self, step: int):
        super().__init__(f"Non-finite composition at step {step}")
        self.step = step


@dataclasses.dataclass(frozen=True)
class Grid:
    """``make_grid`` result (core_shell.jl:17-25): ``kx`` (n x 1) and ``ky``
    (1 x n) with the Nyquist mode zeroed, ``k2`` (n x n) with it kept.  The
    device computes wavenumbers on the fly; these arrays are for callers."""
    kx: np.ndarray
    ky: np.ndarray
    k2: np.ndarray
    spacing: float

    @property
    def n(self) -> int:
        return self.k2.shape[0]


def _fftfreq(n: int, fs: float) -> np.ndarray:
    m = np.arange(n, dtype=np.float64)
    m[m >= (n + 1) // 2] -= n
    return m * (fs / n)


def make_grid(n: int, spacing: float) -> Grid:
    k = (2 * np.pi) * _fftfreq(n, 1 / spacing)
    dk = k.copy()
    if n % 2 == 0:
        dk[n // 2] = 0.0
    return Grid(dk.reshape(n, 1), dk.reshape(1, n), (k ** 2).reshape(n, 1) + (k ** 2).reshape(1, n),
                float(spacing))


def _fortran(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def initial_fields(n: int, spacing: float = 1.0, radius: float = 70.0, width: float = 4.0,
                   seed: int = 5749):
    """``initial_fields`` (core_shell.jl:27-37), noise bit-identical to Julia's
    ``rand(MersenneTwister(seed), n, n)``."""
    c = np.empty((n, n), order="F")
    phi = np.empty((n, n), order="F")
    rc = lib().csg_initial_fields(int(n), float(spacing), float(radius), float(width), int(seed),
                                  ptr(c), ptr(phi))
    check(rc)
    return c, phi


_PREC = {"fp64": _capi.CSG_FP64, "mixed": _capi.CSG_MIXED}


PASSES = ("P1_row", "P2_col", "P3_row", "P4_col")


def _sel4(sel) -> "ctypes.Array":
    if len(sel) != 4:
        raise ValueError("a schedule selects one candidate for each of the 4 passes")
    return (ctypes.c_int32 * 4)(*(int(x) for x in sel))


class Session:
    """One device context (grid size, spacing, parameters, precision).

    ``schedule``: None (the default candidates), "cached" (the tuned choice
    stored by ``schedules.save`` for this size and precision, if any) or four
    candidate indices (see ``schedule_info``)."""

    def __init__(self, n: int, spacing: float = 1.0, p: Parameters = Parameters(),
                 precision: str = "fp64", device: int = 0, schedule=None):
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        self.n, self.spacing, self.p, self.precision = int(n), float(spacing), p, precision
        h = _capi._VP()
        pc = p._c()
        check(lib().csg_create(h, self.n, self.spacing, pc, _PREC[precision], int(device)))
        self.handle = h
        self._phi_key = None
        self._fin = weakref.finalize(self, lib().csg_destroy, h)
        if schedule == "cached":
            from . import schedules
            schedule = schedules.load(self.n, precision, self.schedule_info())
        if schedule is not None:
            self.set_schedule(schedule)

    def schedule_info(self) -> list:
        """Per pass, the compiled candidates as dicts (E = elements per thread,
        W = rows or lines per workgroup, G = tile-load group, PF = when the
        second input is loaded (0 where used, 1 with the first loads, 2 once
        they have arrived),
        X = FFT exchange layout, T = 1 for per-stage twiddle tables and 2
        for the half table staged in LDS beyond 16 KiB, S = 1
        when P3 gives each gradient component its own workgroups, 2 when it
        does so with the two components of a row tile on neighbouring
        workgroups, 0 when one workgroup transforms both; 0 for the other
        passes)."""
        keys = ("E", "W", "G", "PF", "X", "T", "S")
        out = []
        for q in range(4):
            cnt = ctypes.c_int32()
            check(lib().csg_schedule_info(self.handle, q, ctypes.byref(cnt), None))
            buf = (ctypes.c_int32 * (len(keys) * cnt.value))()
            check(lib().csg_schedule_info(self.handle, q, ctypes.byref(cnt), buf))
            out.append([dict(zip(keys, buf[len(keys) * i:len(keys) * (i + 1)])) for i in range(cnt.value)])
        return out

    @property
    def schedule(self) -> tuple:
        sel = (ctypes.c_int32 * 4)()
        check(lib().csg_get_schedule(self.handle, sel))
        return tuple(sel)

    def set_schedule(self, sel):
        check(lib().csg_set_schedule(self.handle, _sel4(sel)))

    def set_params(self, p: Parameters):
        """Replace the parameters in place (no reallocation; field kept)."""
        check(lib().csg_set_params(self.handle, p._c()))
        self.p = p

    def autotune(self, dt: float = 0.01, rounds: int = 2):
        """Time the candidates inside the captured loop (needs set_fields; the
        resident field is kept): returns (selection, device us per step)."""
        sel = (ctypes.c_int32 * 4)()
        us = ctypes.c_double()
        check(lib().csg_autotune(self.handle, float(dt), int(rounds), sel, ctypes.byref(us)))
        return tuple(sel), us.value

    def close(self):
        self._fin()

    def set_fields(self, c0, phi):
        c0, phi = _fortran(c0), _fortran(phi)
        self._check_shape(c0)
        self._check_shape(phi)
        check(lib().csg_set_fields(self.handle, ptr(c0), ptr(phi)))
        # phi is not fingerprinted here (hashing 8 MB at 1024^2 took longer
        # than the upload): the next ensure_phi uploads its phi again
        self._phi_key = None

    def ensure_phi(self, phi, key=None):
        """Upload phi (as both fields) unless the phi with fingerprint `key`
        (computed here when None) is the resident one."""
        phi = _fortran(phi)
        k = _key(phi) if key is None else key
        if k != self._phi_key:
            self.set_fields(phi, phi)
            self._phi_key = k

    def _check_shape(self, a):
        if a.shape != (self.n, self.n):
            raise ValueError(f"expected a {self.n}x{self.n} array, got {a.shape}")

    def run(self, dt: float, steps: int, history_every: int = 1) -> np.ndarray:
        rows = steps // history_every + 1
        hist = np.zeros((rows, 6), order="F")
        rc = lib().csg_run(self.handle, float(dt), int(steps), int(history_every), ptr(hist))
        if rc == _capi.CSG_ERR_NONFINITE:
            raise NonFiniteCompositionError(int(lib().csg_nonfinite_step(self.handle)))
        check(rc)
        return hist

    def step(self, dt: float, nsteps: int = 1):
        check(lib().csg_step(self.handle, float(dt), int(nsteps)))

    def field(self) -> np.ndarray:
        c = np.empty((self.n, self.n), order="F")
        check(lib().csg_get_field(self.handle, ptr(c)))
        return c

    def time_passes(self, dt: float, nsteps: int) -> np.ndarray:
        ms = np.zeros(5)
        check(lib().csg_time_passes(self.handle, float(dt), int(nsteps), ptr(ms)))
        return ms

    def device_bytes(self) -> int:
        return int(lib().csg_device_bytes(self.handle))


def copy_bandwidth(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """Measured device copy bandwidth in GB/s (read + write bytes over time)
    of a streaming copy kernel over two buffers of nbytes bytes."""
    out = ctypes.c_double()
    check(lib().csg_copy_bandwidth(int(device), int(nbytes), int(reps), ctypes.byref(out)))
    return out.value


def _key(phi: np.ndarray):
    return (phi.shape, hash(phi.tobytes()))


# Device contexts reused by the function API: one per (n, spacing,
# precision, device), with the Parameters swapped in place (csg_set_params,
# no reallocation) when a call brings others; at most MAX_SESSIONS are kept,
# the least recently used one is closed first.
MAX_SESSIONS = 8
_sessions: "collections.OrderedDict" = collections.OrderedDict()


def _session(n, spacing, p, precision, device=0) -> Session:
    k = (int(n), float(spacing), precision, int(device))
    s = _sessions.get(k)
    if s is None:
        s = Session(n, spacing, p, precision, device)
        _sessions[k] = s
        while len(_sessions) > MAX_SESSIONS:
            _, old = _sessions.popitem(last=False)
            old.close()
    else:
        _sessions.move_to_end(k)
        if s.p != p:
            s.set_params(p)
    return s


@dataclasses.dataclass(frozen=True)
class Shape:
    """``geometry`` result (core_shell.jl:39-43): h, g, M and the phi they came from."""
    h: np.ndarray
    g: np.ndarray
    M: np.ndarray
    phi: np.ndarray
    # fingerprint of phi, taken once by geometry (None: taken when needed)
    key: tuple = dataclasses.field(default=None, repr=False, compare=False)


def geometry(phi, p: Parameters = Parameters(), *, precision: str = "fp64", device: int = 0) -> Shape:
    # a read-only snapshot, like the reference's (h, g, M) computed here: a
    # later change to the caller's array does not reach the device through
    # this Shape
    phi = np.array(phi, dtype=np.float64, order="F", copy=True)
    phi.setflags(write=False)
    key = _key(phi)
    s = _session(phi.shape[0], 1.0, p, precision, device)
    s.ensure_phi(phi, key)
    n = phi.shape[0]
    h, g, M = (np.empty((n, n), order="F") for _ in range(3))
    check(lib().csg_geometry(s.handle, ptr(h), ptr(g), ptr(M)))
    return Shape(h, g, M, phi, key)


def _ctx(shape: Shape, spacing: float, p: Parameters, precision: str, device: int) -> Session:
    s = _session(shape.phi.shape[0], spacing, p, precision, device)
    s.ensure_phi(shape.phi, shape.key)
    return s


def bulk_energy(c, shape: Shape, p: Parameters = Parameters(), *, precision="fp64", device=0):
    """``bulk_energy`` (core_shell.jl:45-49)."""
    s = _ctx(shape, 1.0, p, precision, device)
    c = _fortran(c)
    out = np.empty_like(c, order="F")
    check(lib().csg_bulk_energy(s.handle, ptr(c), ptr(out)))
    return out


def bulk_derivative(c, shape: Shape, p: Parameters = Parameters(), *, precision="fp64", device=0):
    """``bulk_derivative`` (core_shell.jl:51-55)."""
    s = _ctx(shape, 1.0, p, precision, device)
    c = _fortran(c)
    out = np.empty_like(c, order="F")
    check(lib().csg_bulk_derivative(s.handle, ptr(c), ptr(out)))
    return out


def chemical_potential(c, shape: Shape, grid: Grid, p: Parameters = Parameters(), *,
                       precision="fp64", device=0):
    """``chemical_potential`` (core_shell.jl:57-60)."""
    s = _ctx(shape, grid.spacing, p, precision, device)
    c = _fortran(c)
    mu = np.empty_like(c, order="F")
    check(lib().csg_chemical_potential(s.handle, ptr(c), ptr(mu)))
    return mu


def transport_rate(c, shape: Shape, grid: Grid, p: Parameters = Parameters(), *,
                   precision="fp64", device=0):
    """``transport_rate`` (core_shell.jl:62-70): (full n x n complex k-space
    rate, FFTW convention; dissipation)."""
    s = _ctx(shape, grid.spacing, p, precision, device)
    c = _fortran(c)
    re = np.empty_like(c, order="F")
    im = np.empty_like(c, order="F")
    d = np.zeros(1)
    check(lib().csg_transport_rate(s.handle, ptr(c), ptr(re), ptr(im), ptr(d)))
    return re + 1j * im, float(d[0])


def energy(c, shape: Shape, grid: Grid, p: Parameters = Parameters(), *, precision="fp64", device=0):
    """``energy`` (core_shell.jl:78-82)."""
    s = _ctx(shape, grid.spacing, p, precision, device)
    c = _fortran(c)
    F = np.zeros(1)
    check(lib().csg_energy(s.handle, ptr(c), ptr(F)))
    return float(F[0])


def step(c, shape: Shape, grid: Grid, p: Parameters, dt: float, *, precision="fp64", device=0):
    """``step`` (core_shell.jl:72-76)."""
    s = _session(shape.phi.shape[0], grid.spacing, p, precision, device)
    s.set_fields(c, shape.phi)
    s.step(dt, 1)
    return s.field()


def run_simulation(c0, phi, dt: float = 0.01, steps: int = 1000, spacing: float = 1.0,
                   p: Parameters = Parameters(), directory=None, *, precision: str = "fp64",
                   history_every: int = 1, device: int = 0, schedule=None):
    """``run_simulation`` (core_shell.jl:84-103): returns (c, history) with
    history rows [t, mean c, F, min c, max c, dissipation] of c_s for
    s = 0, history_every, ..., before c_s is advanced.  c0 is not modified.
    Raises NonFiniteCompositionError (and writes no files) on non-finite c_s.
    ``schedule``: as for Session (None: the default kernel schedules)."""
    c0 = _fortran(c0)
    s = _session(c0.shape[0], spacing, p, precision, device)
    if schedule == "cached":
        from . import schedules
        schedule = schedules.load(s.n, precision, s.schedule_info())
    s.set_schedule((0, 0, 0, 0) if schedule is None else schedule)
    s.set_fields(c0, phi)
    hist = s.run(dt, steps, history_every)
    c = s.field()
    if directory is not None:
        from .writers import writedlm
        os.makedirs(directory, exist_ok=True)
        writedlm(os.path.join(directory, "history.csv"), hist)
        writedlm(os.path.join(directory, "composition.csv"), c)
        writedlm(os.path.join(directory, "particle.csv"), _fortran(phi))
    return c, hist
