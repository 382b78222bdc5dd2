"""Oracle (test infrastructure only): FP64 numpy restatement of ``core_shell.jl``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module, and only as the reference results are compared against
or as the timed CPU baseline;
the product path (the HIP library) never calls it.

Arrays are indexed like the Julia matrices: ``a[i, j]`` with ``i`` the first
(contiguous in Julia) index, along which ``kx`` varies.  Every function follows
the reference op for op, with full complex 2D FFTs (``numpy.fft``, pocketfft,
FP64) standing in for FFTW's ``fft``/``ifft`` (unnormalised forward, 1/N²
inverse), so the CPU work per iteration matches the reference's (17 complex
N×N FFTs per loop iteration, ``core_shell.jl:90-95``).

Pinning (tests/test_oracle.py): F(t=0), F(t=10), mass drift and final ranges of
``check-results.txt:2-7`` with the Julia-exact noise of ``julia_rng``;
``checks.jl:7-14`` identities; closed forms (phi=0 fixed point, single cosine
mode growth factor, uniform fixed point); numpy.fft against a brute-force DFT.
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from . import julia_rng


@dataclasses.dataclass(frozen=True)
class Parameters:
    """``core_shell.jl:4-15``."""
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


def fftfreq(n: int, fs: float) -> np.ndarray:
    """``FFTW.fftfreq(n, fs)``: [0, 1, ..., ceil(n/2)-1, -floor(n/2), ..., -1] * fs/n."""
    m = np.arange(n, dtype=np.float64)
    m[m >= (n + 1) // 2] -= n
    return m * (fs / n)


def make_grid(n: int, spacing: float):
    """``core_shell.jl:17-25``: derivative wavenumbers with the Nyquist mode zeroed,
    ``k2`` with it kept."""
    k = 2 * np.pi * fftfreq(n, 1 / spacing)
    dk = k.copy()
    if n % 2 == 0:
        dk[n // 2] = 0.0
    kx = dk.reshape(n, 1)
    ky = dk.reshape(1, n)
    k2 = (k ** 2).reshape(n, 1) + (k ** 2).reshape(1, n)
    return dict(kx=kx, ky=ky, k2=k2, spacing=spacing)


def particle(n: int, spacing: float = 1.0, radius: float = 70.0, width: float = 4.0):
    """phi of ``core_shell.jl:29-30`` (0-based centre index n÷2)."""
    idx = (np.arange(n) - n // 2) * spacing
    r = np.hypot(idx.reshape(n, 1), idx.reshape(1, n))
    return 0.5 * (1 - np.tanh((r - radius) / width))


def initial_fields(n: int, spacing=1.0, radius=70.0, width=4.0, seed=5749, noise=None):
    """``core_shell.jl:27-37``.  ``noise`` (uniform [0,1) field, Julia order)
    defaults to ``rand(MersenneTwister(seed), n, n)``."""
    phi = particle(n, spacing, radius, width)
    inside = phi >= 0.5
    if noise is None:
        noise = julia_rng.MersenneTwister(seed).rand(n, n)
    noise = 0.01 * (2 * noise - 1)
    # mean(noise[inside]): Julia's logical indexing gathers in column-major order
    sel = noise.T[inside.T]
    noise = noise - julia_mean(sel)
    c = np.full((n, n), 0.5)
    c[inside] += noise[inside]
    return c, phi


def julia_sum(x: np.ndarray) -> float:
    """``sum`` of a Float64 vector in the order of Julia's ``mapreduce_impl``:
    pairwise halving above 1024 values, sequential below (Julia vectorises the
    sequential part, so results may differ in the last ulp)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    return _pairwise(x, 0, len(x))


def _pairwise(x, lo, hi):
    n = hi - lo
    if n <= 1024:
        s = 0.0
        for v in x[lo:hi]:
            s += v
        return s
    mid = lo + ((n - 1) >> 1) + 1
    return _pairwise(x, lo, mid) + _pairwise(x, mid, hi)


def julia_mean(x: np.ndarray) -> float:
    x = np.asarray(x).ravel(order="F")
    return julia_sum(x) / len(x)


def geometry(phi: np.ndarray, p: Parameters):
    """``core_shell.jl:39-43``."""
    h = phi ** 3 * (10 - 15 * phi + 6 * phi ** 2)
    g = phi ** 2 * (1 - phi) ** 2
    return dict(h=h, g=g, M=p.mobility * h)


def bulk_energy(c, shape, p: Parameters):
    """``core_shell.jl:45-49``."""
    h, g = shape["h"], shape["g"]
    return (p.A * (1 - h) * (c - p.ca) ** 2
            + p.B * h * (c - p.cb) ** 2 * (c - p.cc) ** 2
            + p.P * (1 - p.chi * c) * g)


def bulk_derivative(c, shape, p: Parameters):
    """``core_shell.jl:51-55``."""
    h, g = shape["h"], shape["g"]
    return (2 * p.A * (1 - h) * (c - p.ca)
            + 2 * p.B * h * (c - p.cb) * (c - p.cc) * (2 * c - p.cb - p.cc)
            - p.chi * p.P * g)


# 2D transforms (FFTW's fft/ifft convention: unnormalised forward, 1/N^2
# inverse), numpy pocketfft on one thread.
def fft(x):
    return np.fft.fft2(x)


def ifft(x):
    return np.fft.ifft2(x)


def chemical_potential(c, shape, grid, p: Parameters):
    """``core_shell.jl:57-60``."""
    return bulk_derivative(c, shape, p) + np.real(ifft(2 * p.kappa * grid["k2"] * fft(c)))


def transport_rate(c, shape, grid, p: Parameters):
    """``core_shell.jl:62-70``: (full N×N k-space rate, dissipation)."""
    mu_hat = fft(chemical_potential(c, shape, grid, p))
    grad_x = np.real(ifft(1j * grid["kx"] * mu_hat))
    grad_y = np.real(ifft(1j * grid["ky"] * mu_hat))
    divergence_hat = (1j * grid["kx"] * fft(shape["M"] * grad_x)
                      + 1j * grid["ky"] * fft(shape["M"] * grad_y))
    dissipation = np.sum(shape["M"] * (grad_x ** 2 + grad_y ** 2)) * grid["spacing"] ** 2
    return divergence_hat, dissipation


def step(c, shape, grid, p: Parameters, dt: float):
    """``core_shell.jl:72-76``."""
    rate, _ = transport_rate(c, shape, grid, p)
    denominator = 1 + 2 * p.Q * p.kappa * grid["k2"] ** 2 * dt
    return np.real(ifft(fft(c) + dt * rate / denominator))


def energy(c, shape, grid, p: Parameters):
    """``core_shell.jl:78-82``."""
    gradient_energy = p.kappa * np.sum(grid["k2"] * np.abs(fft(c)) ** 2) / c.size
    return (np.sum(bulk_energy(c, shape, p)) + gradient_energy) * grid["spacing"] ** 2


class NonFiniteError(RuntimeError):
    pass


def run_simulation(c0, phi, dt=0.01, steps=1000, spacing=1.0, p: Parameters = Parameters(),
                   directory=None, history_every=1):
    """``core_shell.jl:84-103``.  History rows are [t, mean c, F, min c, max c,
    dissipation] of c_s before it is advanced; ``history_every`` keeps every
    k-th row (1 = the reference's every-step record)."""
    n = c0.shape[0]
    grid = make_grid(n, spacing)
    shape = geometry(phi, p)
    c = np.array(c0, dtype=np.float64, copy=True)
    history = np.zeros((steps + 1, 6))
    for s in range(steps + 1):
        _, dissipation = transport_rate(c, shape, grid, p)
        history[s] = [s * dt, c.mean(), energy(c, shape, grid, p), c.min(), c.max(), dissipation]
        if not np.all(np.isfinite(c)):
            raise NonFiniteError(f"Non-finite composition at step {s}")
        if s < steps:
            c = step(c, shape, grid, p, dt)
    if directory is not None:
        os.makedirs(directory, exist_ok=True)
    return c, history[::history_every]


def run_simulation_fast(c0, phi, dt=0.01, steps=1000, spacing=1.0, p: Parameters = Parameters()):
    """Same loop, one ``transport_rate`` per iteration (the reference evaluates it
    twice on the same c: ``core_shell.jl:91`` and ``:73``).  Identical results;
    used to produce oracle fields at larger sizes in reasonable CPU time."""
    n = c0.shape[0]
    grid = make_grid(n, spacing)
    shape = geometry(phi, p)
    den = 1 + 2 * p.Q * p.kappa * grid["k2"] ** 2 * dt
    c = np.array(c0, dtype=np.float64, copy=True)
    history = np.zeros((steps + 1, 6))
    for s in range(steps + 1):
        rate, dissipation = transport_rate(c, shape, grid, p)
        history[s] = [s * dt, c.mean(), energy(c, shape, grid, p), c.min(), c.max(), dissipation]
        if not np.all(np.isfinite(c)):
            raise NonFiniteError(f"Non-finite composition at step {s}")
        if s < steps:
            c = np.real(ifft(fft(c) + dt * rate / den))
    return c, history
