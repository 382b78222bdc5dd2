"""Gradients of a final-morphology loss through the time loop (HIP adjoint).

    L = 1/2 ||c(T) - c_target||^2 + lam F(c(T))

with c(T) the result of `steps` applications of step (core_shell.jl:72-76)
from c0.  Returns L, dL/dc0 and dL/d(kappa, P, chi).  The backward sweep runs
on the GPU (csg_adjoint_forward / csg_adjoint_backward) with checkpoints every
`ckpt_every` steps.  By default every state is kept when the run's states fit
KEEP_ALL_BYTES of device memory (the backward sweep then recomputes nothing),
else one checkpoint per ~sqrt(steps) steps; the loss's direct parameter terms
lam dF/dtheta use that F is affine in kappa, P and chi, evaluated with the
GPU energy.  FP64 only.
"""
from __future__ import annotations

import collections
import dataclasses
import math

import numpy as np

from ._capi import CSG_ERR_HIP, check, lib, ptr

# Device memory the default checkpointing may spend on keeping every state of
# the forward run: (steps + 1) FP64 fields of N^2 values (2000 steps of 512^2
# take 4.2 GB).
KEEP_ALL_BYTES = 32 << 30


def default_ckpt_every(n: int, steps: int) -> int:
    """1 (keep every state) when (steps + 1) N^2 doubles fit KEEP_ALL_BYTES,
    else ~sqrt(steps)."""
    if (steps + 1) * n * n * 8 <= KEEP_ALL_BYTES:
        return 1
    return max(1, int(round(math.sqrt(steps))))
from .core_shell import Parameters, Session, _fortran


@dataclasses.dataclass
class Gradient:
    loss: float
    dc0: np.ndarray
    dkappa: float
    dP: float
    dchi: float
    cT: np.ndarray


def _energy(s: Session, c) -> float:
    F = np.zeros(1)
    check(lib().csg_energy(s.handle, ptr(_fortran(c)), ptr(F)))
    return float(F[0])


def _energy_param_grads(s: Session, c, p: Parameters):
    """dF/dtheta for theta in (kappa, P, chi): F is affine in each, so each is
    F(theta = 1) - F(theta = 0).  Uses the session with its parameters swapped
    in place and restores them."""
    out = []
    try:
        for name in ("kappa", "P", "chi"):
            vals = []
            for v in (1.0, 0.0):
                s.set_params(dataclasses.replace(p, **{name: v}))
                vals.append(_energy(s, c))
            out.append(vals[0] - vals[1])
    finally:
        s.set_params(p)
    return np.array(out)


# One FP64 context per (n, spacing, device), kept between calls: the checkpoint
# and segment buffers and the captured adjoint graphs are then reused.  At
# most MAX_CONTEXTS are kept (their checkpoints can hold gigabytes); the least
# recently used one is closed first.
MAX_CONTEXTS = 2
_contexts: "collections.OrderedDict" = collections.OrderedDict()


def _context(n: int, spacing: float, p: Parameters, device: int) -> Session:
    key = (int(n), float(spacing), int(device))
    s = _contexts.get(key)
    if s is None:
        s = _contexts[key] = Session(n, spacing, p, "fp64", device)
        while len(_contexts) > MAX_CONTEXTS:
            _, old = _contexts.popitem(last=False)
            old.close()
    else:
        _contexts.move_to_end(key)
        if s.p != p:
            s.set_params(p)
    return s


def gradient(c0, phi, target, *, dt: float = 0.01, steps: int = 1000, lam: float = 0.0,
             spacing: float = 1.0, p: Parameters = Parameters(), ckpt_every: int | None = None,
             device: int = 0) -> Gradient:
    c0, phi, target = _fortran(c0), _fortran(phi), _fortran(target)
    n = c0.shape[0]
    K = ckpt_every or default_ckpt_every(n, steps)
    s = _context(n, spacing, p, device)
    s.set_fields(c0, phi)
    rc = lib().csg_adjoint_forward(s.handle, float(dt), int(steps), int(K))
    if rc == CSG_ERR_HIP and ckpt_every is None and K == 1:
        # not enough free device memory to keep every state: checkpoint instead
        # (the failed call has freed the checkpoint buffers it allocated; the
        # retry keeps only the ~sqrt(steps) it needs)
        K = max(1, int(round(math.sqrt(steps))))
        s.set_fields(c0, phi)
        rc = lib().csg_adjoint_forward(s.handle, float(dt), int(steps), int(K))
    check(rc)
    cT = s.field()
    L = 0.5 * float(np.sum((cT - target) ** 2))
    gbar = cT - target
    dtheta = np.zeros(3)
    if lam != 0.0:
        L += lam * _energy(s, cT)
        mu = np.empty((n, n), order="F")
        check(lib().csg_chemical_potential(s.handle, ptr(_fortran(cT)), ptr(mu)))
        gbar = gbar + lam * spacing ** 2 * mu
    gbar = _fortran(gbar)
    dc0 = np.empty((n, n), order="F")
    dp = np.zeros(3)
    check(lib().csg_adjoint_backward(s.handle, ptr(gbar), ptr(dc0), ptr(dp)))
    dtheta += dp
    if lam != 0.0:
        # after the backward sweep: swapping parameters discards the checkpoints
        dtheta += lam * _energy_param_grads(s, cT, p)
    return Gradient(L, dc0, float(dtheta[0]), float(dtheta[1]), float(dtheta[2]), cT)


def loss(c0, phi, target, *, dt: float = 0.01, steps: int = 1000, lam: float = 0.0,
         spacing: float = 1.0, p: Parameters = Parameters(), device: int = 0) -> float:
    """The forward loss alone (for finite-difference checks); the same kernels
    as the adjoint's forward run, through the captured loop."""
    c0, phi = _fortran(c0), _fortran(phi)
    s = _context(c0.shape[0], spacing, p, device)
    s.set_fields(c0, phi)
    s.run(dt, steps, history_every=max(1, steps))
    cT = s.field()
    val = 0.5 * float(np.sum((cT - np.asarray(target)) ** 2))
    if lam != 0.0:
        val += lam * _energy(s, cT)
    return val
