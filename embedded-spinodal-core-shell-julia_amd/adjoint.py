"""Gradients of a final-morphology loss through the time loop (HIP adjoint).

    L = 1/2 ||c(T) - c_target||^2 + lam F(c(T))

with c(T) the result of `steps` applications of step (core_shell.jl:72-76)
from c0.  Returns L, dL/dc0 and dL/d(kappa, P, chi).  The backward sweep runs
on the GPU (csg_adjoint_forward / csg_adjoint_backward) with checkpoints every
`ckpt_every` steps (default ~sqrt(steps)); the loss's direct parameter terms
lam dF/dtheta use that F is affine in kappa, P and chi, evaluated with the
GPU energy.  FP64 only.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from ._capi import check, lib, ptr
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


def _energy_param_grads(c, phi, spacing, p: Parameters, device: int):
    """dF/dtheta for theta in (kappa, P, chi): F is affine in each."""
    out = []
    for name in ("kappa", "P", "chi"):
        vals = []
        for v in (1.0, 0.0):
            s = Session(c.shape[0], spacing, dataclasses.replace(p, **{name: v}), "fp64", device)
            s.set_fields(c, phi)
            vals.append(_energy(s, c))
            s.close()
        out.append(vals[0] - vals[1])
    return np.array(out)


def gradient(c0, phi, target, *, dt: float = 0.01, steps: int = 1000, lam: float = 0.0,
             spacing: float = 1.0, p: Parameters = Parameters(), ckpt_every: int | None = None,
             device: int = 0) -> Gradient:
    c0, phi, target = _fortran(c0), _fortran(phi), _fortran(target)
    n = c0.shape[0]
    K = ckpt_every or max(1, int(round(math.sqrt(steps))))
    s = Session(n, spacing, p, "fp64", device)
    try:
        s.set_fields(c0, phi)
        check(lib().csg_adjoint_forward(s.handle, float(dt), int(steps), int(K)))
        cT = s.field()
        L = 0.5 * float(np.sum((cT - target) ** 2))
        gbar = cT - target
        dtheta = np.zeros(3)
        if lam != 0.0:
            L += lam * _energy(s, cT)
            mu = np.empty((n, n), order="F")
            check(lib().csg_chemical_potential(s.handle, ptr(_fortran(cT)), ptr(mu)))
            gbar = gbar + lam * spacing ** 2 * mu
            dtheta += lam * _energy_param_grads(cT, phi, spacing, p, device)
        gbar = _fortran(gbar)
        dc0 = np.empty((n, n), order="F")
        dp = np.zeros(3)
        check(lib().csg_adjoint_backward(s.handle, ptr(gbar), ptr(dc0), ptr(dp)))
        dtheta += dp
    finally:
        s.close()
    return Gradient(L, dc0, float(dtheta[0]), float(dtheta[1]), float(dtheta[2]), cT)


def loss(c0, phi, target, *, dt: float = 0.01, steps: int = 1000, lam: float = 0.0,
         spacing: float = 1.0, p: Parameters = Parameters(), device: int = 0) -> float:
    """The forward loss alone (for finite-difference checks)."""
    c0, phi = _fortran(c0), _fortran(phi)
    s = Session(c0.shape[0], spacing, p, "fp64", device)
    try:
        s.set_fields(c0, phi)
        s.step(dt, steps)
        cT = s.field()
        val = 0.5 * float(np.sum((cT - np.asarray(target)) ** 2))
        if lam != 0.0:
            val += lam * _energy(s, cT)
    finally:
        s.close()
    return val
