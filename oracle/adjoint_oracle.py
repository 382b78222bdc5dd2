"""Oracle (test infrastructure only): CPU adjoint of the reference step.

Only ``tests/`` may import this module, as the reference the HIP adjoint's
gradients are compared against.

The step of ``core_shell.jl:72-76`` written with real linear operators
(spectral derivatives D_x, D_y with the Nyquist mode zeroed, D^T = -D;
Lambda = IFFT K^2 FFT and P_den = IFFT (1/den) FFT, both symmetric;
M = diag(M0 h(phi))):

    mu(c) = f'(c) + 2 kappa Lambda c
    r(c)  = D_x M D_x mu + D_y M D_y mu          (= Re ifft(transport_rate))
    c'    = c + dt P_den r(c)

so its vector-Jacobian product with gbar = dL/dc' is

    v     = P_den gbar
    r_v   = D_x M D_x v + D_y M D_y v            (the same flux operator)
    dL/dc = gbar + dt (f''(c) r_v + 2 kappa Lambda r_v)

and the per-step parameter terms are
    dL/dP     += -dt chi <r_v, g(phi)>,   dL/dchi += -dt P <r_v, g(phi)>,
    dL/dkappa += dt <r_v, 2 Lambda c> + dt <gbar, IFFT(-2Q K^4 dt / den^2 rate^)>.

Loss (the 512^2 adjoint workload): L = 1/2 ||c(T) - c*||^2 + lam F(c(T)),
dL/dc(T) = c(T) - c* + lam spacing^2 mu(c(T)) (the derivative pinned by
checks.jl:7-11), and the direct terms lam dF/dtheta at c(T).

Pinned by central finite differences of the oracle's own forward loss
(tests/test_adjoint.py), which restates the reference step.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import core_shell_oracle as o


def bulk_second_derivative(c, shape, p):
    """d/dc of bulk_derivative (core_shell.jl:53-54)."""
    h = shape["h"]
    return (2 * p.A * (1 - h)
            + 2 * p.B * h * ((2 * c - p.cb - p.cc) ** 2 + 2 * (c - p.cb) * (c - p.cc)))


def flux_operator(v, shape, grid):
    """D_x M D_x v + D_y M D_y v (real field)."""
    vh = o.fft(v)
    gx = np.real(o.ifft(1j * grid["kx"] * vh))
    gy = np.real(o.ifft(1j * grid["ky"] * vh))
    M = shape["M"]
    return np.real(o.ifft(1j * grid["kx"] * o.fft(M * gx) + 1j * grid["ky"] * o.fft(M * gy)))


def step_vjp(c, gbar, shape, grid, p, dt):
    """(dL/dc_s, (dkappa, dP, dchi)) for one step given gbar = dL/dc_{s+1}."""
    k2 = grid["k2"]
    den = 1 + 2 * p.Q * p.kappa * k2 ** 2 * dt
    v = np.real(o.ifft(o.fft(gbar) / den))
    rv = flux_operator(v, shape, grid)
    u = bulk_second_derivative(c, shape, p) * rv + 2 * p.kappa * np.real(o.ifft(k2 * o.fft(rv)))
    gprev = gbar + dt * u
    g = shape["g"]
    s_rg = np.sum(rv * g)
    d_P = -dt * p.chi * s_rg
    d_chi = -dt * p.P * s_rg
    lam_c = np.real(o.ifft(k2 * o.fft(c)))
    rate, _ = o.transport_rate(c, shape, grid, p)
    dden = np.real(o.ifft(rate * (-2 * p.Q * k2 ** 2 * dt) / den ** 2))
    d_kappa = dt * np.sum(rv * 2 * lam_c) + dt * np.sum(gbar * dden)
    return gprev, np.array([d_kappa, d_P, d_chi])


@dataclasses.dataclass
class Loss:
    target: np.ndarray
    lam: float = 0.0


def loss_value(cT, shape, grid, p, loss: Loss):
    return 0.5 * np.sum((cT - loss.target) ** 2) + loss.lam * o.energy(cT, shape, grid, p)


def energy_param_grads(c, shape, grid, p):
    """dF/d(kappa, P, chi) at c (F is affine in each)."""
    sp2 = grid["spacing"] ** 2
    ch = o.fft(c)
    d_kappa = np.sum(grid["k2"] * np.abs(ch) ** 2) / c.size * sp2
    d_P = np.sum((1 - p.chi * c) * shape["g"]) * sp2
    d_chi = -p.P * np.sum(c * shape["g"]) * sp2
    return np.array([d_kappa, d_P, d_chi])


def forward(c0, shape, grid, p, dt, steps):
    cs = [np.array(c0, dtype=np.float64)]
    for _ in range(steps):
        cs.append(o.step(cs[-1], shape, grid, p, dt))
    return cs


def gradient(c0, phi, p, dt, steps, loss: Loss, spacing=1.0):
    """(L, dL/dc0, dL/d(kappa, P, chi)) by the adjoint sweep over stored states."""
    grid = o.make_grid(c0.shape[0], spacing)
    shape = o.geometry(phi, p)
    cs = forward(c0, shape, grid, p, dt, steps)
    cT = cs[-1]
    L = loss_value(cT, shape, grid, p, loss)
    gbar = (cT - loss.target) + loss.lam * spacing ** 2 * o.chemical_potential(cT, shape, grid, p)
    dtheta = loss.lam * energy_param_grads(cT, shape, grid, p)
    for s in range(steps - 1, -1, -1):
        gbar, d = step_vjp(cs[s], gbar, shape, grid, p, dt)
        dtheta = dtheta + d
    return L, gbar, dtheta


def loss_of(c0, phi, p, dt, steps, loss: Loss, spacing=1.0):
    grid = o.make_grid(c0.shape[0], spacing)
    shape = o.geometry(phi, p)
    cT = forward(c0, shape, grid, p, dt, steps)[-1]
    return loss_value(cT, shape, grid, p, loss)
