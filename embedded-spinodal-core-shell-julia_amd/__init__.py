"""MI355X (gfx950) core-shell Cahn-Hilliard step library: host package.

The compute path is ``libcsg.so`` (HIP kernels behind the C ABI of
``include/csg.h``); this package binds it with ctypes and mirrors the API of
``core_shell.jl`` (``Parameters``, ``make_grid``, ``initial_fields``,
``geometry``, ``bulk_energy``, ``bulk_derivative``, ``chemical_potential``,
``transport_rate``, ``step``, ``energy``, ``run_simulation``), plus the
adjoint (``adjoint.gradient``) and ensembles (``run_ensemble``).
"""
from .core_shell import (Grid, NonFiniteCompositionError, Parameters, Session, Shape,  # noqa: F401
                         bulk_derivative, bulk_energy, chemical_potential, copy_bandwidth, energy, geometry,
                         initial_fields, make_grid, run_simulation, step, transport_rate)
from ._capi import CsgError, lib  # noqa: F401
from .writers import julia_float, writedlm  # noqa: F401
from . import adjoint, schedules  # noqa: F401
from .ensemble import (Ensemble, EnsembleNonFiniteError, Member, ensemble_members,  # noqa: F401
                       member_fields, run_ensemble)
