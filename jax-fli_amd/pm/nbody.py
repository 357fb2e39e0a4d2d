"""``nbody`` entry point (src/jax_fli/pm/nbody.py:28-171): status and shape checks, lightcone
geometry, integration, near-to-far ordering and lightcone metadata."""

from __future__ import annotations

import dataclasses

import numpy as np

from .. import autodiff
from .. import cosmology as _c
from ..fields import FieldStatus, ParticleField
from ..painting import PaintingOptions
from ._resolve_geometry import resolve_ts_geometry
from .integrate import integrate
from .interp import NoInterp
from .solvers import AbstractNBodySolver, EfficientDriftDoubleKick


def nbody(cosmo, dx_field: ParticleField, p_field: ParticleField, *, t0: float = 0.1, t1: float = 1.0,
          dt0: float = 0.05, ts=None, nb_shells: int | None = None, density_widths=None,
          solver: AbstractNBodySolver | None = None, adjoint: str = "checkpointed", checkpoints: int | None = None,
          equal_vol: bool = False, min_width: float = 50.0):
    if solver is None:
        solver = EfficientDriftDoubleKick(interp_kernel=NoInterp(painting=PaintingOptions(target="particles")))
    assert dx_field.status in (FieldStatus.LPT1, FieldStatus.LPT2), "dx_field must have status FieldStatus.LPT1 or FieldStatus.LPT2."
    assert p_field.status in (FieldStatus.LPT1, FieldStatus.LPT2), "p_field must have status FieldStatus.LPT1 or FieldStatus.LPT2."
    if dx_field.mesh_size != p_field.mesh_size:
        raise ValueError("dx_field and p_field must have matching mesh_size")
    if dx_field.box_size != p_field.box_size:
        raise ValueError("dx_field and p_field must have matching box_size")
    lpt_t0 = np.atleast_1d(np.asarray(dx_field.scale_factors, dtype=np.float64)).squeeze()
    if lpt_t0.ndim != 0:
        raise ValueError("Starting scale factor t0 and LPT fields' scale factor must be scalars.")
    if np.float32(lpt_t0) != np.float32(t0):
        raise ValueError(f"Starting scale factor t0={t0} does not match LPT fields' scale factor {lpt_t0}.")
    if ts is None and nb_shells is None:
        ts = np.array([t1])
    cosmo_f, _, _ = autodiff.split_cosmo(cosmo)
    ts_resolved, r_centers, widths, _ = resolve_ts_geometry(cosmo_f, dx_field, painting=solver.interp_kernel.painting,
                                                            ts=ts, nb_shells=nb_shells, density_widths=density_widths,
                                                            equal_vol=equal_vol, min_width=min_width)
    interp = solver.interp_kernel.update_geometry(ts=ts_resolved, r_centers=np.asarray(r_centers),
                                                  density_widths=np.asarray(widths),
                                                  max_comoving_distance=dx_field.max_comoving_radius)
    solver = dataclasses.replace(solver, interp_kernel=interp)
    lc = integrate(dx_field, p_field, cosmo, ts_resolved, solver, t0=t0, t1=t1, dt0=dt0, adjoint=adjoint,
                   checkpoints=checkpoints)
    r_centers = np.atleast_1d(np.asarray(r_centers, dtype=np.float64))
    widths = np.atleast_1d(np.asarray(widths, dtype=np.float64))
    if ts_resolved.ndim == 1 and ts_resolved.size == 1:
        lc = lc.replace(array=lc.array[0], scale_factors=lc.scale_factors[0])
        r_centers, widths = r_centers[0], widths[0]
    elif lc.is_batched():
        lc = lc.replace(array=lc.array.flip(0), scale_factors=np.asarray(lc.scale_factors)[::-1].copy())
        r_centers, widths = r_centers[::-1].copy(), widths[::-1].copy()
    sf = np.asarray(lc.scale_factors, dtype=np.float64)
    return lc.replace(z_sources=_c.a2z(sf), comoving_centers=r_centers, density_width=widths,
                      status=FieldStatus.LIGHTCONE)
