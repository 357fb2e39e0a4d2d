"""``lpt`` entry point (src/jax_fli/pm/lpt.py:19-214): first-order LPT at the lattice,
dx = D(a) F(delta_L), p = a^2 f(a) E(a) dx (jaxpm.pm.lpt, lpt.py:119-127), as GRID_RELATIVE
ParticleFields; optional density painting of the result."""

from __future__ import annotations

import numpy as np
import torch

from .. import autodiff
from .. import cosmology as _c
from ..fields import DensityField, FieldStatus, ParticleField, PositionUnit
from ..painting import PaintingOptions
from ._resolve_geometry import resolve_ts_geometry


def _lattice_particles(initial_particles, mesh_size):
    """Particle grid (npx, npy, npz) of user-given initial particles (GRID_ABSOLUTE, (npx, npy, npz, 3)),
    which must be the regular sub-lattice (i, j, k) * mesh / np: the displacements lpt returns are
    GRID_RELATIVE to that lattice (lpt.py:136-153), which the relative paint reconstructs from the
    array's shape. Other particle sets are refused."""
    arr = getattr(initial_particles, "array", initial_particles)
    arr = torch.as_tensor(arr)
    if arr.ndim != 4 or arr.shape[-1] != 3:
        raise ValueError("initial_particles must have shape (X, Y, Z, 3)")
    npart = tuple(int(n) for n in arr.shape[:3])
    if any(m % n for m, n in zip(mesh_size, npart)):
        raise NotImplementedError("initial particles must form a lattice whose size divides the mesh")
    r = [m // n for m, n in zip(mesh_size, npart)]
    axes = [torch.arange(n, dtype=torch.float32, device=arr.device) * ri for n, ri in zip(npart, r)]
    lattice = torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1)
    if not torch.equal(arr.to(torch.float32), lattice):
        raise NotImplementedError("only the regular particle sub-lattice is supported as initial_particles")
    return npart


def lpt(cosmo, initial_field: DensityField, *, ts=None, nb_shells=None, density_widths=None, order: int = 1,
        initial_particles=None, painting: PaintingOptions = PaintingOptions(target="particles"),
        equal_vol: bool = False, min_width: float = 50.0):
    if not isinstance(initial_field, DensityField):
        raise TypeError("initial_field must be a DensityField instance.")
    if order not in (1, 2):
        raise ValueError("order must be either 1 or 2.")
    if initial_field.status != FieldStatus.INITIAL_FIELD:
        raise ValueError("initial_field must have status FieldStatus.INITIAL_FIELD.")
    if order == 2:
        raise NotImplementedError("second-order LPT is not implemented on the pmx engine")
    particles = None
    if initial_particles is not None:
        particles = _lattice_particles(initial_particles, initial_field.mesh_size)
    cosmo_in = cosmo
    cosmo, _, _ = autodiff.split_cosmo(cosmo_in)
    ts_resolved, r_centers, widths, is_lightcone = resolve_ts_geometry(
        cosmo, initial_field, painting=painting, ts=ts, nb_shells=nb_shells, density_widths=density_widths,
        equal_vol=equal_vol, min_width=min_width)
    if is_lightcone:
        raise NotImplementedError("lightcone LPT (per-particle scale factors) is not supported")
    if painting.target not in ("particles", "density"):
        raise ValueError(f"Painting target {painting.target} is incompatible with snapshot mode. Use 'particles' or 'density'.")
    a = float(np.asarray(ts_resolved, dtype=np.float32).reshape(-1)[0])
    # dx = D(a) F(delta), p = a^2 f E dx; differentiable in the field and in tensor-valued parameters
    dx, p = autodiff.lpt1(initial_field.array, cosmo_in, a, particles=particles)
    meta = dict(status=FieldStatus.LPT1, scale_factors=a, comoving_centers=float(_c.radial_comoving_distance(cosmo, a)),
                z_sources=1.0 / a - 1.0, unit=PositionUnit.GRID_RELATIVE)
    dx_field = ParticleField.FromDensityMetadata(array=dx, field=initial_field, **meta)
    p_field = ParticleField.FromDensityMetadata(array=p, field=initial_field, **meta)
    if painting.target == "density":
        dx_field = dx_field.paint(weights=painting.weights, chunk_size=painting.chunk_size,
                                  batch_size=painting.batch_size)
    return dx_field, p_field
