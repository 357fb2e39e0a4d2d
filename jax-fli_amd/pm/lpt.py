"""``lpt`` entry point (src/jax_fli/pm/lpt.py:19-214): LPT of order 1 or 2 at the particle lattice
(jaxpm.pm.lpt, lpt.py:119-127; upstream JaxPM, Jenkins 2010 eqs. 2-3),
    dx = D(a) F(delta_L) [+ 3/7 D2(a)/D2(1) F(delta2)],  p = a^2 f E D F [+ a^2 f2 E 3/7 D2 F(delta2)],
delta2 = sum_{i<j} (phi_ii phi_jj - phi_ij^2), as GRID_RELATIVE ParticleFields: at one scale factor
(snapshot), or on the lightcone with each particle at the scale factor of its comoving distance from
the observer (utils.py:234-244), then optionally painted onto flat or spherical shells or a density."""

from __future__ import annotations

import numpy as np
import torch

from .. import autodiff, engine
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


def _lightcone_table(cosmo, field: DensityField, particles, ntab: int = 4096):
    """(coef (4, ntab), r_max, r (X,Y,Z) float64) of the lattice particles' comoving distances: columns
    D(a(r)), a^2 f E, 3/7 D2/D2(1), a^2 f2 E on the uniform grid [0, r_max]."""
    mesh = np.asarray(field.mesh_size)
    npart = np.asarray(particles if particles is not None else field.mesh_size)
    cell = np.asarray(field.box_size, dtype=np.float64) / mesh
    obs = np.asarray(field.observer_position_mpc, dtype=np.float64)
    step = (mesh // npart) * cell
    axes = [np.arange(n) * s - o for n, s, o in zip(npart, step, obs)]
    r = np.sqrt(axes[0][:, None, None] ** 2 + axes[1][None, :, None] ** 2 + axes[2][None, None, :] ** 2)
    r_max = float(r.max()) * (1.0 + 1e-12) + 1e-9
    rg = np.linspace(0.0, r_max, ntab)
    a = _c.a_of_chi(cosmo, rg)
    E = _c.E(cosmo, a)
    D2 = 3.0 / 7.0 * _c.growth_factor_second(cosmo, a)
    coef = np.stack([_c.growth_factor(cosmo, a), a**2 * _c.growth_rate(cosmo, a) * E, D2,
                     a**2 * _c.growth_rate_second(cosmo, a) * E])
    return coef, r_max, r, rg, a


def lpt(cosmo, initial_field: DensityField, *, ts=None, nb_shells=None, density_widths=None, order: int = 1,
        initial_particles=None, painting: PaintingOptions = PaintingOptions(target="particles"),
        equal_vol: bool = False, min_width: float = 50.0):
    if not isinstance(initial_field, DensityField):
        raise TypeError("initial_field must be a DensityField instance.")
    if order not in (1, 2):
        raise ValueError("order must be either 1 or 2.")
    if initial_field.status != FieldStatus.INITIAL_FIELD:
        raise ValueError("initial_field must have status FieldStatus.INITIAL_FIELD.")
    particles = None
    if initial_particles is not None:
        particles = _lattice_particles(initial_particles, initial_field.mesh_size)
    cosmo_in = cosmo
    cosmo, _, tensors = autodiff.split_cosmo(cosmo_in)
    ts_resolved, r_centers, widths, is_lightcone = resolve_ts_geometry(
        cosmo, initial_field, painting=painting, ts=ts, nb_shells=nb_shells, density_widths=density_widths,
        equal_vol=equal_vol, min_width=min_width)
    if not is_lightcone and painting.target not in ("particles", "density"):
        raise ValueError(f"Painting target {painting.target} is incompatible with snapshot mode. Use 'particles' or 'density'.")
    ic = initial_field.array
    if (order == 2 or is_lightcone) and (ic.requires_grad or tensors):
        raise NotImplementedError("gradients of second-order or lightcone LPT are not implemented")
    if is_lightcone:
        coef, r_max, r, rg, a_tab = _lightcone_table(cosmo, initial_field, particles)
        dx, p = engine.lpt(ic.contiguous(), order, coef, particles=particles, r_max=r_max,
                           box=initial_field.box_size, observer=initial_field.observer_position_mpc)
        a = np.interp(r, rg, a_tab).astype(np.float32)[..., None]
        meta = dict(status=FieldStatus.LIGHTCONE, scale_factors=a, comoving_centers=r.astype(np.float32)[..., None],
                    z_sources=(1.0 / a - 1.0), unit=PositionUnit.GRID_RELATIVE)
    else:
        a = float(np.asarray(ts_resolved, dtype=np.float32).reshape(-1)[0])
        if order == 1:
            # dx = D(a) F(delta), p = a^2 f E dx; differentiable in the field and in tensor-valued parameters
            dx, p = autodiff.lpt1(ic, cosmo_in, a, particles=particles)
        else:
            E = float(_c.E(cosmo, a))
            coef = [float(_c.growth_factor(cosmo, a)), float(a**2 * _c.growth_rate(cosmo, a) * E),
                    float(3.0 / 7.0 * _c.growth_factor_second(cosmo, a)),
                    float(a**2 * _c.growth_rate_second(cosmo, a) * E)]
            dx, p = engine.lpt(ic.contiguous(), 2, coef, particles=particles)
        meta = dict(status=FieldStatus.LPT1 if order == 1 else FieldStatus.LPT2, scale_factors=a,
                    comoving_centers=float(_c.radial_comoving_distance(cosmo, a)), z_sources=1.0 / a - 1.0,
                    unit=PositionUnit.GRID_RELATIVE)
    dx_field = ParticleField.FromDensityMetadata(array=dx, field=initial_field, **meta)
    p_field = ParticleField.FromDensityMetadata(array=p, field=initial_field, **meta)
    if is_lightcone and painting.target in ("flat", "spherical"):
        if painting.target == "flat":
            dx_field = dx_field.paint_2d(center=r_centers, density_plane_width=widths, weights=painting.weights,
                                         batch_size=painting.batch_size)
        else:
            dx_field = dx_field.paint_spherical(
                center=r_centers, density_plane_width=widths, scheme=painting.scheme, weights=painting.weights,
                kernel_width_arcmin=painting.kernel_width_arcmin,
                smoothing_interpretation=painting.smoothing_interpretation, paint_nside=painting.paint_nside,
                ud_grade_power=painting.ud_grade_power, ud_grade_order_in=painting.ud_grade_order_in,
                ud_grade_order_out=painting.ud_grade_order_out, ud_grade_pess=painting.ud_grade_pess,
                batch_size=painting.batch_size)
        a_snap = _c.a_of_chi(cosmo, np.asarray(r_centers, dtype=np.float64))
        dx_field = dx_field.replace(scale_factors=a_snap, z_sources=1.0 / a_snap - 1.0)
        if dx_field.is_batched():
            dx_field = dx_field[::-1]
    elif is_lightcone and painting.target not in ("particles", "density"):
        raise ValueError(f"Unknown painting target {painting.target}.")
    if painting.target == "density":
        dx_field = dx_field.paint(weights=painting.weights, chunk_size=painting.chunk_size,
                                  batch_size=painting.batch_size)
    return dx_field, p_field
