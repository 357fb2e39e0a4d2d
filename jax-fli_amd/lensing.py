"""Born convergence from a lightcone of density shells: ``born`` (src/jax_fli/lensing/born.py:14-89,
src/jax_fli/_src/lensing/_born.py:29-190, _metadata.py:17-47), spherical (HEALPix) and flat-sky.

kappa_k = sum_s C[k, s] (rho_s / mean(rho_s) - 1), with the host building C in float64:
C[k, s] = 3/2 Om (H0/c)^2 d_r r_s / a_s sum_z nz_k(z) w_z clip(1 - r_s / chi(z), 0, 1000) for a
redshift distribution (Simpson weights w_z on linspace(min_z, max_z, n_integrate + 1)), or
clip(1 - r_s / chi(z_k), 0, 1000) for a scalar source redshift. The einsum of _born_spherical is
linear in the per-z convergence, so it folds into C. d_r = max_comoving_radius / n_planes, as the
reference takes it (born.py:36-37), whatever the shells' own widths. The maps are reduced on the
device by pmx_born; gradients flow to the maps (pmx_born_vjp) and to tensor-valued cosmological
parameters (float64 central differences of C). Flat lightcones (FlatDensity) go through
pmx_born_flat: each plane resampled by periodic bilinear interpolation on its chi-scaled angular
grid inside the same sum; their parameter gradients add the sampling positions' dependence on chi.
"""

from __future__ import annotations

import numpy as np
import torch

from . import autodiff, engine
from . import cosmology as _c
from .fields import ConvergenceUnit, FieldStatus, FlatKappaField, SphericalDensity, SphericalKappaField
from .redshift import redshift_distribution, simps_weights

H0_OVER_C = 100.0 / 299792.458  # h/Mpc: H0 = 100 km/s/(Mpc/h), c in km/s (jax_cosmo.constants)


def _normalize_sources(nz_shear):
    """("distribution", list) or ("redshift", float64 array) (_normalize_nz.py:12-38)."""
    entries = list(nz_shear) if isinstance(nz_shear, (list, tuple)) else [nz_shear]
    if not entries:
        raise ValueError("nz_shear must contain at least one entry")
    if isinstance(entries[0], redshift_distribution):
        if not all(isinstance(e, redshift_distribution) for e in entries):
            raise ValueError("Cannot mix redshift distributions with scalar sources")
        return "distribution", entries
    z = np.asarray(entries, dtype=np.float64).squeeze()
    z = np.atleast_1d(z)
    if z.ndim != 1:
        raise ValueError("Scalar redshift sources must be 1D array-like")
    return "redshift", z


def born_coefficients(cosmo, r_center, scale_factors, nz_shear, d_r, min_z=0.01, max_z=1.5, n_integrate=32):
    """C (n_sources, n_planes), float64."""
    kind, sources = _normalize_sources(nz_shear)
    r = np.asarray(r_center, dtype=np.float64).reshape(-1)
    a = np.asarray(scale_factors, dtype=np.float64).reshape(-1)
    base = 1.5 * cosmo.Omega_m * H0_OVER_C**2 * d_r * r / a
    if kind == "distribution":
        z = np.linspace(min_z, max_z, n_integrate + 1)
        chi_s = np.asarray(_c.radial_comoving_distance(cosmo, _c.z2a(z)), dtype=np.float64)
        eff = np.clip(1.0 - r[:, None] / chi_s[None, :], 0.0, 1000.0)  # (S, nz)
        nzw = np.stack([np.asarray(nz(z), dtype=np.float64) for nz in sources]) * simps_weights(min_z, max_z,
                                                                                                 n_integrate)
        return (nzw @ eff.T) * base[None, :]
    chi_s = np.atleast_1d(np.asarray(_c.radial_comoving_distance(cosmo, _c.z2a(sources)), dtype=np.float64))
    return np.clip(1.0 - r[None, :] / chi_s[:, None], 0.0, 1000.0) * base[None, :]


def _source_redshifts(kind, sources, min_z, max_z, n_integrate):
    """Effective source redshifts (_metadata.py:9-33): Simpson-weighted mean z of each n(z)."""
    if kind != "distribution":
        return np.atleast_1d(sources)
    z = np.linspace(min_z, max_z, n_integrate + 1)
    w = simps_weights(min_z, max_z, n_integrate)
    return np.array([float((w * nz(z) * z).sum() / (w * nz(z)).sum()) for nz in sources])


class _Born(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, spec, *params):
        coef_fn, cosmo, names = spec
        ctx.spec = spec
        coef = engine.to_device(coef_fn(cosmo), maps.device)
        kappa = engine.born(maps.contiguous(), coef)
        ctx.save_for_backward(maps)
        ctx.coef, ctx.params = coef, params
        return kappa

    @staticmethod
    def backward(ctx, kappa_bar):
        (maps,) = ctx.saved_tensors
        names = bool(ctx.params)
        maps_bar, dots = engine.born_vjp(maps, ctx.coef, kappa_bar.contiguous(), want_dots=names)
        grads = []
        if names:
            coef_fn, cosmo, pnames = ctx.spec
            # host derivatives of C, computed while the device runs the adjoint
            ders = [engine.to_device(autodiff.fd(coef_fn, cosmo, n), maps.device) for n in pnames]
            D, E, mean = dots
            g = D / mean[None, :] - E[:, None]  # d<kappa_bar, kappa>/d C
            for dC, t in zip(ders, ctx.params):
                grads.append(autodiff._param_grad(t, (dC * g).sum()))
        return (maps_bar if ctx.needs_input_grad[0] else None, None, *grads)


def flat_scales(lightcone, r, shape_yx):
    """(nshell, 2) float64: row and column sampling coordinate per output pixel index of each plane at
    comoving distance r (_born_core_impl, _born.py:53-98, with pixel_size from _born_flat, :154-159):
    x angles linspace(0, fx, nx) index rows over pixel_size[0] = box_y / ny, y angles linspace(0, fy, ny)
    index columns over pixel_size[1] = box_x / nx."""
    fsz = lightcone.field_size
    fx, fy = (tuple(float(f) for f in fsz) if isinstance(fsz, (tuple, list, np.ndarray)) else (float(fsz),) * 2)
    ny_m, nx_m = (int(n) for n in lightcone.flatsky_npix)
    py, px = lightcone.box_size[1] / ny_m, lightcone.box_size[0] / nx_m
    ny, nx = (int(n) for n in shape_yx)
    deg = np.pi / 180.0
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    return np.stack([(fx / nx) * deg * r / py, (fy / ny) * deg * r / px], axis=1)


class _BornFlat(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, spec, *params):
        coef_fn, scale_fn, chi_fn, cosmo, names = spec
        ctx.spec = spec
        coef = engine.to_device(coef_fn(cosmo), maps.device)
        scales = engine.to_device(scale_fn(cosmo), maps.device)
        kappa = engine.born_flat(maps.contiguous(), coef, scales)
        ctx.save_for_backward(maps)
        ctx.coef, ctx.scales, ctx.params = coef, scales, params
        return kappa

    @staticmethod
    def backward(ctx, kappa_bar):
        (maps,) = ctx.saved_tensors
        coef_fn, scale_fn, chi_fn, cosmo, pnames = ctx.spec
        want = bool(ctx.params)
        maps_bar, dots = engine.born_flat_vjp(maps, ctx.coef, ctx.scales, kappa_bar.contiguous(), want_dots=want)
        grads = []
        if want:
            # host derivatives of C and of the plane distances, computed while the device runs the adjoint
            chi = np.asarray(chi_fn(cosmo), dtype=np.float64)
            ders = [(engine.to_device(autodiff.fd(coef_fn, cosmo, n), maps.device),
                     engine.to_device(autodiff.fd(chi_fn, cosmo, n) / chi, maps.device)) for n in pnames]
            D, E, mean, P = dots
            g = D / mean[None, :] - E[:, None]  # d<kappa_bar, kappa>/d C
            for (dC, dchi), t in zip(ders, ctx.params):
                grads.append(autodiff._param_grad(t, (dC * g).sum() + (P * dchi).sum()))
        return (maps_bar if ctx.needs_input_grad[0] else None, None, *grads)


def born(cosmo, lightcone, nz_shear, min_z=0.01, max_z=1.5, n_integrate=32):
    """Born convergence maps (n_sources, npix) of a spherical lightcone (SphericalKappaField)."""
    if nz_shear is None:
        raise ValueError("nz_shear must be provided for lensing")
    if lightcone.status != FieldStatus.LIGHTCONE:
        raise ValueError(f"Expected lightcone with status=LIGHTCONE, got {lightcone.status}")
    a = np.atleast_1d(np.asarray(lightcone.scale_factors, dtype=np.float64))
    n_planes = a.size
    arr = lightcone.array
    if arr.ndim not in (2, 3):
        raise ValueError(f"Lightcone array must be 2D (spherical) or 3D (flat), got {arr.ndim}D")
    cosmo_f, names, tensors = autodiff.split_cosmo(cosmo)
    if not isinstance(lightcone, SphericalDensity):
        return _born_flat(cosmo, cosmo_f, names, tensors, lightcone, a, nz_shear, min_z, max_z, n_integrate)
    d_r = lightcone.max_comoving_radius / n_planes
    kind, sources = _normalize_sources(nz_shear)

    def coef_fn(c):
        r = _c.radial_comoving_distance(c, a)
        return born_coefficients(c, r, a, nz_shear, d_r, min_z, max_z, n_integrate)

    maps = arr.reshape(n_planes, -1)
    if maps.requires_grad or tensors:
        kappa = _Born.apply(maps, (coef_fn, cosmo_f, names), *tensors)
    else:
        kappa = engine.born(maps.contiguous(), torch.as_tensor(coef_fn(cosmo_f), dtype=torch.float64))
    zs = _source_redshifts(kind, sources, min_z, max_z, n_integrate)
    af = _c.z2a(zs)
    return SphericalKappaField.FromDensityMetadata(
        array=kappa, field=lightcone, status=FieldStatus.KAPPA, unit=ConvergenceUnit.DIMENSIONLESS, z_sources=zs,
        scale_factors=af, comoving_centers=np.asarray(_c.radial_comoving_distance(cosmo_f, af)),
        density_width=np.zeros_like(zs))


def _born_flat(cosmo, cosmo_f, names, tensors, lightcone, a, nz_shear, min_z, max_z, n_integrate):
    """Flat-sky branch of born (_born_flat, src/jax_fli/_src/lensing/_born.py:131-190)."""
    if lightcone.field_size is None:
        raise ValueError("field_size is required on lightcone for flat-sky convergence")
    if lightcone.flatsky_npix is None:
        raise ValueError("flatsky_npix must be set on lightcone for flat geometry")
    arr = lightcone.array
    if arr.ndim != 3:
        raise ValueError(f"flat lightcone array must be (n_planes, ny, nx), got {tuple(arr.shape)}")
    n_planes = a.size
    d_r = lightcone.max_comoving_radius / n_planes
    kind, sources = _normalize_sources(nz_shear)
    shape_yx = tuple(arr.shape[-2:])

    def chi_fn(c):
        return np.asarray(_c.radial_comoving_distance(c, a), dtype=np.float64).reshape(-1)

    def coef_fn(c):
        return born_coefficients(c, chi_fn(c), a, nz_shear, d_r, min_z, max_z, n_integrate)

    def scale_fn(c):
        return flat_scales(lightcone, chi_fn(c), shape_yx)

    if arr.requires_grad or tensors:
        kappa = _BornFlat.apply(arr, (coef_fn, scale_fn, chi_fn, cosmo_f, names), *tensors)
    else:
        kappa = engine.born_flat(arr.contiguous(), coef_fn(cosmo_f), scale_fn(cosmo_f))
    zs = _source_redshifts(kind, sources, min_z, max_z, n_integrate)
    af = _c.z2a(zs)
    return FlatKappaField.FromDensityMetadata(
        array=kappa, field=lightcone, status=FieldStatus.KAPPA, unit=ConvergenceUnit.DIMENSIONLESS, z_sources=zs,
        scale_factors=af, comoving_centers=np.asarray(_c.radial_comoving_distance(cosmo_f, af)),
        density_width=np.zeros_like(zs))
