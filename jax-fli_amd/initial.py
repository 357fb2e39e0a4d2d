"""Gaussian initial conditions: white noise coloured by the linear power spectrum
(src/jax_fli/initial.py:33-199). Noise is Philox4x32-10 on the GPU (counter = cell index, key =
seed) in place of ``normal_field(jax.random key)``; the colouring is one r2c/c2r FFT round trip
with the sqrt(P(|k|) N^3 / V) multiply inside the pmx x pass."""

from __future__ import annotations

import numpy as np
import torch

from . import autodiff
from . import cosmology as _c
from . import engine
from .fields import DensityField, DensityUnit, FieldStatus


def _pk_table(cosmo, pk_fn, mesh_size, box_size):
    if pk_fn is None:
        if cosmo is None:
            raise ValueError("Either pk_fn or cosmo must be provided to compute the power spectrum.")
        k = np.logspace(-4, 1, 128)  # initial.py:171-173
        return k, _c.linear_matter_power(cosmo, k)
    kmax = np.sqrt(sum((np.pi * n / b) ** 2 for n, b in zip(mesh_size, box_size))) * 1.01
    kmin = 2 * np.pi / max(box_size) * 0.99
    k = np.concatenate([[0.0], np.geomspace(kmin, kmax, 4096)])
    return k, np.asarray(pk_fn(k), dtype=np.float64)


def interpolate_initial_conditions(initial_field, mesh_size, box_size, *, cosmo=None, pk_fn=None,
                                   observer_position=(0.5, 0.5, 0.5), flatsky_npix=None, nside=None,
                                   field_size=None, halo_size=(0, 0), sharding=None) -> DensityField:
    mesh_size = tuple(int(n) for n in mesh_size)
    box_size = tuple(float(b) for b in box_size)
    field = initial_field.array if isinstance(initial_field, DensityField) else initial_field
    if not isinstance(field, torch.Tensor):
        field = torch.as_tensor(np.asarray(field, dtype=np.float32), device="cuda")
    factor = np.prod(mesh_size) / np.prod(box_size)

    def table(c):
        k, pk = _pk_table(c, pk_fn, mesh_size, box_size)
        return k, pk * factor

    def colour_fn(noise, c):
        k, pt = table(c)
        return engine.color_field(noise.to(torch.float32).contiguous(), box_size, torch.from_numpy(k),
                                  torch.from_numpy(pt))

    # differentiable in the noise and in tensor-valued cosmological parameters (white noise -> IC)
    out = autodiff.colour(field, cosmo, mesh_size, box_size, colour_fn, table)
    return DensityField(array=out, mesh_size=mesh_size, box_size=box_size, observer_position=observer_position,
                        sharding=sharding, halo_size=tuple(halo_size) if not np.isscalar(halo_size) else (halo_size,) * 2,
                        nside=nside, field_size=field_size, flatsky_npix=flatsky_npix, scale_factors=0.0,
                        status=FieldStatus.INITIAL_FIELD, unit=DensityUnit.DENSITY)


def gaussian_initial_conditions(key, mesh_size, box_size, *, cosmo=None, pk_fn=None,
                                observer_position=(0.5, 0.5, 0.5), flatsky_npix=None, nside=None, field_size=None,
                                halo_size=(0, 0), sharding=None) -> DensityField:
    """`key` is an integer seed (Philox key)."""
    if not isinstance(key, (int, np.integer)):
        raise ValueError("key must be an integer seed")
    noise = engine.white_noise(tuple(int(n) for n in mesh_size), int(key))
    return interpolate_initial_conditions(noise, mesh_size, box_size, cosmo=cosmo, pk_fn=pk_fn,
                                          observer_position=observer_position, flatsky_npix=flatsky_npix, nside=nside,
                                          field_size=field_size, halo_size=halo_size, sharding=sharding)
