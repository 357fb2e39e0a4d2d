"""3D power spectrum on the GPU: the estimator of ``_power`` (src/jax_fli/_src/power/_compute.py:23-119),
auto spectrum monopole, with the FFT and the |k| binning done by the pmx x-pass kernel in float64
accumulators. Returns (k, P(k)) as numpy arrays, like ``jax_fli.power``."""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np
import torch

from . import engine


@dataclasses.dataclass(frozen=True)
class PowerSpectrum:
    wavenumber: Any
    spectra: Any
    name: str = "pk"
    scale_factors: Any = None


def default_kedges(mesh_shape, box_shape) -> np.ndarray:
    ms = np.asarray(mesh_shape, dtype=np.float64)
    bs = np.asarray(box_shape, dtype=np.float64)
    kmax = np.pi * np.min(ms / bs)
    dk = 2 * np.pi / np.min(bs) * 2
    kedges = np.arange(dk, kmax, dk) + dk / 2
    if kedges.size < 2:
        kedges = np.linspace(kmax / 4, kmax * 0.9, 2)
    return kedges


def power(mesh, mesh2=None, *, box_shape, kedges=None, multipoles=0, los=(0.0, 0.0, 1.0)):
    if mesh2 is not None or multipoles not in (0, (0,), [0]):
        raise NotImplementedError("only the auto-spectrum monopole is implemented on the GPU path")
    if not isinstance(mesh, torch.Tensor):
        mesh = torch.as_tensor(np.asarray(mesh, dtype=np.float32), device="cuda")
    mesh = mesh.to(torch.float32).contiguous()
    shape = tuple(mesh.shape)
    if kedges is None:
        kedges = default_kedges(shape, box_shape)
    elif np.ndim(kedges) == 0:
        n = int(kedges)
        kmax = np.pi * np.min(np.asarray(shape) / np.asarray(box_shape))
        kedges = np.linspace(0.0, kmax, n + 1)
    kedges = np.asarray(kedges, dtype=np.float64)
    pk, ks, kc = engine.power_sums(mesh, tuple(float(b) for b in box_shape), torch.from_numpy(kedges))
    pk, ks, kc = pk.cpu().numpy(), ks.cpu().numpy(), kc.cpu().numpy()
    ncell = float(np.prod(shape))
    kavg = (ks / np.where(kc == 0, 1, kc))[1:-1]
    norm = np.where(kc > 0, kc, 1)
    spectra = (pk / ncell / norm)[1:-1] * np.prod(np.asarray(box_shape, dtype=np.float64) / np.asarray(shape))
    return kavg, spectra
