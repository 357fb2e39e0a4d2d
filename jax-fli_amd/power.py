"""3D power spectra on the GPU: the estimator of ``_power`` (src/jax_fli/_src/power/_compute.py:5-119) --
auto and cross spectra, Legendre multipoles about a line of sight -- and the monopole ``transfer`` and
``coherence`` built on it (_compute.py:240-252; public API src/jax_fli/power/compute.py:12-66). The FFT
and the |k| / mu binning run in the pmx x-pass kernel with float64 accumulators
(pmx_power_multipoles). Returns (k, P) as numpy arrays, P of shape (nbins,) for one pole and
(npoles, nbins) for several, like ``jax_fli.power``."""

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


def _as_mesh(mesh, device=None):
    if not isinstance(mesh, torch.Tensor):
        mesh = torch.as_tensor(np.asarray(mesh, dtype=np.float32), device=device or "cuda")
    return mesh.to(torch.float32).contiguous()


def _kedges(kedges, shape, box_shape) -> np.ndarray:
    if kedges is None:
        kedges = default_kedges(shape, box_shape)
    elif np.ndim(kedges) == 0:
        n = int(kedges)
        kmax = np.pi * np.min(np.asarray(shape) / np.asarray(box_shape))
        kedges = np.linspace(0.0, kmax, n + 1)
    return np.asarray(kedges, dtype=np.float64)


def power(mesh, mesh2=None, *, box_shape, kedges=None, multipoles=0, los=(0.0, 0.0, 1.0)):
    """(k, P): auto spectrum of `mesh` or cross spectrum with `mesh2`, for the pole(s) `multipoles`
    about the line of sight `los` (ignored when multipoles == 0). A cross spectrum's bin sum is the
    modulus of the complex sum of m1 conj(m2), as in the reference."""
    mesh = _as_mesh(mesh)
    if mesh2 is not None:
        mesh2 = _as_mesh(mesh2, mesh.device)
        if tuple(mesh2.shape) != tuple(mesh.shape):
            raise ValueError("mesh2 must have the mesh's shape")
    shape = tuple(mesh.shape)
    poles = tuple(int(l) for l in multipoles) if isinstance(multipoles, (list, tuple)) else (int(multipoles),)
    use_los = not (not isinstance(multipoles, (list, tuple)) and int(multipoles) == 0)
    kedges = _kedges(kedges, shape, box_shape)
    re, im, ks, kc = engine.power_multipole_sums(mesh, mesh2, tuple(float(b) for b in box_shape),
                                                 torch.from_numpy(kedges), poles,
                                                 tuple(float(v) for v in los) if use_los else None)
    re, ks, kc = re.cpu().numpy(), ks.cpu().numpy(), kc.cpu().numpy()
    psum = re if im is None else np.sqrt(re**2 + im.cpu().numpy() ** 2)
    ncell = float(np.prod(shape))
    kavg = (ks / np.where(kc == 0, 1, kc))[1:-1]
    norm = np.where(kc > 0, kc, 1)
    spectra = (psum / ncell / norm)[:, 1:-1] * np.prod(np.asarray(box_shape, dtype=np.float64) / np.asarray(shape))
    return (kavg, spectra[0]) if len(poles) == 1 else (kavg, spectra)


def transfer(mesh0, mesh1, *, box_shape, kedges=None):
    """(k, sqrt(P1 / P0)), monopoles (src/jax_fli/_src/power/_compute.py:240-244)."""
    k, p0 = power(mesh0, box_shape=box_shape, kedges=kedges)
    _, p1 = power(mesh1, box_shape=box_shape, kedges=kedges)
    return k, (p1 / p0) ** 0.5


def coherence(mesh0, mesh1, *, box_shape, kedges=None):
    """(k, P01 / sqrt(P0 P1)), monopoles (src/jax_fli/_src/power/_compute.py:247-252)."""
    k, p01 = power(mesh0, mesh1, box_shape=box_shape, kedges=kedges)
    _, p0 = power(mesh0, box_shape=box_shape, kedges=kedges)
    _, p1 = power(mesh1, box_shape=box_shape, kedges=kedges)
    return k, p01 / (p0 * p1) ** 0.5
