"""Unfused per-op GPU form of one PM step, the baseline the fused step is measured against.

It mirrors the op structure of the reference force ``jaxpm.pm.pm_forces`` as called by ``_forces``
(src/jax_fli/pm/solvers.py:19-32): a separate CIC paint by per-corner global atomic adds (the form an
XLA scatter-add takes on a GPU), a complex-to-complex forward 3D FFT of the
density (rocFFT through ``torch.fft.fftn``), a separate k-space multiply per force component
(Green's function -1/k^2 times the FastPM 4-point gradient), three complex-to-complex inverse FFTs,
three separate single-component CIC readouts, then separate kick and drift updates
(``EfficientDriftDoubleKick.step``, src/jax_fli/pm/solvers.py:165-195). Every intermediate is
materialised in HBM. The atomic paint and the single-component readout are this engine's kernels.

No JAX-on-ROCm exists on the target machines, so this form stands in for the reference's plain per-op
JAX-on-GPU path.
"""

from __future__ import annotations

import math

import torch

from .. import _lib, engine


class UnfusedPM:
    """Per-op PM force and kick/drift on a fixed mesh (cell units, k = 2 pi fftfreq(n))."""

    def __init__(self, mesh_shape, device):
        self.mesh_shape = tuple(int(n) for n in mesh_shape)
        self.plan = _lib.Plan(self.mesh_shape, device, flags=_lib.PAINT_ATOMIC | _lib.PLAN_NO_RESERVE)
        ks = [2.0 * math.pi * torch.fft.fftfreq(n, d=1.0, device=device, dtype=torch.float64) for n in self.mesh_shape]
        kx, ky, kz = torch.meshgrid(*ks, indexing="ij")
        k2 = kx**2 + ky**2 + kz**2
        green = torch.where(k2 > 0, -1.0 / torch.where(k2 > 0, k2, torch.ones_like(k2)), torch.zeros_like(k2))
        # -gradient_i * invlaplace; gradient_i = i (8 sin k - sin 2k) / 6 (zero at the Nyquist bin)
        self.kernels = []
        for k in (kx, ky, kz):
            g = (8.0 * torch.sin(k) - torch.sin(2.0 * k)) / 6.0
            self.kernels.append((-(1j) * g * green).to(torch.complex64))

    def paint(self, x: torch.Tensor) -> torch.Tensor:
        rho = torch.empty(self.mesh_shape, dtype=torch.float32, device=x.device)
        npx, npy, npz = x.shape[:3]
        _lib.check(_lib.load().pmx_paint_cic(self.plan.handle, _lib.ptr(x), npx, npy, npz, 0, None, 1.0,
                                             _lib.ptr(rho), 0, _lib.stream_ptr()), "pmx_paint_cic")
        return rho

    def forces(self, x: torch.Tensor, scale: float) -> torch.Tensor:
        rho = self.paint(x)
        delta_k = torch.fft.fftn(rho.to(torch.complex64))
        comps = []
        for kern in self.kernels:
            f_k = delta_k * kern
            f = torch.fft.ifftn(f_k).real.contiguous()
            comps.append(engine.readout(f, x))
        return torch.stack(comps, dim=-1) * scale

    def step(self, x: torch.Tensor, v: torch.Tensor, force_scale: float, c_kick: float, c_drift_post: float,
             c_drift_pre: float | None = None):
        if c_drift_pre is not None:
            x.add_(v, alpha=c_drift_pre)
        f = self.forces(x, force_scale)
        v.add_(f, alpha=c_kick)
        x.add_(v, alpha=c_drift_post)
        return x, v
