"""Snapshot interpolation kernels (src/jax_fli/pm/interp.py:46-209): the abstract interface and
``NoInterp`` (paint the particles as they are at each snapshot). Drift/Onion/Telephoto tilers are
not implemented on this engine."""

from __future__ import annotations

import dataclasses
import warnings
from typing import Any

import numpy as np

from ..painting import PaintingOptions


@dataclasses.dataclass(frozen=True)
class InterpTilerState:
    rotation_idx: int = -1
    shell_idx: int = -1
    rotations: Any = None


@dataclasses.dataclass(frozen=True)
class AbstractInterp:
    painting: PaintingOptions
    ts: Any = None
    r_centers: Any = None
    density_widths: Any = None
    max_comoving_distance: float | None = None

    def update_geometry(self, ts, r_centers, density_widths, max_comoving_distance):
        return dataclasses.replace(self, ts=ts, r_centers=r_centers, density_widths=density_widths,
                                   max_comoving_distance=max_comoving_distance)

    def init(self) -> InterpTilerState:
        raise NotImplementedError

    def advance(self, state: InterpTilerState, t) -> InterpTilerState:
        raise NotImplementedError

    def rewind(self, state: InterpTilerState, t) -> InterpTilerState:
        raise NotImplementedError

    def paint(self, state: InterpTilerState, t, y, cosmo):
        raise NotImplementedError


@dataclasses.dataclass(frozen=True)
class NoInterp(AbstractInterp):
    def init(self) -> InterpTilerState:
        if self.r_centers is not None and self.density_widths is not None and self.max_comoving_distance is not None:
            far = float(np.asarray(self.r_centers).reshape(-1)[0] + np.asarray(self.density_widths).reshape(-1)[0] / 2)
            if far > self.max_comoving_distance:
                warnings.warn(
                    f"NoInterp does not support tiling. Your furthest shell is at {far:.2f} Mpc/h, but your box only "
                    f"extends to {self.max_comoving_distance:.2f} Mpc/h. Your simulations will run but will have "
                    f"artifacts for the far shells.")
        return InterpTilerState(rotation_idx=-1, shell_idx=-1, rotations=None)

    def advance(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx + 1) if hit else state

    def rewind(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx - 1) if hit else state

    def paint(self, state, t, y, cosmo):
        dx, p = y
        r_center = float(np.asarray(self.r_centers).reshape(-1)[state.shell_idx])
        width = float(np.asarray(self.density_widths).reshape(-1)[state.shell_idx])
        tgt = None if self.painting is None else self.painting.target
        if tgt is None or tgt == "particles":
            result = dx.replace(array=dx.array.clone())
        elif tgt == "spherical":
            pt = self.painting
            result = dx.paint_spherical(center=r_center, density_plane_width=width, scheme=pt.scheme,
                                        weights=pt.weights, kernel_width_arcmin=pt.kernel_width_arcmin,
                                        smoothing_interpretation=pt.smoothing_interpretation,
                                        paint_nside=pt.paint_nside, batch_size=pt.batch_size)
        elif tgt == "flat":
            result = dx.paint_2d(center=r_center, density_plane_width=width, weights=self.painting.weights,
                                 batch_size=self.painting.batch_size)
        elif tgt == "density":
            result = dx.paint(weights=self.painting.weights, chunk_size=self.painting.chunk_size,
                              batch_size=self.painting.batch_size)
        else:
            result = dx
        return result.replace(scale_factors=float(t), comoving_centers=r_center, density_width=width)


class DriftInterp(AbstractInterp):
    def __init__(self, *a, **k):
        raise NotImplementedError("DriftInterp is not implemented on the pmx engine")


class OnionTiler(AbstractInterp):
    def __init__(self, *a, **k):
        raise NotImplementedError("OnionTiler is not implemented on the pmx engine")


class TelephotoInterp(AbstractInterp):
    def __init__(self, *a, **k):
        raise NotImplementedError("TelephotoInterp is not implemented on the pmx engine")
