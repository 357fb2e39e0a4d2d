"""Correction kernels (src/jax_fli/pm/correction.py:25-128). The PM path uses NoCorrection; the PGD
and Sharpening kernels are not implemented on this engine and raise."""

from __future__ import annotations

import dataclasses


@dataclasses.dataclass(frozen=True)
class AbstractCorrection:
    reversible = True

    def apply(self, t, displacement, velocities, full_drift_factor=1.0):
        raise NotImplementedError

    def rewind(self, t, displacement, velocities, full_drift_factor=1.0):
        raise NotImplementedError


@dataclasses.dataclass(frozen=True)
class NoCorrection(AbstractCorrection):
    """Identity (correction.py:97-128)."""

    reversible = True

    def apply(self, t, displacement, velocities, full_drift_factor=1.0):
        return displacement, velocities

    def rewind(self, t, displacement, velocities, full_drift_factor=1.0):
        return displacement, velocities


class PGDKernel(AbstractCorrection):
    def __init__(self, *a, **k):
        raise NotImplementedError("PGDKernel is not implemented on the pmx engine")


class SharpeningKernel(AbstractCorrection):
    def __init__(self, *a, **k):
        raise NotImplementedError("SharpeningKernel is not implemented on the pmx engine")
