"""Correction kernels (src/jax_fli/pm/correction.py:25-333): NoCorrection, the PGD kernel (a
displacement correction after the drift, not reversible) and the Sharpening kernel (the same
displacement as a velocity boost before the drift, reversible).

Both corrections evaluate S(x) = alpha * F_pgd(x): the PM force of paint(x) with the Green's function
multiplied by exp(-kl^2/k^2) exp(-k^4/ks^4) (k in cell units, 0 at k = 0), through ``pmx_pgd`` (the
same paint, FFT passes, gradient stencil and readout as the force, with the filter applied in the
k-space pass). ``op(full_drift_factor)`` gives the record the fused integration loop applies between
a step's leading drift and its kick.
"""

from __future__ import annotations

import dataclasses

from .. import engine
from ..fields import PositionUnit


def _mode(field) -> str:
    if field.unit == PositionUnit.GRID_RELATIVE:
        return "relative"
    if field.unit == PositionUnit.GRID_ABSOLUTE:
        return "absolute"
    raise ValueError("PM corrections need GRID_RELATIVE or GRID_ABSOLUTE particles")


@dataclasses.dataclass(frozen=True)
class CorrectionOp:
    """x += S(x) (target "x") or v += S(x) * v_scale (target "v")."""

    target: str
    alpha: float
    kl: float
    ks: float
    v_scale: float = 1.0

    def run(self, x, v, mesh_size, mode) -> None:
        engine.pgd(x, v, mesh_size, mode=mode, alpha=self.alpha, kl=self.kl, ks=self.ks, target=self.target,
                   v_scale=self.v_scale)


@dataclasses.dataclass(frozen=True)
class AbstractCorrection:
    @property
    def reversible(self) -> bool:
        return True

    def apply(self, t, pos, vel, full_drift_factor=1.0):
        raise NotImplementedError

    def rewind(self, t, pos, vel, full_drift_factor=1.0):
        raise NotImplementedError

    def op(self, full_drift_factor):
        """The correction as a CorrectionOp (None: nothing to do)."""
        raise NotImplementedError

    def rewind_op(self, full_drift_factor):
        raise NotImplementedError


@dataclasses.dataclass(frozen=True)
class NoCorrection(AbstractCorrection):
    """Identity (correction.py:95-128)."""

    def apply(self, t, pos, vel, full_drift_factor=1.0):
        return pos, vel

    def rewind(self, t, pos, vel, full_drift_factor=1.0):
        return pos, vel

    def op(self, full_drift_factor):
        return None

    def rewind_op(self, full_drift_factor):
        return None


def _run(op: CorrectionOp, pos, vel):
    x = pos.array.contiguous().clone()
    v = vel.array.contiguous().clone() if op.target == "v" else vel.array
    op.run(x, v, pos.mesh_size, _mode(pos))
    return pos.replace(array=x), (vel.replace(array=v) if op.target == "v" else vel)


@dataclasses.dataclass(frozen=True)
class PGDKernel(AbstractCorrection):
    """x += alpha F_pgd(x) after the drift (correction.py:131-228); not reversible."""

    alpha: float = 0.2
    kl: float = 0.3
    ks: float = 1.0

    @property
    def reversible(self) -> bool:
        return False

    def op(self, full_drift_factor):
        return CorrectionOp("x", float(self.alpha), float(self.kl), float(self.ks))

    def rewind_op(self, full_drift_factor):
        raise NotImplementedError("PGDKernel is not reversible.")

    def apply(self, t, pos, vel, full_drift_factor=1.0):
        return _run(self.op(full_drift_factor), pos, vel)

    def rewind(self, t, pos, vel, full_drift_factor=1.0):
        raise NotImplementedError("PGDKernel is not reversible.")


@dataclasses.dataclass(frozen=True)
class SharpeningKernel(AbstractCorrection):
    """v += alpha F_pgd(x) / full_drift_factor before the drift (correction.py:231-333); the rewind
    recomputes S at the recovered position and subtracts it."""

    alpha: float = 0.2
    kl: float = 0.3
    ks: float = 1.0

    def op(self, full_drift_factor):
        return CorrectionOp("v", float(self.alpha), float(self.kl), float(self.ks), 1.0 / float(full_drift_factor))

    def rewind_op(self, full_drift_factor):
        return CorrectionOp("v", float(self.alpha), float(self.kl), float(self.ks), -1.0 / float(full_drift_factor))

    def apply(self, t, pos, vel, full_drift_factor=1.0):
        return _run(self.op(full_drift_factor), pos, vel)

    def rewind(self, t, pos, vel, full_drift_factor=1.0):
        return _run(self.rewind_op(full_drift_factor), pos, vel)
