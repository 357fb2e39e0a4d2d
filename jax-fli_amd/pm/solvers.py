"""N-body solvers (src/jax_fli/pm/solvers.py:19-411) on the pmx fused step.

The scalar kick/drift factors are computed on the host in float64 from the same float32 time
values the reference's traced arithmetic produces; the particle update runs as one
``pmx_step`` (paint -> FFT force mesh -> readout fused with kick and drift).

Each solver exposes the reference's ``init/step/save_at/reverse`` on ParticleFields (functional,
returns new fields) plus ``init_op/step_op``: the same updates as coefficient records, which
``integrate`` chains so that one step's drift is fused into the previous step's readout.
``ReversibleDoubleKickDrift.reverse`` is the exact algebraic inverse of its ``step`` (the
reference divides its un-drift by dGfa instead of gp, solvers.py:309 vs :364, so its reverse
does not invert its step).
"""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np

from .. import cosmology as _c
from .. import engine
from ..fields import ParticleField, PositionUnit
from .correction import AbstractCorrection, NoCorrection
from .interp import AbstractInterp, InterpTilerState

f32 = np.float32


def _sqrt32(x) -> np.float32:
    return f32(np.sqrt(f32(x)))


@dataclasses.dataclass(frozen=True)
class NBodyState:
    interp_state: InterpTilerState
    t_initial: float


@dataclasses.dataclass
class StepOp:
    """v += F(x) * force_scale * kick (if kick is not None), with optional drifts before/after and an
    optional correction (PGD displacement or sharpening boost) between the leading drift and the kick."""

    pre: float | None = None
    kick: float | None = None
    post: float | None = None
    corr: Any = None


def force_scale(cosmo) -> float:
    """_forces multiplies pm_forces by 1.5 * Omega_m (solvers.py:29-30)."""
    return 1.5 * cosmo.Omega_m


def _mode(field: ParticleField) -> str:
    if field.unit == PositionUnit.GRID_RELATIVE:
        return "relative"
    if field.unit == PositionUnit.GRID_ABSOLUTE:
        return "absolute"
    raise ValueError("PM forces need GRID_RELATIVE or GRID_ABSOLUTE particles")


def run_op(op: StepOp, x, v, mesh_size, mode, cosmo, deconv=False, force_mesh=None) -> None:
    """Apply one StepOp in place on device arrays x, v (force_mesh: keep the kick's force mesh there)."""
    if op.kick is None:
        for c in (op.pre, op.post):
            if c is not None:
                x.add_(v * f32(c))
        return
    pre = op.pre
    if op.corr is not None:  # the correction sees the drifted positions, the kick the corrected ones
        if pre is not None:
            x.add_(v * f32(pre))
            pre = None
        op.corr.run(x, v, mesh_size, mode)
    engine.step(x, v, mesh_size, mode=mode, force_scale=force_scale(cosmo), c_kick=op.kick, c_drift_pre=pre,
                c_drift_post=op.post, deconv=deconv, force_mesh=force_mesh)


@dataclasses.dataclass(frozen=True)
class AbstractNBodySolver:
    interp_kernel: AbstractInterp
    pgd_kernel: AbstractCorrection = NoCorrection()

    # coefficient records ------------------------------------------------------
    def init_op(self, t0, t1, dt, cosmo, raw: bool = False) -> StepOp | None:
        raise NotImplementedError

    def step_op(self, t0, t1, dt, t_initial, cosmo) -> StepOp:
        raise NotImplementedError

    def step_ops(self, grid, dt, t_initial, cosmo, raw: bool = False) -> list:
        """step_op for every (t0, t1) of grid (raw: float64 coefficients, for derivatives w.r.t. the
        cosmology; otherwise rounded to float32 as the reference's traced arithmetic gives them)."""
        return [self.step_op(a, b, dt, t_initial, cosmo) for a, b in grid]

    # reference interface --------------------------------------------------------
    def _apply(self, op, displacement, velocities, cosmo):
        x = displacement.array.clone()
        v = velocities.array.clone()
        if op is not None:
            run_op(op, x, v, displacement.mesh_size, _mode(displacement), cosmo)
        return x, v

    def save_at(self, displacement, velocities, t1, dt, state: NBodyState, cosmo) -> Any:
        return self.interp_kernel.paint(state.interp_state, t1, (displacement, velocities), cosmo)

    def reverse(self, displacement, velocities, t0, t1, dt, state, cosmo):
        raise NotImplementedError(f"{type(self).__name__} is not reversible.")


@dataclasses.dataclass(frozen=True)
class EfficientDriftDoubleKick(AbstractNBodySolver):
    """FastPM KDK (solvers.py:118-225): half kick at init, then drift -> force -> double kick."""

    def init_op(self, t0, t1, dt, cosmo, raw: bool = False) -> StepOp:
        # solvers.py:136-147: ai = ac = t0, af = sqrt(t0 t1) (host floats), kick with F(x0)
        ai = ac = f32(t0)
        af = f32((float(t0) * float(t1)) ** 0.5)
        kick = (_c.Gf(cosmo, af) - _c.Gf(cosmo, ai)) / _c.dGfa(cosmo, ac)
        k = kick / (ac**2 * _c.E(cosmo, ac))
        return StepOp(kick=float(k) if raw else float(f32(k)))

    def step_op(self, t0, t1, dt, t_initial, cosmo) -> StepOp:
        t0, t1 = f32(t0), f32(t1)
        t0t1 = _sqrt32(t0 * t1)
        # drift (solvers.py:165-172)
        drift = (_c.Gp(cosmo, t1) - _c.Gp(cosmo, t0)) / _c.gp(cosmo, t0t1)
        cd = drift / (float(t0t1) ** 3 * _c.E(cosmo, t0t1))
        # double kick at t1 with the nominal next half step (solvers.py:178-193)
        t2 = f32(t1 + f32(dt))
        t1t2 = _sqrt32(t1 * t2)
        k1 = (_c.Gf(cosmo, t1) - _c.Gf(cosmo, t0t1)) / _c.dGfa(cosmo, t1)
        k2 = (_c.Gf(cosmo, t1t2) - _c.Gf(cosmo, t1)) / _c.dGfa(cosmo, t1)
        ck = (k1 + k2) / (float(t1) ** 2 * _c.E(cosmo, t1))
        # solvers.py:175: the correction gets the drift factor without its 1/(a^3 E) prefactor
        return StepOp(pre=float(f32(cd)), kick=float(f32(ck)), corr=self.pgd_kernel.op(float(f32(drift))))

    def step_ops(self, grid, dt, t_initial, cosmo, raw: bool = False) -> list:
        """step_op vectorised over the steps (same float32 time arithmetic and float64 factors)."""
        if len(grid) == 0:
            return []
        t0 = np.asarray([g[0] for g in grid], dtype=np.float32)
        t1 = np.asarray([g[1] for g in grid], dtype=np.float32)
        t0t1 = np.sqrt((t0 * t1).astype(np.float32)).astype(np.float32)
        drift = (_c.Gp(cosmo, t1) - _c.Gp(cosmo, t0)) / _c.gp(cosmo, t0t1)
        cd = drift / (t0t1.astype(np.float64) ** 3 * _c.E(cosmo, t0t1))
        t2 = (t1 + f32(dt)).astype(np.float32)
        t1t2 = np.sqrt((t1 * t2).astype(np.float32)).astype(np.float32)
        dgf = _c.dGfa(cosmo, t1)
        gf1 = _c.Gf(cosmo, t1)
        k1 = (gf1 - _c.Gf(cosmo, t0t1)) / dgf
        k2 = (_c.Gf(cosmo, t1t2) - gf1) / dgf
        ck = (k1 + k2) / (t1.astype(np.float64) ** 2 * _c.E(cosmo, t1))
        r = (lambda q: float(q)) if raw else (lambda q: float(f32(q)))
        return [StepOp(pre=r(a), kick=r(b), corr=self.pgd_kernel.op(r(d))) for a, b, d in zip(cd, ck, drift)]

    def init(self, displacement, velocities, t0, t1, dt, cosmo):
        op = self.init_op(t0, t1, dt, cosmo)
        x, v = self._apply(op, displacement, velocities, cosmo)
        af = float(f32((float(t0) * float(t1)) ** 0.5))
        return (displacement.replace(array=x), velocities.replace(array=v, scale_factors=af),
                NBodyState(interp_state=self.interp_kernel.init(), t_initial=float(t0)))

    def step(self, displacement, velocities, t0, t1, dt, state, cosmo):
        op = self.step_op(t0, t1, dt, state.t_initial, cosmo)
        x, v = self._apply(op, displacement, velocities, cosmo)
        t1 = float(f32(t1))
        new_state = NBodyState(interp_state=self.interp_kernel.advance(state.interp_state, t1),
                               t_initial=state.t_initial)
        return (displacement.replace(array=x, scale_factors=t1), velocities.replace(array=v, scale_factors=t1),
                new_state)


@dataclasses.dataclass(frozen=True)
class ReversibleDoubleKickDrift(AbstractNBodySolver):
    """Reversible KKD (solvers.py:228-411): force at x_i -> K1 + K2 -> drift; init is a no-op."""

    def init_op(self, t0, t1, dt, cosmo, raw: bool = False):
        return None

    def step_op(self, t0, t1, dt, t_initial, cosmo, raw: bool = False) -> StepOp:
        """raw: float64 coefficients (for derivatives w.r.t. the cosmology), else rounded to float32
        as the reference's traced float32 arithmetic produces them."""
        t0, t1 = f32(t0), f32(t1)
        t_prev = max(f32(t0 - f32(dt)), f32(t_initial))  # solvers.py:276-277
        ai_1 = _sqrt32(t_prev * t0)
        af_2 = _sqrt32(t0 * t1)
        k1 = (_c.Gf(cosmo, t0) - _c.Gf(cosmo, ai_1)) / _c.dGfa(cosmo, t0)
        k2 = (_c.Gf(cosmo, af_2) - _c.Gf(cosmo, t0)) / _c.dGfa(cosmo, t0)
        ck = (k1 + k2) / (float(t0) ** 2 * _c.E(cosmo, t0))
        drift = (_c.Gp(cosmo, t1) - _c.Gp(cosmo, t0)) / _c.gp(cosmo, af_2)
        cd = drift / (float(af_2) ** 3 * _c.E(cosmo, af_2))
        if raw:
            return StepOp(kick=float(ck), post=float(cd))
        # solvers.py:311-316: the boost is S(x_i) / full_drift_factor, applied before the drift
        return StepOp(kick=float(f32(ck)), post=float(f32(cd)), corr=self.pgd_kernel.op(float(f32(cd))))

    def step_coeffs(self, grid, dt, t_initial, cosmo, raw: bool = False) -> np.ndarray:
        """(nsteps, 2) array of (kick, drift) for the steps [(t0, t1), ...]: step_op vectorised over the
        steps (the same float32 time arithmetic and float64 factors, element by element)."""
        if len(grid) == 0:
            return np.zeros((0, 2))
        t0 = np.asarray([g[0] for g in grid], dtype=np.float32)
        t1 = np.asarray([g[1] for g in grid], dtype=np.float32)
        t_prev = np.maximum((t0 - f32(dt)).astype(np.float32), f32(t_initial))
        ai_1 = np.sqrt((t_prev * t0).astype(np.float32)).astype(np.float32)
        af_2 = np.sqrt((t0 * t1).astype(np.float32)).astype(np.float32)
        dgf = _c.dGfa(cosmo, t0)
        gf_t0 = _c.Gf(cosmo, t0)
        k1 = (gf_t0 - _c.Gf(cosmo, ai_1)) / dgf
        k2 = (_c.Gf(cosmo, af_2) - gf_t0) / dgf
        ck = (k1 + k2) / (t0.astype(np.float64) ** 2 * _c.E(cosmo, t0))
        drift = (_c.Gp(cosmo, t1) - _c.Gp(cosmo, t0)) / _c.gp(cosmo, af_2)
        cd = drift / (af_2.astype(np.float64) ** 3 * _c.E(cosmo, af_2))
        out = np.stack([ck, cd], axis=-1).astype(np.float64)
        return out if raw else out.astype(np.float32).astype(np.float64)

    def step_ops(self, grid, dt, t_initial, cosmo, raw: bool = False) -> list:
        return [StepOp(kick=float(a), post=float(b), corr=self.pgd_kernel.op(float(b)))
                for a, b in self.step_coeffs(grid, dt, t_initial, cosmo, raw=raw)]

    def reverse_op(self, t0, t1, dt, t_initial, cosmo) -> StepOp:
        """Exact inverse of step_op(t0, t1): x -= v cd, then the boost is rewound at that x
        (solvers.py:372-376), then v -= F(x) ck."""
        op = self.step_op(t0, t1, dt, t_initial, cosmo)
        return StepOp(pre=-op.post, kick=-op.kick, corr=self.pgd_kernel.rewind_op(op.post))

    def init(self, displacement, velocities, t0, t1, dt, cosmo):
        return displacement, velocities, NBodyState(interp_state=self.interp_kernel.init(), t_initial=float(t0))

    def step(self, displacement, velocities, t0, t1, dt, state, cosmo):
        op = self.step_op(t0, t1, dt, state.t_initial, cosmo)
        x, v = self._apply(op, displacement, velocities, cosmo)
        t1f = float(f32(t1))
        af2 = float(_sqrt32(f32(t0) * f32(t1)))
        new_state = NBodyState(interp_state=self.interp_kernel.advance(state.interp_state, t1f),
                               t_initial=state.t_initial)
        return (displacement.replace(array=x, scale_factors=t1f), velocities.replace(array=v, scale_factors=af2),
                new_state)

    def reverse(self, displacement, velocities, t0, t1, dt, state, cosmo):
        op = self.reverse_op(t0, t1, dt, state.t_initial, cosmo)
        x, v = self._apply(op, displacement, velocities, cosmo)
        interp = self.interp_kernel
        tmax = float(np.asarray(interp.ts).reshape(-1)[-1]) if interp.ts is not None else float(t1)
        prev = NBodyState(interp_state=interp.rewind(state.interp_state, min(float(f32(t1)), tmax)),
                          t_initial=state.t_initial)
        return displacement.replace(array=x, scale_factors=float(f32(t0))), velocities.replace(array=v), prev
