"""Time integration (src/jax_fli/pm/integrate.py:17-279).

``time_grid`` reproduces the reference's loop structure exactly, in float32: an outer scan over the
snapshot times ``ts``; inside, steps of ``dt0`` from the current time, the last one clipped to the
snapshot (tolerance 1e-6, ``_clip_to_end``), so the grid restarts at every snapshot and the run stops
at max(ts) (integrate.py:230-276). ``integrate`` then runs the solver's StepOps on device buffers,
folding each step's leading drift into the previous step's readout kernel whenever no snapshot
separates them, and calls ``save_at`` at each snapshot.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Any

import numpy as np
import torch

from .. import autodiff, engine
from ..fields import flat_slab
from .interp import image_snapshot_vjp, paint_drifted_vjp, paint_nside_of, ud_grade_adjoint
from .solvers import AbstractNBodySolver, NBodyState, ReversibleDoubleKickDrift, StepOp, _mode, force_scale, run_op

f32 = np.float32


def time_grid(t0: float, ts, dt0: float, t1: float):
    """Events [('step', t_curr, t_next) | ('save', index, t_target)] in float32 arithmetic."""
    ts = np.atleast_1d(np.asarray(ts, dtype=np.float32))
    max_steps = int(math.ceil(t1 / dt0)) + 5
    dt = f32(dt0)
    tol = f32(1e-6)
    t = f32(t0)
    events = []
    for idx, tt in enumerate(ts):
        tt = f32(tt)
        n = 0
        while t < tt and n < max_steps:
            tn = f32(t + dt)
            if tn > f32(tt - tol):
                tn = tt
            events.append(("step", float(t), float(tn)))
            t = tn
            n += 1
        events.append(("save", idx, float(tt)))
    return events


def count_steps(t0, ts, dt0, t1) -> int:
    return sum(1 for e in time_grid(t0, ts, dt0, t1) if e[0] == "step")


def _stack(snaps):
    first = snaps[0]
    arr = torch.stack([s.array for s in snaps])
    sf = np.asarray([s.scale_factors for s in snaps], dtype=np.float64)
    rc = np.asarray([s.comoving_centers for s in snaps], dtype=np.float64)
    dw = np.asarray([s.density_width for s in snaps], dtype=np.float64)
    return first.replace(array=arr, scale_factors=sf, comoving_centers=rc, density_width=dw)


# HBM the differentiable integration may spend on keeping each step's force mesh for the adjoint
# (12 bytes per cell and step), at most a quarter of the free device memory; beyond it the reverse
# pass recomputes the force at the restored positions.
KEEP_FORCE_MESH_BYTES = 16 << 30


def _force_mesh_store(nsteps, mesh_size, device):
    """(nsteps, nx, ny, nz, 3) float32 buffer for the steps' force meshes, or None over budget."""
    need = nsteps * int(np.prod(mesh_size)) * 12
    if nsteps == 0 or need > KEEP_FORCE_MESH_BYTES or device.type != "cuda":
        return None
    if need > torch.cuda.mem_get_info(device)[0] // 4:
        return None
    return torch.empty((nsteps,) + tuple(mesh_size) + (3,), dtype=torch.float32, device=device)


def _snapshot_vjp(solver, displacements, velocities, tg, cosmo, t0, param_names=()):
    """(state_of(idx), save_vjp(idx, x, v, g)): the interpolation state of snapshot idx and the
    (position, velocity or None, parameter dots or None) cotangents of its paint (of x, and of v and
    the cosmological parameters param_names through the drift factor when the particles are drifted
    to the lightcone) for the snapshot cotangent g."""
    interp = solver.interp_kernel
    drifting = type(interp).__name__ == "DriftInterp" or bool(getattr(interp, "drift_on_lightcone", False))
    target = interp.painting.target if interp.painting is not None else "particles"
    if target not in ("density", "particles", "spherical", "flat"):
        raise NotImplementedError(f"gradients through '{target}' snapshots are not implemented")
    # flat, density and particle snapshots of a drifted or rotated image go through the image's adjoint
    imaged = target != "spherical" and (drifting or type(interp).__name__ == "TelephotoInterp")
    mesh_size = displacements.mesh_size
    mode = _mode(displacements)
    base_state = interp.init()
    save_time = {ev[1]: ev[2] for ev in tg if ev[0] == "save"}

    def state_of(idx):
        return NBodyState(interp_state=dataclasses.replace(base_state, rotation_idx=-1, shell_idx=idx),
                          t_initial=float(t0))

    def save_vjp(idx, x, v, g):
        st = state_of(idx).interp_state
        if imaged:
            return image_snapshot_vjp(interp, st, displacements.replace(array=x), velocities.replace(array=v),
                                      save_time[idx], cosmo, g, param_names=param_names)
        if target == "particles":
            return g, None, None
        weights = 1.0 if interp.painting.weights is None else interp.painting.weights
        if target == "density":
            return engine.paint_vjp(x, g.contiguous(), mode=mode, weights=weights)[0], None, None
        rc, wd = interp.shell(st)
        if target == "flat":
            zlo, zhi, scale = flat_slab(displacements, rc, wd)
            return engine.paint_flat_vjp(x, g.contiguous(), mesh_size, zlo, zhi, scale, mode=mode,
                                         weights=weights), None, None
        tiles = interp.images(st, displacements)
        if drifting:
            return paint_drifted_vjp(displacements.replace(array=x), velocities.replace(array=v), save_time[idx], cosmo,
                                     rc, wd, interp.painting, g, tiles=tiles, param_names=param_names)
        g = ud_grade_adjoint(g.contiguous(), displacements, interp.painting)
        return engine.paint_healpix_vjp(x, g.contiguous(), mesh_size, displacements.box_size,
                                        displacements.observer_position_mpc, rc - wd / 2, rc + wd / 2,
                                        paint_nside_of(displacements, interp.painting), scheme=interp.painting.scheme,
                                        mode=mode, weights=weights, tiles=tiles), None, None

    return state_of, save_vjp


def fold_ops(solver, cosmo, t0, ts, dt0, t1, raw: bool = False) -> list:
    """The forward's item list: ("op", StepOp) with each step's leading drift folded into the
    previous op when no snapshot separates them, and ("save", idx, t) at the snapshots (raw: float64
    coefficients, for derivatives with respect to the cosmology)."""
    ops: list = []
    op0 = solver.init_op(t0, t0 + dt0, dt0, cosmo, raw=raw)
    if op0 is not None:
        ops.append(("op", op0))
    tg = time_grid(t0, ts, dt0, t1)
    step_ops = iter(solver.step_ops([(ev[1], ev[2]) for ev in tg if ev[0] == "step"], dt0, float(t0), cosmo,
                                    raw=raw))
    for ev in tg:
        ops.append(("op", next(step_ops)) if ev[0] == "step" else ev)
    items: list = []
    pending: StepOp | None = None
    for item in ops:
        if item[0] == "op":
            op = StepOp(item[1].pre, item[1].kick, item[1].post, item[1].corr)
            if pending is not None and pending.post is None and op.pre is not None:
                pending.post, op.pre = op.pre, None
            if pending is not None:
                items.append(("op", pending))
            pending = op
        else:
            if pending is not None:
                items.append(("op", pending))
                pending = None
            items.append(item)
    if pending is not None:
        items.append(("op", pending))
    return items


def _integrate_grad(displacements, velocities, cosmo_in, ts, solver, t0, t1, dt0):
    """Differentiable integration: the reversible solver's fused steps forward, the reverse adjoint
    (pmx_step_reverse_vjp per step, snapshot paint VJPs) backward (integrate.py:147-193, 288-538)."""
    if not isinstance(solver, ReversibleDoubleKickDrift):
        raise NotImplementedError("gradients need the reversible solver (ReversibleDoubleKickDrift) with "
                                  "adjoint='reverse'")
    sharpen = None
    if solver.pgd_kernel.op(1.0) is not None:
        kern = solver.pgd_kernel
        if not kern.reversible:
            raise NotImplementedError(f"{type(kern).__name__} is not reversible: the reverse adjoint needs the "
                                      "sharpening kernel")
        sharpen = (float(kern.alpha), float(kern.kl), float(kern.ks))
    interp = solver.interp_kernel
    cosmo, _, _ = autodiff.split_cosmo(cosmo_in)
    mesh_size = displacements.mesh_size
    mode = _mode(displacements)
    tg = time_grid(t0, ts, dt0, t1)
    grid = [(ev[1], ev[2]) for ev in tg if ev[0] == "step"]
    coef = solver.step_coeffs(grid, dt0, float(t0), cosmo)
    events, i = [], 0
    for ev in tg:
        if ev[0] == "step":
            events.append(("step", i, float(coef[i, 0]), float(coef[i, 1])))
            i += 1
        else:
            events.append(("save", ev[1], ev[2]))
    saved_fields = []
    _state, save_vjp = _snapshot_vjp(solver, displacements, velocities, tg, cosmo, t0,
                                     param_names=tuple(autodiff.split_cosmo(cosmo_in)[1]))

    def run(x0, v0, cosmo_):
        x = x0.detach().contiguous().clone()
        v = v0.detach().contiguous().clone()
        fs = force_scale(cosmo_)
        snaps = []
        saved_fields.clear()
        # the sharpened reverse step recomputes its combined force + boost mesh, so none is kept
        meshes = None if sharpen else _force_mesh_store(len(grid), mesh_size, x.device)
        for ev in events:
            if ev[0] == "step":
                if sharpen:  # SharpeningKernel.apply before the kick and drift, as run_op does
                    engine.pgd(x, v, mesh_size, mode=mode, alpha=sharpen[0], kl=sharpen[1], ks=sharpen[2],
                               target="v", v_scale=1.0 / ev[3])
                engine.step(x, v, mesh_size, mode=mode, force_scale=fs, c_kick=ev[2], c_drift_post=ev[3],
                            force_mesh=None if meshes is None else meshes[ev[1]])
            else:
                _, idx, tt = ev
                fld = solver.save_at(displacements.replace(array=x, scale_factors=tt),
                                     velocities.replace(array=v, scale_factors=tt), tt, dt0, _state(idx), cosmo_)
                saved_fields.append(fld)
                snaps.append(fld.array)
        return x, v, torch.stack(snaps), len(grid), meshes

    run.info = {"mesh": mesh_size, "mode": mode, "events": events, "force_scale": force_scale, "save_vjp": save_vjp,
                "sharpen": sharpen}

    def coef_fn(c):
        return solver.step_coeffs(grid, dt0, float(t0), c, raw=True)

    arr = autodiff.nbody_reverse(displacements.array, velocities.array, cosmo_in, run, coef_fn)
    out = _stack(saved_fields)
    return out.replace(array=arr)


def _op_vjp(op: StepOp, x, v, v_after, xb, vb, fs, mesh_size, mode, force_mesh=None):
    """Pull (xb, vb) back in place through one StepOp applied to the state (x, v) (v_after: the
    velocity it produced) and return its coefficient cotangents [dL/dpre, dL/dkick, dL/dpost,
    dL/dv_scale] (float64 device; v_scale: the sharpening boost's factor). The op is x_a = x + pre v; [x_b = x_a + S(x_a) | v += v_scale S(x_a)];
    v_c = v + kick F(x_b); x_d = x_b + post v_c. force_mesh: the kick's force mesh kept by the forward."""
    dots = torch.zeros(4, dtype=torch.float64, device=x.device)  # [pre, kick, post, v_scale]
    x_a = x if op.pre is None else engine.axpy(x, v, f32(op.pre))
    x_b = x_a
    corr = op.corr
    if corr is not None and corr.target == "x":
        x_b = x_a.clone()
        corr.run(x_b, None, mesh_size, mode)
    if op.post is not None:  # dots[2] += <xb, v_after>; vb += post xb
        engine.adjoint_drift(xb, vb, v_after, f32(op.post), dots[2:3], mesh_size)
    if op.kick is not None:  # xb += kick J^T vb; dots[1] += <vb, F(x_b)>
        engine.forces_vjp(x_b, mesh_size, vb, mode=mode, scale=fs, dot=dots[1:2], force_mesh=force_mesh,
                          pos_bar=xb, accumulate=True, bar_scale=f32(op.kick))
    if corr is not None:
        u = xb if corr.target == "x" else vb
        g = engine.pgd_vjp(x_a, mesh_size, u, mode=mode, alpha=corr.alpha, kl=corr.kl, ks=corr.ks,
                           dot=None if corr.target == "x" else dots[3:4])  # <v_bar, S(x_a)>
        if corr.target == "x":
            xb.add_(g)
        else:
            xb.add_(g * f32(corr.v_scale))
    if op.pre is not None:  # dots[0] += <xb, v>; vb += pre xb
        engine.adjoint_drift(xb, vb, v, f32(op.pre), dots[0:1], mesh_size)
    return dots


def _integrate_checkpointed(displacements, velocities, cosmo_in, ts, solver, t0, t1, dt0, checkpoints):
    """Differentiable integration of any solver (integrate.py:147-193 with adjoint="checkpointed"):
    the forward runs the same folded StepOps as the non-differentiable path and keeps the state
    before every op (and snapshot) when that fits the HBM budget, else every stride-th op (stride =
    ceil(n_ops / checkpoints), or about sqrt(n_ops)); the backward recomputes each segment from its
    checkpoint and pulls the cotangents back through every op (the forces' and corrections' VJPs
    at the positions they saw) and every snapshot paint."""
    cosmo, names, _ = autodiff.split_cosmo(cosmo_in)
    mesh_size = displacements.mesh_size
    mode = _mode(displacements)
    tg = time_grid(t0, ts, dt0, t1)
    items = fold_ops(solver, cosmo, t0, ts, dt0, t1)
    state_of, save_vjp = _snapshot_vjp(solver, displacements, velocities, tg, cosmo, t0, param_names=tuple(names))
    op_items = [k for k, it in enumerate(items) if it[0] == "op"]
    op_index = {k: j for j, k in enumerate(op_items)}
    nops = len(op_items)
    state_bytes = 2 * displacements.array.numel() * displacements.array.element_size()
    dev = displacements.array.device
    budget = KEEP_FORCE_MESH_BYTES
    if dev.type == "cuda":
        budget = min(budget, torch.cuda.mem_get_info(dev)[0] // 4)
    kick_items = [k for k in op_items if items[k][1].kick is not None]
    mesh_slot: dict = {}
    if checkpoints is None and (len(items) + 1) * state_bytes <= budget:
        ckpt_at = set(range(len(items)))  # every op's input state: no recomputation
        mesh_bytes = int(np.prod(mesh_size)) * 12
        if (len(items) + 1) * state_bytes + len(kick_items) * mesh_bytes <= budget:
            mesh_slot = {k: j for j, k in enumerate(kick_items)}  # and every kick's force mesh
    else:
        stride = max(1, math.ceil(nops / checkpoints)) if checkpoints else max(1, math.isqrt(max(nops, 1)))
        ckpt_at = {0} | {op_items[j] for j in range(0, nops, stride)}
    saved_fields = []

    def forward(x0, v0, cosmo_):
        x = x0.detach().contiguous().clone()
        v = v0.detach().contiguous().clone()
        ckpts, snaps = {}, []
        saved_fields.clear()
        meshes = (torch.empty((len(mesh_slot),) + tuple(mesh_size) + (3,), dtype=torch.float32, device=x.device)
                  if mesh_slot else None)
        for k, it in enumerate(items):
            if k in ckpt_at:
                ckpts[k] = (x.clone(), v.clone())
            if it[0] == "op":
                run_op(it[1], x, v, mesh_size, mode, cosmo_,
                       force_mesh=meshes[mesh_slot[k]] if k in mesh_slot else None)
            else:
                _, idx, tt = it
                fld = solver.save_at(displacements.replace(array=x, scale_factors=tt),
                                     velocities.replace(array=v, scale_factors=tt), tt, dt0, state_of(idx), cosmo_)
                saved_fields.append(fld)
                snaps.append(fld.array)
        return torch.stack(snaps), (ckpts, meshes), (x, v)

    def backward(store, final, snaps_bar):
        ckpts, meshes = store
        fs = force_scale(cosmo)
        xb = torch.zeros_like(final[0])
        vb = torch.zeros_like(final[1])
        cbars = []
        pdots = None  # parameter cotangents from the snapshots (the drift factor)
        starts = sorted(ckpts)
        for si in reversed(range(len(starts))):
            k0 = starts[si]
            k1 = starts[si + 1] if si + 1 < len(starts) else len(items)
            after = ckpts[k1] if k1 < len(items) else final
            if k1 == k0 + 1:  # a kept state: the pull-backs only read it
                states = [ckpts[k0]]
            else:  # recompute the segment's states from its checkpoint
                x, v = (t.clone() for t in ckpts[k0])
                states = []
                for k in range(k0, k1):
                    states.append((x.clone(), v.clone()) if k + 1 < k1 else (x, v))
                    if items[k][0] == "op" and k + 1 < k1:
                        run_op(items[k][1], x, v, mesh_size, mode, cosmo)
            for k in reversed(range(k0, k1)):
                xs, vs = states[k - k0]
                if items[k][0] == "save":
                    gx, gv, pd = save_vjp(items[k][1], xs, vs, snaps_bar[items[k][1]])
                    xb += gx
                    if gv is not None:
                        vb += gv
                    if pd is not None:
                        pdots = pd if pdots is None else pdots + pd
                else:
                    fm = meshes[mesh_slot[k]] if k in mesh_slot else None
                    cbars.append((op_index[k], _op_vjp(items[k][1], xs, vs, after[1], xb, vb, fs, mesh_size, mode,
                                                       force_mesh=fm)))
                after = (xs, vs)
        return xb, vb, cbars, pdots

    def coef_fn(c):
        return np.asarray([[o.pre or 0.0, o.kick or 0.0, o.post or 0.0,
                            o.corr.v_scale if o.corr is not None and o.corr.target == "v" else 0.0]
                           for it in fold_ops(solver, c, t0, ts, dt0, t1, raw=True) if it[0] == "op"
                           for o in (it[1],)],
                          dtype=np.float64)

    arr = autodiff.nbody_checkpointed(displacements.array, velocities.array, cosmo_in,
                                      (forward, backward, coef_fn, force_scale))
    out = _stack(saved_fields)
    return out.replace(array=arr)


def integrate(displacements, velocities, cosmo, ts, solver: AbstractNBodySolver, t0: float, t1: float,
              dt0: float, adjoint: str = "checkpointed", checkpoints: int | None = None) -> Any:
    if adjoint not in ("checkpointed", "reverse", "lax"):
        raise ValueError(f"Unknown adjoint type: {adjoint}")
    if adjoint == "reverse" and not solver.pgd_kernel.reversible:
        raise ValueError(f"Cannot use adjoint='reverse' with {type(solver.pgd_kernel).__name__}.")
    _, _, grad_params = autodiff.split_cosmo(cosmo)
    if displacements.array.requires_grad or velocities.array.requires_grad or grad_params:
        if adjoint == "reverse":
            return _integrate_grad(displacements, velocities, cosmo, ts, solver, t0, t1, dt0)
        return _integrate_checkpointed(displacements, velocities, cosmo, ts, solver, t0, t1, dt0, checkpoints)
    cosmo, _, _ = autodiff.split_cosmo(cosmo)
    mesh_size = displacements.mesh_size
    mode = _mode(displacements)
    x = displacements.array.contiguous().clone()
    v = velocities.array.contiguous().clone()
    state = NBodyState(interp_state=solver.interp_kernel.init(), t_initial=float(t0))
    snaps = []
    for item in fold_ops(solver, cosmo, t0, ts, dt0, t1):
        if item[0] == "op":
            run_op(item[1], x, v, mesh_size, mode, cosmo)
        else:
            _, idx, tt = item
            st = NBodyState(interp_state=dataclasses.replace(state.interp_state, rotation_idx=-1, shell_idx=idx),
                            t_initial=state.t_initial)
            dx_f = displacements.replace(array=x, scale_factors=tt)
            p_f = velocities.replace(array=v, scale_factors=tt)
            snaps.append(solver.save_at(dx_f, p_f, tt, dt0, st, cosmo))
    return _stack(snaps)
