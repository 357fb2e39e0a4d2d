"""Reverse-mode gradients of the PM path through torch.autograd.

The reference differentiates with jax.grad: a custom_vjp around the integration loop
(src/jax_fli/pm/integrate.py:147-193 forward, 288-538 backward: reverse one step, then jax.vjp of
that step) and plain JAX autodiff for the colouring of the white noise (src/jax_fli/initial.py:
124-199), LPT (src/jax_fli/pm/lpt.py:119) and the snapshot paints. Here each stage is a
torch.autograd.Function whose backward calls the hand-written adjoint kernels of the C ABI:

  interpolate_initial_conditions   noise_bar = C ic_bar (the colouring C is symmetric)
  lpt                              ic_bar = force_mesh^T (D (dx_bar + f p_bar))
  nbody, reversible solver,        one pmx_step_reverse_vjp per step (reverse the state, pull the
  adjoint="reverse"                cotangents back, reusing the reverse pass's force mesh);
                                   snapshot paints through pmx_paint_cic_vjp

Cosmological parameters may be 0-d torch tensors (``Cosmology(Omega_c=torch.tensor(..,
requires_grad=True))``). The host scalars they feed -- growth factors, kick/drift coefficients,
force scale 1.5 Omega_m, the P(k) table -- are differentiated by float64 central differences of the
host functions, and chained with the device cotangents of those scalars (dot products the adjoint
kernels accumulate).
"""

from __future__ import annotations

import dataclasses
import math
import weakref

import numpy as np
import torch

from . import cosmology as _c
from . import engine


# ---------------------------------------------------------------------------- parameters
_PARAM_VALUES: dict = {}


def _param_value(v: torch.Tensor) -> float:
    """Host value of a scalar parameter tensor, read from the device once per tensor version (the
    read waits for the device, so the stages of one forward model share it). Entries are keyed by
    identity (a tensor's == is elementwise and would itself read the device)."""
    hit = _PARAM_VALUES.get(id(v))
    if hit is not None and hit[0]() is v and hit[1] == v._version:
        return hit[2]
    val = float(v.detach().cpu())
    if len(_PARAM_VALUES) >= 64:
        for key in [k for k, e in _PARAM_VALUES.items() if e[0]() is None]:
            del _PARAM_VALUES[key]
    _PARAM_VALUES[id(v)] = (weakref.ref(v), v._version, val)
    return val


def split_cosmo(cosmo):
    """(float-valued Cosmology, names, tensors) of the parameters given as tensors needing grad."""
    if cosmo is None:
        return None, [], []
    vals, names, tensors = {}, [], []
    for f in dataclasses.fields(cosmo):
        v = getattr(cosmo, f.name)
        if isinstance(v, torch.Tensor):
            vals[f.name] = _param_value(v)
            if v.requires_grad:
                names.append(f.name)
                tensors.append(v)
    return (dataclasses.replace(cosmo, **vals) if vals else cosmo), names, tensors


def fd(fn, cosmo, name, rel=1e-4):
    """float64 central difference of a host function of the cosmology w.r.t. one parameter."""
    v = float(getattr(cosmo, name))
    h = rel * max(abs(v), 1e-2)
    hi = np.asarray(fn(dataclasses.replace(cosmo, **{name: v + h})), dtype=np.float64)
    lo = np.asarray(fn(dataclasses.replace(cosmo, **{name: v - h})), dtype=np.float64)
    return (hi - lo) / (2.0 * h)


def _param_grad(t: torch.Tensor, value):
    """Gradient for parameter tensor t from a float or a device scalar (kept asynchronous when t
    lives on the same device)."""
    if isinstance(value, torch.Tensor):
        return value.to(dtype=t.dtype, device=t.device).reshape(t.shape)
    return torch.as_tensor(value, dtype=t.dtype, device=t.device).reshape(t.shape)


# ---------------------------------------------------------------------------- initial conditions
def _kgrid(mesh_size, box_size, device):
    """|k| (h/Mpc, float64) on the r2c grid (nx, ny, nz//2+1), the wavevectors of the colouring."""
    nx, ny, nz = mesh_size
    kx = 2 * math.pi * torch.fft.fftfreq(nx, d=box_size[0] / nx, dtype=torch.float64, device=device)
    ky = 2 * math.pi * torch.fft.fftfreq(ny, d=box_size[1] / ny, dtype=torch.float64, device=device)
    kz = 2 * math.pi * torch.fft.rfftfreq(nz, d=box_size[2] / nz, dtype=torch.float64, device=device)
    return torch.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)


def _interp_setup(x, xp):
    """Bracketing indices and fractions of a linear interpolation with linear extrapolation at both
    ends (the colouring's table lookup); apply to a table fp with _interp_apply."""
    ind = torch.clamp(torch.searchsorted(xp, x, right=False) - 1, 0, xp.numel() - 2)
    x0, x1 = xp[ind], xp[ind + 1]
    return ind, (x - x0) / (x1 - x0)


def _interp_apply(setup, fp):
    ind, t = setup
    f0, f1 = fp[ind], fp[ind + 1]
    return f0 + t * (f1 - f0)


_SETUPS: dict = {}


def _colour_setup(mesh_size, box_size, k, device):
    """Interpolation set-up of the table k at the r2c grid's |k|, the half-spectrum weights
    (1 on the kz = 0 and Nyquist planes, 2 elsewhere) over N, cached per (mesh, box, table, device)."""
    k = np.ascontiguousarray(k, dtype=np.float64)
    key = (tuple(mesh_size), tuple(box_size), k.tobytes(), str(device))
    hit = _SETUPS.get(key)
    if hit is None:
        if len(_SETUPS) >= 16:
            _SETUPS.clear()
        setup = _interp_setup(_kgrid(mesh_size, box_size, device), engine.to_device(k, device))
        nz = mesh_size[2]
        wz = torch.full((nz // 2 + 1,), 2.0, dtype=torch.float64, device=device)
        wz[0] = 1.0
        if nz % 2 == 0:
            wz[-1] = 1.0
        hit = _SETUPS[key] = (setup, wz / float(np.prod(mesh_size)))
    return hit


class _Colour(torch.autograd.Function):
    @staticmethod
    def forward(ctx, noise, spec, *params):
        colour_fn, table_fn, cosmo, names, mesh_size, box_size = spec
        ctx.save_for_backward(noise)
        ctx.spec, ctx.params = spec, params
        return colour_fn(noise, cosmo)

    @staticmethod
    def backward(ctx, g):
        (noise,) = ctx.saved_tensors
        colour_fn, table_fn, cosmo, names, mesh_size, box_size = ctx.spec
        g = g.contiguous()
        gn = colour_fn(g, cosmo) if ctx.needs_input_grad[0] else None
        grads = []
        if names:
            # <g, irfft(rfft(noise) dA/dtheta)> = (1/N) sum_k w_k Re(conj(G_k) E_k) dA/dtheta(k) over the
            # half spectrum (w_k = 1 on the kz = 0 and Nyquist planes, 2 elsewhere), A = sqrt(interp(P N/V))
            # device work is queued first (no host transfer waits on it); the host's float64
            # derivatives of the table are computed while it runs
            dev = noise.device
            cross = (torch.conj(torch.fft.rfftn(g)) * torch.fft.rfftn(noise.to(torch.float32))).real.double()
            k, pt = table_fn(cosmo)
            setup, wz = _colour_setup(mesh_size, box_size, k, dev)
            amp = torch.sqrt(torch.clamp(_interp_apply(setup, engine.to_device(pt, dev)), min=0.0))
            cross = cross * wz
            inv2a = torch.where(amp > 0, 0.5 / torch.where(amp > 0, amp, torch.ones_like(amp)), torch.zeros_like(amp))
            cross = cross * inv2a
            dpts = [fd(lambda c: table_fn(c)[1], cosmo, name) for name in names]
            sums = [(cross * _interp_apply(setup, engine.to_device(d, dev))).sum() for d in dpts]
            grads = [_param_grad(t, v) for t, v in zip(ctx.params, sums)]
        return (gn, None, *grads)


def colour(noise, cosmo_in, mesh_size, box_size, colour_fn, table_fn):
    """Differentiable colouring: colour_fn(noise, cosmo), table_fn(cosmo) = (k, P(k) N / V)."""
    cosmo, names, tensors = split_cosmo(cosmo_in)
    if not (noise.requires_grad or tensors):
        return colour_fn(noise, cosmo)
    return _Colour.apply(noise, (colour_fn, table_fn, cosmo, names, tuple(mesh_size), tuple(box_size)), *tensors)


# ---------------------------------------------------------------------------- LPT
def _lpt_factors(cosmo, a):
    D = float(_c.growth_factor(cosmo, a))
    vf = float(a**2 * _c.growth_rate(cosmo, a) * _c.E(cosmo, a))
    return D, vf


class _LPT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ic, spec, *params):
        cosmo, names, a, particles = spec
        D, vf = _lpt_factors(cosmo, a)
        dx, p = engine.lpt1(ic.contiguous(), D, vf, particles=particles)
        ctx.save_for_backward(dx)
        ctx.spec, ctx.params, ctx.DV = spec, params, (D, vf)
        ctx.mesh = tuple(ic.shape)
        return dx, p

    @staticmethod
    def backward(ctx, dx_bar, p_bar):
        (dx,) = ctx.saved_tensors
        D, vf = ctx.DV
        dx_bar = None if dx_bar is None else dx_bar.contiguous()
        p_bar = None if p_bar is None else p_bar.contiguous()
        ic_bar = None
        if ctx.needs_input_grad[0] and (dx_bar is not None or p_bar is not None):
            ic_bar = engine.lpt1_vjp(dx_bar, p_bar, ctx.mesh, D, vf)
        grads = []
        if ctx.params:
            # dx = D F, p = vf D F:  dL/dD = <dx_bar + vf p_bar, dx> / D,  dL/dvf = <p_bar, dx>
            zero = torch.zeros((), dtype=torch.float64, device=dx.device)
            gx = zero if dx_bar is None else (dx_bar * dx).sum(dtype=torch.float64)
            gv = zero if p_bar is None else (p_bar * dx).sum(dtype=torch.float64)
            gD = (gx + vf * gv) / D
            cosmo, names, a, _particles = ctx.spec
            ders = [fd(lambda c: _lpt_factors(c, a), cosmo, name) for name in names]  # host, while the device runs
            for (dD, dv), t in zip(ders, ctx.params):
                grads.append(_param_grad(t, gD * dD + gv * dv))
        return (ic_bar, None, *grads)


def lpt1(ic, cosmo_in, a, particles=None):
    """(dx, p) of first-order LPT at the particle lattice `particles` (default: one per cell)."""
    cosmo, names, tensors = split_cosmo(cosmo_in)
    if not (ic.requires_grad or tensors):
        D, vf = _lpt_factors(cosmo, a)
        return engine.lpt1(ic.contiguous(), D, vf, particles=particles)
    return _LPT.apply(ic, (cosmo, names, a, particles), *tensors)


# ---------------------------------------------------------------------------- N-body
class _NBodyReverse(torch.autograd.Function):
    """Forward: run `plan` (a list of ('step', t0, t1) / ('save', idx)) with the reversible solver's
    fused steps, returning the stacked snapshots. Backward: the reverse adjoint."""

    @staticmethod
    def forward(ctx, x0, v0, spec, *params):
        run, coef_fn, cosmo, names = spec
        x, v, snaps, steps, meshes = run(x0, v0, cosmo)
        ctx.save_for_backward(x.clone(), v.clone())
        ctx.spec, ctx.params, ctx.steps, ctx.meshes = spec, params, steps, meshes
        return snaps

    @staticmethod
    def backward(ctx, snaps_bar):
        run, coef_fn, cosmo, names = ctx.spec
        x, v = (t.clone() for t in ctx.saved_tensors)
        info = run.info
        mesh, mode = info["mesh"], info["mode"]
        fs = float(info["force_scale"](cosmo))
        xb = torch.zeros_like(x)
        vb = torch.zeros_like(v)
        coef_bars = []
        pdots = None  # parameter cotangents from the snapshots (the drift factor)
        for ev in reversed(info["events"]):
            if ev[0] == "save":
                gx, gv, pd = info["save_vjp"](ev[1], x, v, snaps_bar[ev[1]])
                xb += gx
                if gv is not None:
                    vb += gv
                if pd is not None:
                    pdots = pd if pdots is None else pdots + pd
            else:
                _, i, ck, cd = ev
                sh = info.get("sharpen")
                cb = engine.step_reverse_vjp(x, v, xb, vb, mesh, mode=mode, force_scale=fs, c_kick=ck, c_drift=cd,
                                             force_mesh=None if ctx.meshes is None else ctx.meshes[i],
                                             sharpen=None if sh is None else sh + (1.0 / cd,))
                if sh is not None and names:
                    # the combined step returns <v_bar, ck F + vs S>; split it with B = <v_bar, S(x)> at the
                    # restored x, and carry vs = 1 / cd into dL/dcd
                    S = torch.zeros_like(v)
                    engine.pgd(x, S, mesh, mode=mode, alpha=sh[0], kl=sh[1], ks=sh[2], target="v", v_scale=1.0)
                    B = (vb * S).sum(dtype=torch.float64)
                    cb = torch.stack([cb[0] - B / (cd * cd), (cb[1] - B / cd) / ck])
                coef_bars.append((i, cb))
        grads = []
        if names:
            # host derivatives of the step coefficients, computed while the device runs the reverse steps
            co = np.asarray(coef_fn(cosmo))
            ders = [(fd(coef_fn, cosmo, name), float(fd(info["force_scale"], cosmo, name))) for name in names]
            idx = [i for i, _ in coef_bars]
            allcb = torch.stack([cb for _, cb in coef_bars])  # (nsteps, 2) = (dL/dcd, dL/dck), on device
            for (dco, dfs), t in zip(ders, ctx.params):
                # v' = v + F ck, x' = x + v' cd; F = fs F1 enters every kick
                wck = engine.to_device(dco[idx, 0] + co[idx, 0] / fs * dfs, allcb.device)
                wcd = engine.to_device(dco[idx, 1], allcb.device)
                grads.append(_param_grad(t, (allcb[:, 1] * wck + allcb[:, 0] * wcd).sum()))
            if pdots is not None:
                grads = [g + _param_grad(t, pdots[j]) for j, (g, t) in enumerate(zip(grads, ctx.params))]
        ctx.meshes = None  # the kept force meshes are released with the backward pass
        return (xb, vb, None, *grads)


def nbody_reverse(x0, v0, cosmo_in, run, coef_fn):
    cosmo, names, tensors = split_cosmo(cosmo_in)
    return _NBodyReverse.apply(x0, v0, (run, coef_fn, cosmo, names), *tensors)


class _NBodyCheckpointed(torch.autograd.Function):
    """Forward: the folded StepOps with checkpoints, returning the stacked snapshots. Backward: the
    checkpointed adjoint (segments recomputed from their checkpoints, per-op VJPs)."""

    @staticmethod
    def forward(ctx, x0, v0, spec, *params):
        run_fwd, run_bwd, coef_fn, fs_fn, cosmo, names = spec
        snaps, ckpts, final = run_fwd(x0, v0, cosmo)
        ctx.spec, ctx.params, ctx.ckpts, ctx.final = spec, params, ckpts, final
        return snaps

    @staticmethod
    def backward(ctx, snaps_bar):
        run_fwd, run_bwd, coef_fn, fs_fn, cosmo, names = ctx.spec
        xb, vb, cbars, pdots = run_bwd(ctx.ckpts, ctx.final, snaps_bar)
        ctx.ckpts = ctx.final = None  # the checkpoints are released with the backward pass
        grads = []
        if names:
            # host derivatives of the op coefficients, computed while the device runs the adjoint
            co = coef_fn(cosmo)
            fs = float(fs_fn(cosmo))
            ders = [(fd(coef_fn, cosmo, n), float(fd(fs_fn, cosmo, n))) for n in names]
            idx = [i for i, _ in cbars]
            allcb = torch.stack([cb for _, cb in cbars])  # (n_ops, 4) = dL/d(pre, kick, post, v_scale), on device
            for (dco, dfs), t in zip(ders, ctx.params):
                w = dco[idx].copy()
                w[:, 1] += co[idx, 1] / fs * dfs  # the force carries force_scale into every kick
                grads.append(_param_grad(t, (allcb * engine.to_device(w, allcb.device)).sum()))
            if pdots is not None:
                grads = [g + _param_grad(t, pdots[j]) for j, (g, t) in enumerate(zip(grads, ctx.params))]
        return (xb, vb, None, *grads)


def nbody_checkpointed(x0, v0, cosmo_in, fns):
    cosmo, names, tensors = split_cosmo(cosmo_in)
    return _NBodyCheckpointed.apply(x0, v0, tuple(fns) + (cosmo, names), *tensors)
