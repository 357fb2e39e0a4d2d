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

import numpy as np
import torch

from . import cosmology as _c
from . import engine


# ---------------------------------------------------------------------------- parameters
def split_cosmo(cosmo):
    """(float-valued Cosmology, names, tensors) of the parameters given as tensors needing grad."""
    if cosmo is None:
        return None, [], []
    vals, names, tensors = {}, [], []
    for f in dataclasses.fields(cosmo):
        v = getattr(cosmo, f.name)
        if isinstance(v, torch.Tensor):
            vals[f.name] = float(v.detach().cpu())
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


def _param_grad(t: torch.Tensor, value: float):
    return torch.as_tensor(value, dtype=t.dtype, device=t.device).reshape(t.shape)


# ---------------------------------------------------------------------------- initial conditions
def _kgrid(mesh_size, box_size, device):
    """|k| (h/Mpc, float64) on the r2c grid (nx, ny, nz//2+1), the wavevectors of the colouring."""
    nx, ny, nz = mesh_size
    kx = 2 * math.pi * torch.fft.fftfreq(nx, d=box_size[0] / nx, dtype=torch.float64, device=device)
    ky = 2 * math.pi * torch.fft.fftfreq(ny, d=box_size[1] / ny, dtype=torch.float64, device=device)
    kz = 2 * math.pi * torch.fft.rfftfreq(nz, d=box_size[2] / nz, dtype=torch.float64, device=device)
    return torch.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)


def _interp_lin(x, xp, fp):
    """Linear interpolation with linear extrapolation at both ends (the colouring's table lookup)."""
    ind = torch.clamp(torch.searchsorted(xp, x, right=False) - 1, 0, xp.numel() - 2)
    x0, x1, f0, f1 = xp[ind], xp[ind + 1], fp[ind], fp[ind + 1]
    return f0 + (x - x0) * (f1 - f0) / (x1 - x0)


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
            # d ic / d theta = irfft(rfft(noise) dA/dtheta), A = sqrt(interp(P N / V)) per mode
            kk = _kgrid(mesh_size, box_size, noise.device)
            k, pt = table_fn(cosmo)
            kt = torch.as_tensor(k, dtype=torch.float64, device=noise.device)
            amp = torch.sqrt(torch.clamp(_interp_lin(kk, kt, torch.as_tensor(pt, device=noise.device)), min=0.0))
            nf = torch.fft.rfftn(noise.double())
            for name, t in zip(names, ctx.params):
                dpt = torch.as_tensor(fd(lambda c: table_fn(c)[1], cosmo, name), device=noise.device)
                damp = torch.where(amp > 0, _interp_lin(kk, kt, dpt) / (2.0 * amp), torch.zeros_like(amp))
                dic = torch.fft.irfftn(nf * damp, s=mesh_size)
                grads.append(_param_grad(t, float((g.double() * dic).sum())))
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
        cosmo, names, a = spec
        D, vf = _lpt_factors(cosmo, a)
        dx, p = engine.lpt1(ic.contiguous(), D, vf)
        ctx.save_for_backward(dx)
        ctx.spec, ctx.params, ctx.DV = spec, params, (D, vf)
        return dx, p

    @staticmethod
    def backward(ctx, dx_bar, p_bar):
        (dx,) = ctx.saved_tensors
        cosmo, names, a = ctx.spec
        D, vf = ctx.DV
        dx_bar = torch.zeros_like(dx) if dx_bar is None else dx_bar.contiguous()
        p_bar = torch.zeros_like(dx) if p_bar is None else p_bar.contiguous()
        w = (dx_bar + vf * p_bar) * D
        ic_bar = engine.force_mesh_vjp(w.contiguous(), scale=1.0) if ctx.needs_input_grad[0] else None
        grads = []
        if names:
            # dx = D F, p = vf D F:  dL/dD = <dx_bar + vf p_bar, dx> / D,  dL/dvf = <p_bar, dx>
            gD = float(((dx_bar.double() + vf * p_bar.double()) * dx.double()).sum()) / D
            gv = float((p_bar.double() * dx.double()).sum())
            for name, t in zip(names, ctx.params):
                dD, dv = fd(lambda c: _lpt_factors(c, a), cosmo, name)
                grads.append(_param_grad(t, gD * dD + gv * dv))
        return (ic_bar, None, *grads)


def lpt1(ic, cosmo_in, a):
    cosmo, names, tensors = split_cosmo(cosmo_in)
    if not (ic.requires_grad or tensors):
        D, vf = _lpt_factors(cosmo, a)
        return engine.lpt1(ic.contiguous(), D, vf)
    return _LPT.apply(ic, (cosmo, names, a), *tensors)


# ---------------------------------------------------------------------------- N-body
class _NBodyReverse(torch.autograd.Function):
    """Forward: run `plan` (a list of ('step', t0, t1) / ('save', idx)) with the reversible solver's
    fused steps, returning the stacked snapshots. Backward: the reverse adjoint."""

    @staticmethod
    def forward(ctx, x0, v0, spec, *params):
        run, coef_fn, cosmo, names = spec
        x, v, snaps, steps = run(x0, v0, cosmo)
        ctx.save_for_backward(x.clone(), v.clone())
        ctx.spec, ctx.params, ctx.steps = spec, params, steps
        return snaps

    @staticmethod
    def backward(ctx, snaps_bar):
        run, coef_fn, cosmo, names = ctx.spec
        x, v = (t.clone() for t in ctx.saved_tensors)
        info = run.info
        mesh, mode, target = info["mesh"], info["mode"], info["target"]
        fs = float(info["force_scale"](cosmo))
        xb = torch.zeros_like(x)
        vb = torch.zeros_like(v)
        coef_bars = []
        for ev in reversed(info["events"]):
            if ev[0] == "save":
                g = snaps_bar[ev[1]]
                if target == "density":
                    pb, _ = engine.paint_vjp(x, g.contiguous(), mode=mode)
                    xb += pb
                else:
                    xb += g
            else:
                _, i, ck, cd = ev
                cb = engine.step_reverse_vjp(x, v, xb, vb, mesh, mode=mode, force_scale=fs, c_kick=ck, c_drift=cd)
                coef_bars.append((i, cb.clone()))
        grads = []
        if names:
            cbs = {i: cb.cpu().numpy() for i, cb in coef_bars}
            for name, t in zip(names, ctx.params):
                dco = fd(coef_fn, cosmo, name)            # (nsteps, 2): d(ck, cd)/d name
                dfs = float(fd(info["force_scale"], cosmo, name))
                co = np.asarray(coef_fn(cosmo))
                tot = 0.0
                for i, cb in cbs.items():
                    tot += cb[1] * dco[i, 0] + cb[0] * dco[i, 1]
                    # F = fs F1 enters every kick: dL/dfs = sum_i ck_i <v_bar_i, F_i> / fs
                    tot += co[i, 0] * cb[1] / fs * dfs
                grads.append(_param_grad(t, tot))
        return (xb, vb, None, *grads)


def nbody_reverse(x0, v0, cosmo_in, run, coef_fn):
    cosmo, names, tensors = split_cosmo(cosmo_in)
    return _NBodyReverse.apply(x0, v0, (run, coef_fn, cosmo, names), *tensors)
