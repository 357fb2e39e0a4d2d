"""Gradients of the differentiable forward model (white noise -> ICs -> LPT -> reversible KKD with
the reverse adjoint -> final density -> L2 loss) against float64 central differences of the oracle's
forward model, for the noise field (random directions) and for (Omega_c, sigma8).

The reference's own check is jax.grad against finite differences (config: 16 random directions,
eps = 1e-3); JAX is not available, so the float64 oracle stands in for the reference's FP64 run.
Tolerance 2e-3 relative (float32 forward and adjoint against a float64 finite difference).
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest
import torch

from oracle import nbody_oracle as nbo
from oracle import pm_oracle as orc

pytestmark = pytest.mark.gpu

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
N, BOX, A0, T1, DT0 = 32, 200.0, 0.1, 0.5, 0.1
KT = np.logspace(-4, 1, 128)


def _oracle_loss(noise, target, Oc=P["Omega_c"], s8=P["sigma8"]):
    o = nbo.Cosmo(**{**P, "Omega_c": Oc, "sigma8": s8})
    ic = nbo.color_field(noise, (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
    dx, p = nbo.lpt1(o, ic, A0)
    x, _ = nbo.run_nbody(o, dx, p, A0, T1, DT0, [T1], solver="kkd")
    rho = orc.cic_paint(x, (N,) * 3, "relative", dtype=np.float64)
    delta = rho / rho.mean() - 1.0
    return float(((delta - target) ** 2).sum())


def _gpu_loss(pkg, noise_t, target_t, c):
    ic = pkg.interpolate_initial_conditions(noise_t, (N,) * 3, (BOX,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=A0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")))
    out = pkg.nbody(c, dx, p, t0=A0, t1=T1, dt0=DT0, ts=[T1], solver=solver, adjoint="reverse")
    rho = out.array.double()
    delta = rho / rho.mean() - 1.0
    return ((delta - target_t) ** 2).sum()


def test_forward_model_gradient_matches_oracle_fd(pkg, cuda):
    rng = np.random.default_rng(60)
    noise = rng.standard_normal((N,) * 3).astype(np.float32)
    target = rng.standard_normal((N,) * 3) * 0.5
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    s8 = torch.tensor(P["sigma8"], dtype=torch.float64, requires_grad=True)
    c = pkg.Cosmology(**{**P, "Omega_c": Oc, "sigma8": s8})
    noise_t = torch.from_numpy(noise).to(cuda).requires_grad_(True)
    target_t = torch.from_numpy(target).to(cuda)
    L = _gpu_loss(pkg, noise_t, target_t, c)
    L.backward()
    L0 = _oracle_loss(noise.astype(np.float64), target)
    assert abs(float(L) - L0) / L0 < 1e-5
    g = noise_t.grad.double().cpu().numpy()
    eps = 1e-7
    for j in range(2):
        d = np.random.default_rng(61 + j).standard_normal(noise.shape)
        fd = (_oracle_loss(noise + eps * d, target) - _oracle_loss(noise - eps * d, target)) / (2 * eps)
        got = float((g * d).sum())
        print(f"noise direction {j}: adjoint {got:.6e} fd {fd:.6e}")
        assert abs(got - fd) / abs(fd) < 2e-3
    h = 1e-6
    for name, t in (("Omega_c", Oc), ("sigma8", s8)):
        v = P[name]
        kw_hi, kw_lo = {"Oc" if name == "Omega_c" else "s8": v + h}, {"Oc" if name == "Omega_c" else "s8": v - h}
        fd = (_oracle_loss(noise, target, **kw_hi) - _oracle_loss(noise, target, **kw_lo)) / (2 * h)
        got = float(t.grad)
        print(f"{name}: adjoint {got:.6e} fd {fd:.6e}")
        assert abs(got - fd) / abs(fd) < 2e-3


@pytest.mark.parametrize("keep", [True, False])
def test_nbody_gradient_is_the_reverse_adjoint(pkg, cuda, keep, monkeypatch):
    """d L / d (dx, p) of the N-body stage alone, against the float64 oracle run by finite differences;
    with the steps' force meshes kept for the reverse pass and with the force recomputed there."""
    if not keep:
        monkeypatch.setattr(importlib.import_module(pkg.__name__ + ".pm.integrate"), "KEEP_FORCE_MESH_BYTES", 0)
    rng = np.random.default_rng(62)
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    ic = pkg.gaussian_initial_conditions(7, (N,) * 3, (BOX,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=A0)
    x0, v0 = dx.array.clone(), p.array.clone()
    xb = rng.standard_normal(x0.shape)
    dxf = dx.replace(array=x0.clone().requires_grad_(True))
    pf = p.replace(array=v0.clone().requires_grad_(True))
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="particles")))
    out = pkg.nbody(c, dxf, pf, t0=A0, t1=T1, dt0=DT0, ts=[T1], solver=solver, adjoint="reverse")
    (out.array.double() * torch.from_numpy(xb).to(cuda)).sum().backward()
    x0n, v0n = x0.cpu().numpy().astype(np.float64), v0.cpu().numpy().astype(np.float64)

    def loss(xx, vv):
        x, _ = nbo.run_nbody(o, xx, vv, A0, T1, DT0, [T1], solver="kkd")
        return float((x * xb).sum())

    eps = 1e-7
    for which, grad in (("x", dxf.array.grad), ("v", pf.array.grad)):
        d = rng.standard_normal(x0n.shape)
        if which == "x":
            fd = (loss(x0n + eps * d, v0n) - loss(x0n - eps * d, v0n)) / (2 * eps)
        else:
            fd = (loss(x0n, v0n + eps * d) - loss(x0n, v0n - eps * d)) / (2 * eps)
        got = float((grad.double().cpu().numpy() * d).sum())
        print(f"d/d{which}: adjoint {got:.6e} fd {fd:.6e}")
        assert abs(got - fd) / abs(fd) < 2e-3
