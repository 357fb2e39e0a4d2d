"""The checkpointed adjoint (the reference's default adjoint="checkpointed", integrate.py:147-193)
for any solver on the GPU: gradients of a density loss through the default KDK solver
(EfficientDriftDoubleKick) with and without recomputed segments, through the white noise -> ICs ->
LPT -> N-body forward model including (Omega_c, sigma8), through KDK with the PGD kernel, and through
the reversible solver (where it must agree with the reverse adjoint). References: float64 central
differences of the oracle's forward model; tolerance 2e-3 (float32 forward and adjoint)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import nbody_oracle as nbo
from oracle import pm_oracle as orc

pytestmark = pytest.mark.gpu

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
N, BOX, A0, T1, DT0 = 32, 200.0, 0.1, 0.5, 0.1
KT = np.logspace(-4, 1, 128)


def _oracle_loss(noise, target, Oc=P["Omega_c"], s8=P["sigma8"], solver="kdk", correction=None):
    o = nbo.Cosmo(**{**P, "Omega_c": Oc, "sigma8": s8})
    ic = nbo.color_field(noise, (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
    dx, p = nbo.lpt1(o, ic, A0)
    x, _ = nbo.run_nbody(o, dx, p, A0, T1, DT0, [T1], solver=solver, correction=correction)
    rho = orc.cic_paint(x, (N,) * 3, "relative", dtype=np.float64)
    delta = rho / rho.mean() - 1.0
    return float(((delta - target) ** 2).sum())


def _gpu_loss(pkg, noise_t, target_t, c, solver, **kw):
    ic = pkg.interpolate_initial_conditions(noise_t, (N,) * 3, (BOX,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=A0)
    out = pkg.nbody(c, dx, p, t0=A0, t1=T1, dt0=DT0, ts=[T1], solver=solver, **kw)
    rho = out.array.double()
    delta = rho / rho.mean() - 1.0
    return ((delta - target_t) ** 2).sum()


def _density_solver(pkg, kind="kdk", pgd=None):
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="density"))
    if kind == "kdk":
        return pkg.EfficientDriftDoubleKick(interp_kernel=interp, **({} if pgd is None else {"pgd_kernel": pgd}))
    return pkg.ReversibleDoubleKickDrift(interp_kernel=interp)


@pytest.mark.parametrize("checkpoints", [None, 2])
def test_default_kdk_forward_model_gradient(pkg, cuda, checkpoints):
    """Default solver and adjoint; checkpoints=2 recomputes the steps between two checkpoints."""
    rng = np.random.default_rng(90)
    noise = rng.standard_normal((N,) * 3).astype(np.float32)
    target = rng.standard_normal((N,) * 3) * 0.5
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    s8 = torch.tensor(P["sigma8"], dtype=torch.float64, requires_grad=True)
    c = pkg.Cosmology(**{**P, "Omega_c": Oc, "sigma8": s8})
    noise_t = torch.from_numpy(noise).to(cuda).requires_grad_(True)
    L = _gpu_loss(pkg, noise_t, torch.from_numpy(target).to(cuda), c, _density_solver(pkg), checkpoints=checkpoints)
    L.backward()
    L0 = _oracle_loss(noise.astype(np.float64), target)
    assert abs(float(L) - L0) / L0 < 1e-5
    g = noise_t.grad.double().cpu().numpy()
    eps = 1e-7
    for j in range(2):
        d = np.random.default_rng(91 + j).standard_normal(noise.shape)
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


def test_kdk_with_pgd_gradient(pkg, cuda):
    rng = np.random.default_rng(92)
    noise = rng.standard_normal((N,) * 3).astype(np.float32)
    target = rng.standard_normal((N,) * 3) * 0.5
    c = pkg.Cosmology(**P)
    noise_t = torch.from_numpy(noise).to(cuda).requires_grad_(True)
    solver = _density_solver(pkg, pgd=pkg.PGDKernel(alpha=0.2, kl=0.3, ks=1.0))
    L = _gpu_loss(pkg, noise_t, torch.from_numpy(target).to(cuda), c, solver)
    L.backward()
    corr = ("pgd", 0.2, 0.3, 1.0)
    assert abs(float(L) - _oracle_loss(noise.astype(np.float64), target, correction=corr)) / float(L) < 1e-5
    g = noise_t.grad.double().cpu().numpy()
    d = np.random.default_rng(93).standard_normal(noise.shape)
    eps = 1e-7
    fd = (_oracle_loss(noise + eps * d, target, correction=corr) -
          _oracle_loss(noise - eps * d, target, correction=corr)) / (2 * eps)
    got = float((g * d).sum())
    print(f"pgd: adjoint {got:.6e} fd {fd:.6e}")
    assert abs(got - fd) / abs(fd) < 2e-3


def test_checkpointed_equals_reverse_on_the_reversible_solver(pkg, cuda):
    rng = np.random.default_rng(94)
    noise = torch.from_numpy(rng.standard_normal((N,) * 3).astype(np.float32)).to(cuda)
    target = torch.from_numpy(rng.standard_normal((N,) * 3) * 0.5).to(cuda)
    c = pkg.Cosmology(**P)
    grads = []
    for adjoint in ("checkpointed", "reverse"):
        nz = noise.clone().requires_grad_(True)
        _gpu_loss(pkg, nz, target, c, _density_solver(pkg, "kkd"), adjoint=adjoint).backward()
        grads.append(nz.grad.double().cpu().numpy())
    err = np.linalg.norm(grads[0] - grads[1]) / np.linalg.norm(grads[1])
    print(f"checkpointed vs reverse: rel L2 {err:.2e}")
    assert err < 1e-4
