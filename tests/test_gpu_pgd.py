"""PGD and sharpening corrections (src/jax_fli/pm/correction.py:131-333) on the GPU against the
float64 oracle: the displacement S(x) through pmx_pgd, full runs of the KDK solver with PGDKernel and
of the reversible solver with SharpeningKernel, the sharpened step's exact reverse, its adjoint, and
the reverse-adjoint gradient through a sharpened run. Tolerances: S 1e-5 relative L2 (the force
bound), positions after a run 1e-4, the step adjoint 1e-4, the run's gradient 2e-3 against float64
central differences."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import nbody_oracle as nbo
from oracle import pm_oracle as orc

pytestmark = pytest.mark.gpu

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("mesh,npart", [((32, 32, 32), (32, 32, 32)), ((32, 32, 32), (16, 16, 16))])
def test_pgd_displacement_matches_oracle(pkg, cuda, mesh, npart):
    rng = np.random.default_rng(80)
    dx = (rng.standard_normal(npart + (3,)) * 1.2).astype(np.float32)
    v = rng.standard_normal(npart + (3,)).astype(np.float32)
    ref = orc.pgd_displacement(dx, mesh, 0.2, 0.3, 1.0)
    # S itself, as a boost of zero velocities
    x = torch.from_numpy(dx).to(cuda)
    s = torch.zeros_like(x)
    pkg.engine.pgd(x, s, mesh, alpha=0.2, kl=0.3, ks=1.0, target="v", v_scale=1.0)
    assert torch.equal(x.cpu(), torch.from_numpy(dx))
    assert rel(s.cpu().numpy(), ref) < 1e-5
    # positions: x + S in float32 (the float32 sum itself rounds at ulp(x), far above ulp(S))
    pkg.engine.pgd(x, None, mesh, alpha=0.2, kl=0.3, ks=1.0, target="x")
    assert np.abs(x.cpu().numpy() - (dx + s.cpu().numpy())).max() <= 4 * np.spacing(np.abs(dx).max())
    vd = torch.from_numpy(v).to(cuda)
    pkg.engine.pgd(torch.from_numpy(dx).to(cuda), vd, mesh, alpha=0.2, kl=0.3, ks=1.0, target="v", v_scale=-2.5)
    assert rel(vd.cpu().numpy() - v, -2.5 * ref) < 1e-4


@pytest.mark.parametrize("kind", ["pgd", "sharpen"])
def test_corrected_runs_match_oracle(pkg, cuda, kind):
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    n, box, a0, dt0 = 32, 256.0, 0.1, 0.1
    ic = pkg.gaussian_initial_conditions(9, (n, n, n), (box,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=a0)
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="particles"))
    if kind == "pgd":
        solver = pkg.EfficientDriftDoubleKick(interp_kernel=interp, pgd_kernel=pkg.PGDKernel(alpha=0.2, kl=0.3, ks=1.0))
        sname = "kdk"
    else:
        solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp,
                                               pgd_kernel=pkg.SharpeningKernel(alpha=0.2, kl=0.3, ks=1.0))
        sname = "kkd"
    out = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=dt0, ts=[1.0], solver=solver)
    x_ref, _ = nbo.run_nbody(o, dx.array.cpu().numpy(), p.array.cpu().numpy(), a0, 1.0, dt0, [1.0], solver=sname,
                             correction=(kind, 0.2, 0.3, 1.0))
    got = out.array.cpu().numpy().reshape(x_ref.shape)
    err = rel(got, x_ref)
    # the correction moves the run measurably: compare with the uncorrected oracle run
    x_plain, _ = nbo.run_nbody(o, dx.array.cpu().numpy(), p.array.cpu().numpy(), a0, 1.0, dt0, [1.0], solver=sname)
    print(f"{kind}: rel L2 vs oracle {err:.2e}; correction effect {rel(x_ref, x_plain):.2e}")
    assert err < 1e-4
    assert rel(x_ref, x_plain) > 100 * err


def test_sharpened_step_then_reverse_roundtrip(pkg, cuda):
    c = pkg.Cosmology(**P)
    n, box, a0 = 32, 256.0, 0.2
    ic = pkg.gaussian_initial_conditions(41, (n, n, n), (box,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=a0)
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="density")).update_geometry(
        ts=np.array([1.0], np.float32), r_centers=np.zeros(1), density_widths=np.ones(1), max_comoving_distance=1.0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp, pgd_kernel=pkg.SharpeningKernel())
    d0, p0, st = solver.init(dx, p, a0, a0 + 0.01, 0.01, c)
    d1, p1, st1 = solver.step(d0, p0, a0, a0 + 0.01, 0.01, st, c)
    d2, p2, _ = solver.reverse(d1, p1, a0, a0 + 0.01, 0.01, st1, c)
    assert rel(d2.array.cpu().numpy(), d0.array.cpu().numpy()) < 1e-6
    assert rel(p2.array.cpu().numpy(), p0.array.cpu().numpy()) < 1e-5


def test_correction_interfaces(pkg, cuda):
    assert pkg.SharpeningKernel().reversible and not pkg.PGDKernel().reversible
    c = pkg.Cosmology(**P)
    ic = pkg.gaussian_initial_conditions(3, (16,) * 3, (100.0,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=0.1)
    with pytest.raises(NotImplementedError):
        pkg.PGDKernel().rewind(0.5, dx, p, 1.0)
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="particles"))
    with pytest.raises(ValueError):
        pkg.nbody(c, dx, p, t0=0.1, t1=0.5, dt0=0.1, ts=[0.5], adjoint="reverse",
                  solver=pkg.ReversibleDoubleKickDrift(interp_kernel=interp, pgd_kernel=pkg.PGDKernel()))
    # the functional apply equals the in-place engine call
    x2, v2 = pkg.SharpeningKernel(alpha=0.3).apply(0.5, dx, p, full_drift_factor=0.5)
    vv = p.array.clone()
    pkg.engine.pgd(dx.array.clone(), vv, (16,) * 3, alpha=0.3, kl=0.3, ks=1.0, target="v", v_scale=2.0)
    assert torch.equal(v2.array, vv) and torch.equal(x2.array, dx.array)


def test_sharpened_step_reverse_vjp(pkg, cuda):
    """One sharpened KKD step (v += F ck + S vs; x += v cd): the reverse restores the state, and
    x_bar = xb + J_F^T (ck v_bar) + J_S^T (vs v_bar) against the float64 oracle adjoints."""
    n = 32
    mesh = (n, n, n)
    rng = np.random.default_rng(82)
    x = (rng.standard_normal(mesh + (3,)) * 1.2).astype(np.float32)
    v = (rng.standard_normal(mesh + (3,)) * 0.4).astype(np.float32)
    fs, ck, cd, alpha, kl, ks = 0.45, 0.7, 0.3, 0.2, 0.3, 1.0
    vs = 1.0 / cd
    xd, vd = torch.from_numpy(x).to(cuda), torch.from_numpy(v).to(cuda)
    pkg.engine.pgd(xd, vd, mesh, alpha=alpha, kl=kl, ks=ks, target="v", v_scale=vs)
    pkg.engine.step(xd, vd, mesh, force_scale=fs, c_kick=ck, c_drift_post=cd)
    xb = rng.standard_normal(x.shape).astype(np.float32)
    vb = rng.standard_normal(x.shape).astype(np.float32)
    xbd, vbd = torch.from_numpy(xb).to(cuda), torch.from_numpy(vb).to(cuda)
    pkg.engine.step_reverse_vjp(xd, vd, xbd, vbd, mesh, force_scale=fs, c_kick=ck, c_drift=cd,
                                sharpen=(alpha, kl, ks, vs))
    assert np.abs(xd.cpu().numpy() - x).max() < 1e-4
    assert np.abs(vd.cpu().numpy() - v).max() < 1e-4
    vmid = vb.astype(np.float64) + cd * xb
    ref = (xb + orc.pm_forces_vjp(x, mesh, ck * vmid, scale=fs)
           + orc.pm_forces_vjp(x, mesh, vs * vmid, scale=alpha, pgd=(kl, ks)))
    assert rel(vbd.cpu().numpy(), vmid) < 1e-6
    err = rel(xbd.cpu().numpy(), ref)
    print(f"sharpened reverse vjp: x_bar rel L2 {err:.2e}")
    assert err < 1e-4


def test_sharpened_nbody_gradient(pkg, cuda):
    """d L / d (dx, p) through a reversible N-body run with the sharpening boost (reverse adjoint),
    against float64 central differences of the oracle's sharpened run."""
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    n, box, a0, t1, dt0 = 16, 100.0, 0.1, 0.6, 0.1
    ic = pkg.gaussian_initial_conditions(13, (n,) * 3, (box,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=a0)
    x0, v0 = dx.array.clone(), p.array.clone()
    dxf = dx.replace(array=x0.clone().requires_grad_(True))
    pf = p.replace(array=v0.clone().requires_grad_(True))
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="particles")),
                                           pgd_kernel=pkg.SharpeningKernel(alpha=0.2, kl=0.3, ks=1.0))
    out = pkg.nbody(c, dxf, pf, t0=a0, t1=t1, dt0=dt0, ts=[t1], solver=solver, adjoint="reverse")
    rng = np.random.default_rng(83)
    xb = rng.standard_normal(tuple(x0.shape))
    (out.array.double() * torch.from_numpy(xb).to(cuda).reshape(out.array.shape)).sum().backward()
    x0n, v0n = x0.cpu().numpy().astype(np.float64), v0.cpu().numpy().astype(np.float64)

    def loss(xx, vv):
        x, _ = nbo.run_nbody(o, xx, vv, a0, t1, dt0, [t1], solver="kkd", correction=("sharpen", 0.2, 0.3, 1.0))
        return float((x.reshape(xb.shape) * xb).sum())

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


@pytest.mark.parametrize("adjoint", ["reverse", "checkpointed"])
def test_sharpened_cosmology_gradient(pkg, cuda, adjoint):
    """d L / d Omega_c through white noise -> ICs -> LPT -> sharpened reversible N-body -> density,
    against float64 central differences of the oracle's forward model (the boost's 1 / drift factor
    carries the cosmology into the correction)."""
    n, box, a0, t1, dt0 = 16, 100.0, 0.1, 0.6, 0.1
    KT = np.logspace(-4, 1, 128)
    rng = np.random.default_rng(84)
    noise = rng.standard_normal((n,) * 3).astype(np.float32)
    target = rng.standard_normal((n,) * 3) * 0.5
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    c = pkg.Cosmology(**{**P, "Omega_c": Oc})
    ic = pkg.interpolate_initial_conditions(torch.from_numpy(noise).to(cuda), (n,) * 3, (box,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=a0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")),
                                           pgd_kernel=pkg.SharpeningKernel(alpha=0.2, kl=0.3, ks=1.0))
    rho = pkg.nbody(c, dx, p, t0=a0, t1=t1, dt0=dt0, ts=[t1], solver=solver, adjoint=adjoint).array.double()
    L = ((rho / rho.mean() - 1.0 - torch.from_numpy(target).to(cuda)) ** 2).sum()
    L.backward()

    def oracle_loss(oc):
        o = nbo.Cosmo(**{**P, "Omega_c": oc})
        ic0 = nbo.color_field(noise.astype(np.float64), (box,) * 3, KT, nbo.linear_matter_power(o, KT))
        x0, v0 = nbo.lpt1(o, ic0, a0)
        x, _ = nbo.run_nbody(o, x0, v0, a0, t1, dt0, [t1], solver="kkd", correction=("sharpen", 0.2, 0.3, 1.0))
        r = orc.cic_paint(x, (n,) * 3, "relative", dtype=np.float64)
        return float(((r / r.mean() - 1.0 - target) ** 2).sum())

    assert abs(float(L) - oracle_loss(P["Omega_c"])) / float(L) < 1e-5
    h = 1e-6
    fd = (oracle_loss(P["Omega_c"] + h) - oracle_loss(P["Omega_c"] - h)) / (2 * h)
    print(f"{adjoint}: d/dOmega_c adjoint {float(Oc.grad):.6e} fd {fd:.6e}")
    assert abs(float(Oc.grad) - fd) / abs(fd) < 2e-3
