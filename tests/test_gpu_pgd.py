"""PGD and sharpening corrections (src/jax_fli/pm/correction.py:131-333) on the GPU against the
float64 oracle: the displacement S(x) through pmx_pgd, full runs of the KDK solver with PGDKernel and
of the reversible solver with SharpeningKernel, and the sharpened step's exact reverse.
Tolerances: S 1e-5 relative L2 (the force bound), positions after a run 1e-4."""

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
