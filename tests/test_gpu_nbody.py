"""End-to-end parity of the reference-shaped API on the GPU against the CPU oracle:
initial conditions, LPT, full nbody runs (both solvers), snapshot/lightcone painting, P(k),
reversibility. Bounds: mesh relative L2 <= 1e-5 against the oracle's float64 run,
particle positions <= 1e-4, P(k) within 0.1% below k_Nyq/2. Each test also reports the oracle's own
float32-vs-float64 error for scale."""

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


@pytest.fixture(scope="module")
def setup(pkg, cuda):
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    return pkg, c, o


def _ic(pkg, c, n, box, seed):
    return pkg.gaussian_initial_conditions(seed, (n, n, n), (box,) * 3, cosmo=c)


def test_initial_conditions_match_oracle(setup):
    pkg, c, o = setup
    n, box = 32, 256.0
    noise = pkg.engine.white_noise((n, n, n), 3)
    k = np.logspace(-4, 1, 128)
    ic = pkg.interpolate_initial_conditions(noise, (n, n, n), (box,) * 3, cosmo=c)
    ref = nbo.color_field(noise.cpu().numpy(), (box,) * 3, k, pkg.cosmology.linear_matter_power(c, k))
    assert rel(ic.array.cpu().numpy(), ref) < 2e-6


def test_lpt_matches_oracle(setup):
    pkg, c, o = setup
    n, box, a0 = 32, 256.0, 0.1
    ic = _ic(pkg, c, n, box, 5)
    dx, p = pkg.lpt(c, ic, ts=a0)
    rdx, rp = nbo.lpt1(o, ic.array.cpu().numpy().astype(np.float64), a0)
    assert rel(dx.array.cpu().numpy(), rdx) < 1e-5
    assert rel(p.array.cpu().numpy(), rp) < 1e-5
    assert dx.status == pkg.FieldStatus.LPT1 and dx.unit == pkg.PositionUnit.GRID_RELATIVE


@pytest.mark.parametrize("solver_name", ["kdk", "kkd"])
@pytest.mark.parametrize("n,box,steps", [(32, 256.0, 10), (64, 250.0, 10)])
def test_nbody_final_state_matches_oracle(setup, solver_name, n, box, steps):
    pkg, c, o = setup
    a0 = 0.1
    ic = _ic(pkg, c, n, box, 11)
    dx, p = pkg.lpt(c, ic, ts=a0)
    dt0 = (1.0 - a0) / (steps - 1)
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="particles"))
    solver = (pkg.EfficientDriftDoubleKick if solver_name == "kdk" else pkg.ReversibleDoubleKickDrift)(interp_kernel=interp)
    out = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=dt0, solver=solver)
    x_gpu = out.array.cpu().numpy()
    x0 = dx.array.cpu().numpy()
    p0 = p.array.cpu().numpy()
    x64, _ = nbo.run_nbody(o, x0, p0, a0, 1.0, dt0, [1.0], solver=solver_name, dtype=np.float64)
    x32, _ = nbo.run_nbody(o, x0, p0, a0, 1.0, dt0, [1.0], solver=solver_name, dtype=np.float32)
    e_pos, e_pos32 = rel(x_gpu, x64), rel(x32, x64)
    # positions relative to the box: absolute error in cells / n
    abs_err = np.abs(x_gpu - x64).max()
    m_gpu = orc.cic_paint(x_gpu, (n, n, n), dtype=np.float64)
    m64 = orc.cic_paint(x64, (n, n, n), dtype=np.float64)
    e_mesh = rel(m_gpu, m64)
    print(f"{solver_name} {n}^3: displacement rel L2 {e_pos:.2e} (oracle f32 {e_pos32:.2e}), max |dx| err "
          f"{abs_err:.2e} cells, density rel L2 {e_mesh:.2e}")
    assert e_pos < 1e-4
    assert abs_err / n < 1e-4
    assert e_mesh < 1e-5


def test_density_snapshots_and_power(setup):
    pkg, c, o = setup
    n, box, a0 = 64, 250.0, 0.1
    ic = _ic(pkg, c, n, box, 21)
    dx, p = pkg.lpt(c, ic, ts=a0)
    ts = np.array([0.5, 1.0])
    solver = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")))
    out = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=0.09, ts=ts, solver=solver)
    assert tuple(out.array.shape) == (2, n, n, n)
    assert out.status == pkg.FieldStatus.LIGHTCONE
    # near -> far: last snapshot (a = 1) first
    assert out.scale_factors[0] == pytest.approx(1.0) and out.scale_factors[1] == pytest.approx(0.5)
    snaps = {}
    nbo.run_nbody(o, dx.array.cpu().numpy(), p.array.cpu().numpy(), a0, 1.0, 0.09, ts, solver="kdk",
                  on_save=lambda i, t, x, v: snaps.__setitem__(i, orc.cic_paint(x, (n, n, n), dtype=np.float64)))
    assert rel(out.array[0].cpu().numpy(), snaps[1]) < 1e-5
    assert rel(out.array[1].cpu().numpy(), snaps[0]) < 1e-5
    kedges = np.geomspace(2 * np.pi / box, np.pi * n / box, 33)
    k, pk = pkg.power(out.array[0], box_shape=(box,) * 3, kedges=kedges)
    k_ref, pk_ref = orc.power(snaps[1], (box,) * 3, kedges)
    sel = (k_ref < 0.5 * np.pi * n / box) & (pk_ref > 0)
    assert np.all(np.abs(pk[sel] / pk_ref[sel] - 1) < 1e-3)


def test_spherical_lightcone_matches_oracle(setup):
    pkg, c, o = setup
    n, box, a0, nside = 32, 400.0, 0.1, 16
    ic = pkg.gaussian_initial_conditions(31, (n, n, n), (box,) * 3, cosmo=c, nside=nside)
    dx, p = pkg.lpt(c, ic, ts=a0)
    painting = pkg.PaintingOptions(target="spherical", scheme="bilinear")
    solver = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=painting))
    # 4 shells: with nb_shells=3 the reference's int(R // (R / 3)) rounds down to 2 shells
    lc = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=0.1, nb_shells=4, solver=solver)
    assert tuple(lc.array.shape) == (4, 12 * nside**2)
    r_far_to_near, a_c = pkg.compute_lightcone_shells(c, dx, nb_shells=4)
    widths = pkg.distances(r_far_to_near, 0.0)
    maps = {}

    def save(i, t, x, v):
        maps[i] = orc.paint_healpix(x.astype(np.float32), (n, n, n), (box,) * 3, (box / 2,) * 3,
                                    r_far_to_near[i] - widths[i] / 2, r_far_to_near[i] + widths[i] / 2, nside)

    nbo.run_nbody(o, dx.array.cpu().numpy(), p.array.cpu().numpy(), a0, 1.0, 0.1, a_c.astype(np.float32), on_save=save)
    got = lc.array.cpu().numpy()[::-1]  # back to far -> near
    for i in range(4):
        assert rel(got[i], maps[i]) < 1e-4, i


def test_kkd_step_then_reverse_roundtrip(setup):
    pkg, c, o = setup
    n, box, a0 = 32, 256.0, 0.2
    ic = _ic(pkg, c, n, box, 41)
    dx, p = pkg.lpt(c, ic, ts=a0)
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="density")).update_geometry(
        ts=np.array([1.0], np.float32), r_centers=np.zeros(1), density_widths=np.ones(1), max_comoving_distance=1.0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    d0, p0, st = solver.init(dx, p, a0, a0 + 0.01, 0.01, c)
    d1, p1, st1 = solver.step(d0, p0, a0, a0 + 0.01, 0.01, st, c)
    d2, p2, _ = solver.reverse(d1, p1, a0, a0 + 0.01, 0.01, st1, c)
    assert rel(d2.array.cpu().numpy(), d0.array.cpu().numpy()) < 1e-6
    assert rel(p2.array.cpu().numpy(), p0.array.cpu().numpy()) < 1e-5


def test_standalone_steps_equal_fused_integrate(setup):
    pkg, c, o = setup
    n, box, a0 = 32, 256.0, 0.1
    ic = _ic(pkg, c, n, box, 51)
    dx, p = pkg.lpt(c, ic, ts=a0)
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="particles"))
    solver = pkg.EfficientDriftDoubleKick(interp_kernel=interp)
    fused = pkg.nbody(c, dx, p, t0=a0, t1=0.5, dt0=0.1, ts=np.array([0.5]), solver=solver).array
    d, v, st = solver.init(dx, p, a0, a0 + 0.1, 0.1, c)
    for _, ta, tb in [e for e in pkg.pm.time_grid(a0, [0.5], 0.1, 0.5) if e[0] == "step"]:
        d, v, st = solver.step(d, v, ta, tb, 0.1, st, c)
    assert rel(d.array.cpu().numpy(), fused.cpu().numpy()) < 1e-6
