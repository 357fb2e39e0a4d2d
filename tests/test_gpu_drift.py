"""Drift to the lightcone on the GPU against the CPU oracle: the HEALPix shell paint with every
particle (or every rotated image of it) drifted from the snapshot time to the time of its own
distance from the observer (_drift_to_time, src/jax_fli/pm/interp.py:110-121, used by DriftInterp
and OnionTiler with drift_on_lightcone), its adjoint in positions and velocities, both interpolation
kernels inside a full N-body run, and the reverse-adjoint gradient through drifted shells.

The device evaluates the drift factor from float64 tables (4096 points, linear interpolation) built
by the package's own cosmology; the oracle uses its own distance quadrature and growth ODE through
cubic splines. Tolerances: a single paint 1e-5 relative L2; lightcone shells of an N-body run 1e-4
(float32 particle arithmetic over the run against float64); the paint adjoint 1e-5 (velocities) and
1e-4 (positions: the slope of linear tables against spline derivatives); the N-body gradient 2e-3
against float64 central differences."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import nbody_oracle as nbo
from oracle import sky_oracle as so

pytestmark = pytest.mark.gpu

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("tiled", [False, True])
@pytest.mark.parametrize("scheme", ["ngp", "bilinear"])
def test_drifted_paint_matches_oracle(pkg, cuda, scheme, tiled):
    mesh, box, obs = (16, 16, 16), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0)
    rng = np.random.default_rng(41)
    dx = (rng.standard_normal(mesh + (3,)) * 1.2).astype(np.float32)
    vel = (rng.standard_normal(mesh + (3,)) * 30.0).astype(np.float32)
    rmin, rmax, a_now = (40.0, 90.0, 0.9) if tiled else (10.0, 45.0, 0.95)
    tiles = None
    if tiled:
        rots = np.random.default_rng(5).integers(1, 48, size=27).astype(np.int32)
        tiles = (rots, so.onion_shifts((0.5, 0.5, 0.5), box))
    c = pkg.Cosmology(**P)
    tab = pkg.pm.interp.drift_tables(c, a_now, 2.0 * rmax + 100.0, cuda)
    got = pkg.engine.paint_healpix_drift(torch.from_numpy(dx).to(cuda), torch.from_numpy(vel).to(cuda), mesh, box,
                                         obs, rmin, rmax, 8, tab, scheme=scheme, tiles=tiles).cpu().numpy()
    drift = so.Drift(nbo.Cosmo(**P), a_now)
    ref = so.paint_healpix_tiled(dx, mesh, box, obs, rmin, rmax, 8, scheme=scheme, vel=vel, drift=drift,
                                 tiles=None if tiles is None else list(zip(*tiles)))
    undrifted = so.paint_healpix_tiled(dx, mesh, box, obs, rmin, rmax, 8, scheme=scheme,
                                       tiles=None if tiles is None else list(zip(*tiles)))
    assert rel(ref, undrifted) > 1e-2  # the drift moves particles across pixels
    assert rel(got, ref) < 1e-5


def test_drift_vanishes_at_the_snapshot_distance(pkg, cuda):
    """A shell at the comoving distance of the snapshot time: a particle there does not move."""
    c = pkg.Cosmology(**P)
    a_now = 0.97
    r0 = float(np.asarray(pkg.cosmology.radial_comoving_distance(c, a_now)).reshape(-1)[0])
    tab = pkg.pm.interp.drift_tables(c, a_now, 400.0, cuda)
    mesh, box = (8, 8, 8), (800.0, 800.0, 800.0)
    u = np.array([0.6, 0.48, 0.64])  # a unit vector off pixel boundaries
    obs = tuple(400.0 - r0 * u)  # the lattice point (4, 4, 4) of cell 100 sits at distance r0
    dx = np.zeros(mesh + (3,), np.float32)
    vel = np.zeros(mesh + (3,), np.float32)
    vel[4, 4, 4] = (0.0, 500.0, 500.0)
    w = np.zeros(mesh, np.float32)
    w[4, 4, 4] = 1.0
    args = (torch.from_numpy(dx).to(cuda), torch.from_numpy(vel).to(cuda), mesh, box, obs, r0 - 5.0, r0 + 5.0, 4, tab)
    a = pkg.engine.paint_healpix_drift(*args, scheme="ngp", weights=torch.from_numpy(w).to(cuda)).cpu().numpy()
    b = pkg.engine.paint_healpix(args[0], mesh, box, obs, r0 - 5.0, r0 + 5.0, 4, scheme="ngp",
                                 weights=torch.from_numpy(w).to(cuda)).cpu().numpy()
    assert a.sum() > 0 and np.array_equal(a, b)


@pytest.mark.parametrize("kernel", ["drift", "onion"])
def test_nbody_lightcone_with_drift_matches_oracle(pkg, cuda, kernel):
    """A reversible N-body run that saves HEALPix shells through DriftInterp (shells inside the box)
    or OnionTiler with drift_on_lightcone (two of three shells from 27 rotated images), against the
    oracle's run painting its own snapshots with the same drift."""
    N, BOX, A0, DT0, NSIDE = 16, 100.0, 0.1, 0.15, 8
    KT = np.logspace(-4, 1, 128)
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    edges = np.array([10.0, 25.0, 45.0]) if kernel == "drift" else np.array([20.0, 45.0, 90.0, 130.0])
    r_c = 0.5 * (edges[1:] + edges[:-1])[::-1]
    widths = (edges[1:] - edges[:-1])[::-1]
    ts = np.stack([pkg.cosmology.a_of_chi(c, r_c - widths / 2), pkg.cosmology.a_of_chi(c, r_c + widths / 2)])
    painting = pkg.PaintingOptions(target="spherical", scheme="bilinear")
    if kernel == "drift":
        interp = pkg.DriftInterp(painting=painting)
        rots = None
    else:
        interp = pkg.OnionTiler(painting=painting, drift_on_lightcone=True)
        rots = interp.init().rotations
    noise = np.random.default_rng(43).standard_normal((N,) * 3).astype(np.float32)
    ic = pkg.interpolate_initial_conditions(torch.from_numpy(noise).to(cuda), (N,) * 3, (BOX,) * 3, cosmo=c,
                                            nside=NSIDE)
    dx, p = pkg.lpt(c, ic, ts=A0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    lc = pkg.nbody(c, dx, p, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver)
    got = lc.array.reshape(len(r_c), -1).double().cpu().numpy()

    t_snap = np.asarray(lc.scale_factors, dtype=np.float64)
    r_snap = np.asarray(lc.comoving_centers, dtype=np.float64)
    w_snap = np.asarray(lc.density_width, dtype=np.float64)
    order = np.argsort(t_snap)
    ic0 = nbo.color_field(noise.astype(np.float64), (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
    x0, v0 = nbo.lpt1(o, ic0, A0)
    maps = {}

    def save(i, t, x, v):
        s = order[i]
        rc, wd = r_snap[s], w_snap[s]
        tiles = None
        if rots is not None and rc + wd / 2 > lc.max_comoving_radius:
            tiles = list(zip(rots, so.onion_shifts((0.5, 0.5, 0.5), (BOX,) * 3)))
        maps[s] = so.paint_healpix_tiled(x, (N,) * 3, (BOX,) * 3, (BOX / 2,) * 3, rc - wd / 2, rc + wd / 2, NSIDE,
                                         tiles=tiles, vel=v, drift=so.Drift(o, float(t)))

    nbo.run_nbody(o, x0, v0, A0, 1.0, DT0, np.sort(t_snap).astype(np.float32), solver="kkd", on_save=save)
    for s in range(len(t_snap)):
        assert rel(got[s], maps[s]) < 1e-4, s


@pytest.mark.parametrize("tiled", [False, True])
def test_drifted_paint_vjp_matches_oracle(pkg, cuda, tiled):
    mesh, box, obs = (16, 16, 16), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0)
    rng = np.random.default_rng(45)
    dx = (rng.standard_normal(mesh + (3,)) * 1.2).astype(np.float32)
    vel = (rng.standard_normal(mesh + (3,)) * 30.0).astype(np.float32)
    rmin, rmax, a_now = (40.0, 90.0, 0.9) if tiled else (10.0, 45.0, 0.95)
    tiles = None
    if tiled:
        rots = np.random.default_rng(5).integers(1, 48, size=27).astype(np.int32)
        tiles = (rots, so.onion_shifts((0.5, 0.5, 0.5), box))
    c = pkg.Cosmology(**P)
    tab = pkg.pm.interp.drift_tables(c, a_now, 2.0 * rmax + 100.0, cuda)
    mb = rng.standard_normal(12 * 8 * 8).astype(np.float32)
    gx, gv = pkg.engine.paint_healpix_drift_vjp(torch.from_numpy(dx).to(cuda), torch.from_numpy(vel).to(cuda),
                                                torch.from_numpy(mb).to(cuda), mesh, box, obs, rmin, rmax, 8, tab,
                                                tiles=tiles)
    rx, rv = so.paint_healpix_vjp(dx, mb, mesh, box, obs, rmin, rmax, 8, tiles=None if tiles is None else
                                  list(zip(*tiles)), vel=vel, drift=so.Drift(nbo.Cosmo(**P), a_now))
    # the velocity cotangent carries f itself (tables agree to 1e-9); the position cotangent also
    # f' = df/dr, the slope of the device's linear tables against the oracle's spline derivative
    assert rel(gv.cpu().numpy(), rv) < 1e-5
    assert rel(gx.cpu().numpy(), rx) < 1e-4


def test_nbody_gradient_through_the_drifted_lightcone(pkg, cuda):
    """d L / d (dx, p) for L = sum_s <W_s, shell_s> of OnionTiler(drift_on_lightcone=True) shells
    (one inside the box, one from 27 rotated images) of a reversible N-body run, against float64
    central differences of the oracle's run painting its drifted snapshots."""
    N, BOX, A0, DT0, NSIDE = 16, 100.0, 0.1, 0.15, 8
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    edges = np.array([20.0, 45.0, 90.0])
    r_c = 0.5 * (edges[1:] + edges[:-1])[::-1]
    widths = (edges[1:] - edges[:-1])[::-1]
    ts = np.stack([pkg.cosmology.a_of_chi(c, r_c - widths / 2), pkg.cosmology.a_of_chi(c, r_c + widths / 2)])
    interp = pkg.OnionTiler(painting=pkg.PaintingOptions(target="spherical", scheme="bilinear"),
                            drift_on_lightcone=True)
    rots = interp.init().rotations
    ic = pkg.gaussian_initial_conditions(11, (N,) * 3, (BOX,) * 3, cosmo=c, nside=NSIDE)
    dx, p = pkg.lpt(c, ic, ts=A0)
    x0, v0 = dx.array.clone(), p.array.clone()
    dxf = dx.replace(array=x0.clone().requires_grad_(True))
    pf = p.replace(array=v0.clone().requires_grad_(True))
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    lc = pkg.nbody(c, dxf, pf, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver, adjoint="reverse")
    rng = np.random.default_rng(46)
    Wt = rng.standard_normal(tuple(lc.array.shape))
    (lc.array.double() * torch.from_numpy(Wt).to(cuda)).sum().backward()
    t_snap = np.asarray(lc.scale_factors, dtype=np.float64)
    r_snap = np.asarray(lc.comoving_centers, dtype=np.float64)
    w_snap = np.asarray(lc.density_width, dtype=np.float64)
    order = np.argsort(t_snap)
    max_r = lc.max_comoving_radius

    def loss(xx, vv):
        maps = {}

        def save(i, t, x, v):
            s = order[i]
            rc, wd = r_snap[s], w_snap[s]
            tiles = None if rc + wd / 2 <= max_r else list(zip(rots, so.onion_shifts((0.5, 0.5, 0.5), (BOX,) * 3)))
            maps[s] = so.paint_healpix_tiled(x, (N,) * 3, (BOX,) * 3, (BOX / 2,) * 3, rc - wd / 2, rc + wd / 2, NSIDE,
                                             tiles=tiles, vel=v, drift=so.Drift(o, float(t)))

        nbo.run_nbody(o, xx, vv, A0, 1.0, DT0, np.sort(t_snap).astype(np.float32), solver="kkd", on_save=save)
        return float(sum((maps[s] * Wt[s]).sum() for s in range(len(t_snap))))

    x0n, v0n = x0.cpu().numpy().astype(np.float64), v0.cpu().numpy().astype(np.float64)
    eps = 1e-6
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
def test_cosmology_gradient_through_the_drifted_lightcone(pkg, cuda, adjoint):
    """d L / d Omega_c through white noise -> ICs -> LPT -> N-body with OnionTiler drift shells (the
    drift factor's cosmology included), against float64 central differences of the oracle's forward
    model at fixed shell geometry."""
    N, BOX, A0, DT0, NSIDE = 16, 100.0, 0.1, 0.15, 8
    KT = np.logspace(-4, 1, 128)
    c0 = pkg.Cosmology(**P)
    edges = np.array([20.0, 45.0, 90.0])
    r_c = 0.5 * (edges[1:] + edges[:-1])[::-1]
    widths = (edges[1:] - edges[:-1])[::-1]
    ts = np.stack([pkg.cosmology.a_of_chi(c0, r_c - widths / 2), pkg.cosmology.a_of_chi(c0, r_c + widths / 2)])
    interp = pkg.OnionTiler(painting=pkg.PaintingOptions(target="spherical", scheme="bilinear"),
                            drift_on_lightcone=True)
    rots = interp.init().rotations
    rng = np.random.default_rng(47)
    noise = rng.standard_normal((N,) * 3).astype(np.float32)
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    c = pkg.Cosmology(**{**P, "Omega_c": Oc})
    ic = pkg.interpolate_initial_conditions(torch.from_numpy(noise).to(cuda), (N,) * 3, (BOX,) * 3, cosmo=c,
                                            nside=NSIDE)
    dx, p = pkg.lpt(c, ic, ts=A0)
    lc = pkg.nbody(c, dx, p, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=pkg.ReversibleDoubleKickDrift(interp_kernel=interp),
                   adjoint=adjoint)
    Wt = rng.standard_normal(tuple(lc.array.shape))
    L = (lc.array.double() * torch.from_numpy(Wt).to(cuda)).sum()
    L.backward()
    t_snap = np.asarray(lc.scale_factors, dtype=np.float64)
    r_snap = np.asarray(lc.comoving_centers, dtype=np.float64)
    w_snap = np.asarray(lc.density_width, dtype=np.float64)
    order = np.argsort(t_snap)
    max_r = lc.max_comoving_radius

    def oracle_loss(oc):
        o = nbo.Cosmo(**{**P, "Omega_c": oc})
        ic0 = nbo.color_field(noise.astype(np.float64), (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
        x0, v0 = nbo.lpt1(o, ic0, A0)
        maps = {}

        def save(i, t, x, v):
            s = order[i]
            rc, wd = r_snap[s], w_snap[s]
            tiles = None if rc + wd / 2 <= max_r else list(zip(rots, so.onion_shifts((0.5, 0.5, 0.5), (BOX,) * 3)))
            maps[s] = so.paint_healpix_tiled(x, (N,) * 3, (BOX,) * 3, (BOX / 2,) * 3, rc - wd / 2, rc + wd / 2, NSIDE,
                                             tiles=tiles, vel=v, drift=so.Drift(o, float(t)))

        nbo.run_nbody(o, x0, v0, A0, 1.0, DT0, np.sort(t_snap).astype(np.float32), solver="kkd", on_save=save)
        return float(sum((maps[s] * Wt[s]).sum() for s in range(len(t_snap))))

    scale = float((lc.array.detach().double().abs() * torch.from_numpy(np.abs(Wt)).to(cuda)).sum())
    assert abs(float(L) - oracle_loss(P["Omega_c"])) / scale < 1e-4  # the random weights cancel in L
    h = 1e-6
    fd = (oracle_loss(P["Omega_c"] + h) - oracle_loss(P["Omega_c"] - h)) / (2 * h)
    print(f"{adjoint}: d/dOmega_c adjoint {float(Oc.grad):.6e} fd {fd:.6e}")
    assert abs(float(Oc.grad) - fd) / abs(fd) < 2e-3
