"""Drift to the lightcone on the GPU against the CPU oracle: the HEALPix shell paint with every
particle (or every rotated image of it) drifted from the snapshot time to the time of its own
distance from the observer (_drift_to_time, src/jax_fli/pm/interp.py:110-121, used by DriftInterp
and OnionTiler with drift_on_lightcone), and both interpolation kernels inside a full N-body run.

The device evaluates the drift factor from float64 tables (4096 points, linear interpolation) built
by the package's own cosmology; the oracle uses its own distance quadrature and growth ODE through
cubic splines. Tolerances: a single paint 1e-5 relative L2; lightcone shells of an N-body run 1e-4
(float32 particle arithmetic over the run against float64)."""

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


def test_gradient_through_the_drift_raises(pkg, cuda):
    N, BOX = 8, 100.0
    c = pkg.Cosmology(**P)
    noise = torch.from_numpy(np.random.default_rng(44).standard_normal((N,) * 3).astype(np.float32)).to(cuda)
    noise.requires_grad_(True)
    ic = pkg.interpolate_initial_conditions(noise, (N,) * 3, (BOX,) * 3, cosmo=c, nside=4)
    dx, p = pkg.lpt(c, ic, ts=0.1)
    ts = np.array([[0.995], [0.98]])
    painting = pkg.PaintingOptions(target="spherical", scheme="bilinear")
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.DriftInterp(painting=painting))
    with pytest.raises(NotImplementedError):
        pkg.nbody(c, dx, p, t0=0.1, t1=1.0, dt0=0.15, ts=ts, solver=solver, adjoint="reverse")
