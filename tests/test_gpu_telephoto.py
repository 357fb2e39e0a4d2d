"""Flat-sky slab painting (ParticleField.paint_2d, src/jax_fli/fields/particles.py:234-322 and
_single_paint_2d, src/jax_fli/_src/fields/_painting.py:117-202) and TelephotoInterp (src/jax_fli/pm/
interp.py:507-699) on the GPU against the CPU oracle.

pmx_image_positions (one rotated, shifted and optionally drifted image of the particles as absolute
grid positions) against sky_oracle.image_positions; pmx_paint_flat against sky_oracle.paint_flat in
both position modes, on a particle sub-lattice, with per-particle weights and a flat grid finer than
the mesh; and full N-body runs that save flat (with and without the drift) and HEALPix shells through
TelephotoInterp, and through DriftInterp with flat, density and particle targets, against the oracle's
own run painting its snapshots the same way.

Tolerances: image positions 1e-5 grid units per 16 of coordinate size without the drift and 1e-4 with it (float64 tables
with linear interpolation against the oracle's splines, times velocities of tens of grid units); a
single flat paint 1e-6 relative L2 (float32 atomics); lightcone maps of an N-body run 1e-4 (float32
particle arithmetic over the run against float64)."""

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


@pytest.mark.parametrize("drift", [False, True])
@pytest.mark.parametrize("mode,npart", [("relative", 16), ("relative", 8), ("absolute", 12)])
def test_image_positions_match_oracle(pkg, cuda, mode, npart, drift):
    mesh, box, obs = (16, 16, 16), (100.0, 100.0, 100.0), (50.0, 50.0, 0.0)
    rng = np.random.default_rng(7 + npart)
    if mode == "relative":
        pos = (rng.standard_normal((npart,) * 3 + (3,)) * 1.5).astype(np.float32)
    else:
        pos = (rng.uniform(-1.0, 17.0, (npart,) * 3 + (3,))).astype(np.float32)
    vel = (rng.standard_normal(pos.shape) * 20.0).astype(np.float32)
    c = pkg.Cosmology(**P)
    a_now = 0.8
    for code, shift in ((0, (0.0, 0.0, 0.0)), (7, (0.0, 0.0, 150.0)), (46, (0.0, 0.0, 300.0))):
        tab = pkg.pm.interp.drift_tables(c, a_now, 600.0, cuda) if drift else None
        got = pkg.engine.image_positions(torch.from_numpy(pos).to(cuda), mesh, box, obs, rot=code, shift=shift,
                                         mode=mode, vel=torch.from_numpy(vel).to(cuda) if drift else None,
                                         drift=tab).cpu().numpy().reshape(-1, 3)
        ref = so.image_positions(pos, mesh, box, obs, code=code, shift=shift, mode=mode,
                                 vel=vel if drift else None, drift=so.Drift(nbo.Cosmo(**P), a_now) if drift else None)
        assert np.abs(got - ref).max() < (1e-4 if drift else 1e-5 * max(1.0, np.abs(ref).max() / 16)), code
        if drift:
            still = so.image_positions(pos, mesh, box, obs, code=code, shift=shift, mode=mode)
            assert np.abs(ref - still).max() > 1e-2  # the drift moves the particles


@pytest.mark.parametrize("mode,npart,flat", [("relative", 16, (16, 16)), ("relative", 8, (32, 24)),
                                             ("absolute", 12, (20, 16))])
def test_paint_flat_matches_oracle(pkg, cuda, mode, npart, flat):
    mesh = (16, 16, 16)
    rng = np.random.default_rng(11 + npart)
    if mode == "relative":
        pos = (rng.standard_normal((npart,) * 3 + (3,)) * 1.5).astype(np.float32)
    else:
        pos = (rng.uniform(-3.0, 19.0, (npart,) * 3 + (3,))).astype(np.float32)
    w = rng.uniform(0.5, 1.5, (npart,) * 3).astype(np.float32)
    for zlo, zhi, weights in ((-1e6, 1e6, 1.0), (3.25, 9.5, w), (12.0, 20.0, 2.0)):
        wt = torch.from_numpy(w).to(cuda) if isinstance(weights, np.ndarray) else weights
        got = pkg.engine.paint_flat(torch.from_numpy(pos).to(cuda), mesh, flat, zlo, zhi, 0.125, mode=mode,
                                    weights=wt).cpu().numpy()
        ref = so.paint_flat(pos, mesh, flat, zlo, zhi, 0.125, mode=mode, weight=weights)
        assert ref.sum() > 0
        assert rel(got, ref) < 1e-6, (zlo, zhi)


def test_paint_2d_field_matches_oracle(pkg, cuda):
    """ParticleField.paint_2d: the slab (c - w/2, c + w/2] in grid units of z and 1 / pixel volume."""
    N, BOX = 16, 200.0
    noise = np.random.default_rng(3).standard_normal((N,) * 3).astype(np.float32)
    c = pkg.Cosmology(**P)
    ic = pkg.interpolate_initial_conditions(torch.from_numpy(noise).to(cuda), (N,) * 3, (BOX,) * 3, cosmo=c,
                                            flatsky_npix=(24, 20))
    dx, _ = pkg.lpt(c, ic, ts=0.5)
    center, width = 80.0, 30.0
    got = dx.paint_2d(center=center, density_plane_width=width).array.cpu().numpy()
    cg, wg = center * N / BOX, width * N / BOX
    scale = 1.0 / ((BOX / 24) * (BOX / 20) * width)
    ref = so.paint_flat(dx.array.cpu().numpy(), (N,) * 3, (24, 20), np.float32(cg - wg / 2), np.float32(cg + wg / 2),
                        scale, mode="relative")
    assert rel(got, ref) < 1e-6


@pytest.mark.parametrize("target,drift", [("flat", False), ("flat", True), ("spherical", True)])
def test_nbody_telephoto_matches_oracle(pkg, cuda, target, drift):
    """A reversible N-body run saving three shells through TelephotoInterp (two beyond the box
    radius, painted from rotated images shifted along z; one inside, from the box), against the
    oracle's run painting its own snapshots the same way."""
    N, BOX, A0, DT0, NSIDE, FLAT = 16, 100.0, 0.1, 0.15, 8, (16, 16)
    KT = np.logspace(-4, 1, 128)
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    edges = np.array([20.0, 40.0, 75.0, 120.0])
    r_c = 0.5 * (edges[1:] + edges[:-1])[::-1]
    widths = (edges[1:] - edges[:-1])[::-1]
    ts = np.stack([pkg.cosmology.a_of_chi(c, r_c - widths / 2), pkg.cosmology.a_of_chi(c, r_c + widths / 2)])
    painting = pkg.PaintingOptions(target=target, scheme="bilinear")
    interp = pkg.TelephotoInterp(painting=painting, drift_on_lightcone=drift)
    noise = np.random.default_rng(47).standard_normal((N,) * 3).astype(np.float32)
    ic = pkg.interpolate_initial_conditions(torch.from_numpy(noise).to(cuda), (N,) * 3, (BOX,) * 3, cosmo=c,
                                            nside=NSIDE, flatsky_npix=FLAT)
    dx, p = pkg.lpt(c, ic, ts=A0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    with pytest.warns(UserWarning) if target == "spherical" else _nullcontext():
        lc = pkg.nbody(c, dx, p, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver)
    got = lc.array.reshape(len(r_c), -1).double().cpu().numpy()

    radius = dx.max_comoving_radius
    t_snap = np.asarray(lc.scale_factors, dtype=np.float64)
    r_snap = np.asarray(lc.comoving_centers, dtype=np.float64)
    w_snap = np.asarray(lc.density_width, dtype=np.float64)
    order = np.argsort(t_snap)
    far_first = np.sort(r_snap)[::-1]
    assert (r_snap > radius).sum() == 2
    ic0 = nbo.color_field(noise.astype(np.float64), (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
    x0, v0 = nbo.lpt1(o, ic0, A0)
    obs = (BOX / 2,) * 3
    maps = {}

    def save(i, t, x, v):
        s = order[i]
        rc, wd = r_snap[s], w_snap[s]
        code, shift = 0, np.zeros(3)
        if rc > radius:
            code = int((np.sum(far_first[:list(far_first).index(rc) + 1] > radius) - 1) % 47) + 1
            shift = np.array([0.0, 0.0, np.floor(rc / radius) * radius])
        dr = so.Drift(o, float(t)) if drift else None
        if target == "flat":
            g = so.image_positions(x, (N,) * 3, (BOX,) * 3, obs, code=code, shift=shift, vel=v, drift=dr)
            cg, wg = rc * N / BOX, wd * N / BOX
            maps[s] = so.paint_flat(g.reshape((N,) * 3 + (3,)), (N,) * 3, FLAT, np.float32(cg - wg / 2),
                                    np.float32(cg + wg / 2), 1.0 / ((BOX / FLAT[0]) * (BOX / FLAT[1]) * wd),
                                    mode="absolute").reshape(-1)
        else:
            tiles = None if rc <= radius else [(code, shift)]
            maps[s] = so.paint_healpix_tiled(x, (N,) * 3, (BOX,) * 3, obs, rc - wd / 2, rc + wd / 2, NSIDE,
                                             tiles=tiles, vel=v, drift=dr)

    nbo.run_nbody(o, x0, v0, A0, 1.0, DT0, np.sort(t_snap).astype(np.float32), solver="kkd", on_save=save)
    for s in range(len(t_snap)):
        assert np.abs(maps[s]).sum() > 0, s
        assert rel(got[s], maps[s]) < 1e-4, s


@pytest.mark.parametrize("target", ["flat", "density", "particles"])
def test_nbody_drift_interp_targets_match_oracle(pkg, cuda, target):
    """DriftInterp (src/jax_fli/pm/interp.py:212-320) with flat, density and particle targets: every
    particle drifted to its own lightcone time, then the target's paint (or the positions in Mpc/h),
    against the oracle's run drifting and painting its own snapshots."""
    from oracle import pm_oracle as orc

    N, BOX, A0, DT0, FLAT = 16, 100.0, 0.1, 0.15, (16, 16)
    KT = np.logspace(-4, 1, 128)
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    edges = np.array([10.0, 25.0, 45.0])
    r_c = 0.5 * (edges[1:] + edges[:-1])[::-1]
    widths = (edges[1:] - edges[:-1])[::-1]
    ts = np.stack([pkg.cosmology.a_of_chi(c, r_c - widths / 2), pkg.cosmology.a_of_chi(c, r_c + widths / 2)])
    interp = pkg.DriftInterp(painting=pkg.PaintingOptions(target=target))
    noise = np.random.default_rng(53).standard_normal((N,) * 3).astype(np.float32)
    ic = pkg.interpolate_initial_conditions(torch.from_numpy(noise).to(cuda), (N,) * 3, (BOX,) * 3, cosmo=c,
                                            flatsky_npix=FLAT)
    dx, p = pkg.lpt(c, ic, ts=A0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    lc = pkg.nbody(c, dx, p, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver)
    got = lc.array.reshape(len(r_c), -1).double().cpu().numpy()

    t_snap = np.asarray(lc.scale_factors, dtype=np.float64).reshape(-1)
    r_snap = np.asarray(lc.comoving_centers, dtype=np.float64).reshape(-1)
    w_snap = np.asarray(lc.density_width, dtype=np.float64).reshape(-1)
    order = np.argsort(t_snap)
    ic0 = nbo.color_field(noise.astype(np.float64), (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
    x0, v0 = nbo.lpt1(o, ic0, A0)
    obs = (BOX / 2,) * 3
    out = {}

    def save(i, t, x, v):
        s = order[i]
        rc, wd = r_snap[s], w_snap[s]
        g = so.image_positions(x, (N,) * 3, (BOX,) * 3, obs, vel=v, drift=so.Drift(o, float(t)))
        if target == "flat":
            cg, wg = rc * N / BOX, wd * N / BOX
            out[s] = so.paint_flat(g.reshape((N,) * 3 + (3,)), (N,) * 3, FLAT, np.float32(cg - wg / 2),
                                   np.float32(cg + wg / 2), 1.0 / ((BOX / FLAT[0]) * (BOX / FLAT[1]) * wd),
                                   mode="absolute").reshape(-1)
        elif target == "density":
            out[s] = orc.cic_paint(g.reshape((N,) * 3 + (3,)), (N,) * 3, mode="absolute",
                                   dtype=np.float64).reshape(-1)
        else:
            out[s] = (g * (BOX / N)).reshape(-1)

    nbo.run_nbody(o, x0, v0, A0, 1.0, DT0, np.sort(t_snap).astype(np.float32), solver="kkd", on_save=save)
    for s in range(len(t_snap)):
        assert np.abs(out[s]).sum() > 0, s
        assert rel(got[s], out[s]) < 1e-4, s
    if target == "particles":
        assert lc.unit == pkg.PositionUnit.MPC_H


class _nullcontext:
    def __enter__(self):
        return None

    def __exit__(self, *a):
        return False


@pytest.mark.parametrize("mode,npart,flat", [("relative", 16, (16, 16)), ("relative", 8, (32, 24)),
                                             ("absolute", 12, (20, 16))])
def test_paint_flat_vjp_matches_oracle(pkg, cuda, mode, npart, flat):
    mesh = (16, 16, 16)
    rng = np.random.default_rng(21 + npart)
    if mode == "relative":
        pos = (rng.standard_normal((npart,) * 3 + (3,)) * 1.5).astype(np.float32)
    else:
        pos = (rng.uniform(-3.0, 19.0, (npart,) * 3 + (3,))).astype(np.float32)
    w = rng.uniform(0.5, 1.5, (npart,) * 3).astype(np.float32)
    mb = rng.standard_normal(flat).astype(np.float32)
    base = rng.standard_normal(pos.shape).astype(np.float32)
    acc = torch.from_numpy(base).to(cuda)
    pkg.engine.paint_flat_vjp(torch.from_numpy(pos).to(cuda), torch.from_numpy(mb).to(cuda), mesh, 3.25, 12.5, 0.125,
                              mode=mode, weights=torch.from_numpy(w).to(cuda), pos_bar=acc, accumulate=True)
    ref = so.paint_flat_vjp(pos, mb, mesh, 3.25, 12.5, 0.125, mode=mode, weight=w).reshape(pos.shape)
    assert np.abs(ref).sum() > 0
    assert rel(acc.cpu().numpy() - base, ref) < 1e-5


@pytest.mark.parametrize("adjoint", ["reverse", "checkpointed"])
def test_nbody_gradient_through_a_flat_snapshot(pkg, cuda, adjoint):
    """d <W, flat map> / d (dx, p) for a flat slab painted at the end of a KKD run (NoInterp, target
    flat), through both adjoints, against float64 central differences of the oracle's run and flat paint."""
    N, BOX, A0, DT0, FLAT = 16, 100.0, 0.1, 0.2, (16, 16)
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    ic = pkg.gaussian_initial_conditions(9, (N,) * 3, (BOX,) * 3, cosmo=c, flatsky_npix=FLAT)
    dx, p = pkg.lpt(c, ic, ts=A0)
    x0, v0 = dx.array.clone(), p.array.clone()
    rng = np.random.default_rng(63)
    W = rng.standard_normal(FLAT)
    dxf = dx.replace(array=x0.clone().requires_grad_(True))
    pf = p.replace(array=v0.clone().requires_grad_(True))
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="flat"))
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    center, width = 50.0, 40.0
    ts = np.stack([pkg.cosmology.a_of_chi(c, np.array([center - width / 2])),
                   pkg.cosmology.a_of_chi(c, np.array([center + width / 2]))])
    out = pkg.nbody(c, dxf, pf, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver, adjoint=adjoint)
    r_c = float(np.asarray(out.comoving_centers).reshape(-1)[0])
    w_c = float(np.asarray(out.density_width).reshape(-1)[0])
    t_snap = float(np.asarray(out.scale_factors, dtype=np.float64).reshape(-1)[0])
    (out.array.double().reshape(FLAT) * torch.from_numpy(W).to(cuda)).sum().backward()
    zlo, zhi, scale = pkg.fields.flat_slab(dx, r_c, w_c)
    x0n, v0n = x0.cpu().numpy().astype(np.float64), v0.cpu().numpy().astype(np.float64)

    def loss(xx, vv):
        snap = {}

        def save(i, t, x, v):
            snap["x"] = x

        nbo.run_nbody(o, xx, vv, A0, 1.0, DT0, np.array([t_snap], np.float32), solver="kkd", on_save=save)
        return float((so.paint_flat(snap["x"], (N,) * 3, FLAT, zlo, zhi, scale, mode="relative") * W).sum())

    assert loss(x0n, v0n) != 0.0
    eps = 1e-6
    for which, grad in (("x", dxf.array.grad), ("v", pf.array.grad)):
        d = rng.standard_normal(x0n.shape)
        if which == "x":
            fd = (loss(x0n + eps * d, v0n) - loss(x0n - eps * d, v0n)) / (2 * eps)
        else:
            fd = (loss(x0n, v0n + eps * d) - loss(x0n, v0n - eps * d)) / (2 * eps)
        got = float((grad.double().cpu().numpy() * d).sum())
        print(f"{adjoint} d/d{which}: adjoint {got:.6e} fd {fd:.6e}")
        assert abs(got - fd) / abs(fd) < 2e-3


@pytest.mark.parametrize("drift", [False, True])
def test_image_positions_vjp_matches_differences(pkg, cuda, drift):
    """pmx_image_positions_vjp against float64 central differences of the oracle's image transform
    (positions, and velocities through the drift)."""
    mesh, box, obs = (16, 16, 16), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0)
    rng = np.random.default_rng(71)
    pos = (rng.standard_normal((8, 8, 8, 3)) * 1.5).astype(np.float32)
    vel = (rng.standard_normal(pos.shape) * 20.0).astype(np.float32)
    gbar = rng.standard_normal(pos.shape).astype(np.float32)
    c = pkg.Cosmology(**P)
    a_now, code, shift = 0.8, 46, (0.0, 0.0, 150.0)
    tab = pkg.pm.interp.drift_tables(c, a_now, 600.0, cuda) if drift else None
    gx, gv = pkg.engine.image_positions_vjp(torch.from_numpy(pos).to(cuda), torch.from_numpy(gbar).to(cuda), mesh, box,
                                            obs, rot=code, shift=shift,
                                            vel=torch.from_numpy(vel).to(cuda) if drift else None, drift=tab)
    dr = so.Drift(nbo.Cosmo(**P), a_now) if drift else None

    def f(pp, vv):
        return float((so.image_positions(pp, mesh, box, obs, code=code, shift=shift, vel=vv if drift else None,
                                         drift=dr).reshape(pos.shape) * gbar).sum())

    p64, v64 = pos.astype(np.float64), vel.astype(np.float64)
    eps = 1e-5
    d = rng.standard_normal(pos.shape)
    fd = (f(p64 + eps * d, v64) - f(p64 - eps * d, v64)) / (2 * eps)
    assert float((gx.double().cpu().numpy() * d).sum()) == pytest.approx(fd, rel=1e-4)
    if drift:
        fdv = (f(p64, v64 + eps * d) - f(p64, v64 - eps * d)) / (2 * eps)
        assert float((gv.double().cpu().numpy() * d).sum()) == pytest.approx(fdv, rel=1e-4)
    else:
        assert gv is None


@pytest.mark.parametrize("kind,target,adjoint", [("drift", "flat", "reverse"), ("drift", "density", "reverse"),
                                                 ("drift", "particles", "checkpointed"),
                                                 ("telephoto", "flat", "reverse")])
def test_nbody_gradient_through_image_snapshots(pkg, cuda, kind, target, adjoint):
    """d <W, snapshot> / d (dx, p) for one shell painted from a drifted (DriftInterp) or rotated, shifted
    and drifted (TelephotoInterp, a shell beyond the box radius) image, against float64 central
    differences of the oracle's run, image transform and paint."""
    from oracle import pm_oracle as orc

    N, BOX, A0, DT0, FLAT = 16, 100.0, 0.1, 0.2, (16, 16)
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    ic = pkg.gaussian_initial_conditions(11, (N,) * 3, (BOX,) * 3, cosmo=c, flatsky_npix=FLAT)
    dx, p = pkg.lpt(c, ic, ts=A0)
    x0, v0 = dx.array.clone(), p.array.clone()
    rng = np.random.default_rng(64)
    center, width = (30.0, 30.0) if kind == "drift" else (70.0, 30.0)
    painting = pkg.PaintingOptions(target=target)
    interp = (pkg.DriftInterp(painting=painting) if kind == "drift"
              else pkg.TelephotoInterp(painting=painting, drift_on_lightcone=True))
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    ts = np.stack([pkg.cosmology.a_of_chi(c, np.array([center - width / 2])),
                   pkg.cosmology.a_of_chi(c, np.array([center + width / 2]))])
    dxf = dx.replace(array=x0.clone().requires_grad_(True))
    pf = p.replace(array=v0.clone().requires_grad_(True))
    out = pkg.nbody(c, dxf, pf, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver, adjoint=adjoint)
    W = rng.standard_normal(tuple(out.array.shape))
    (out.array.double() * torch.from_numpy(W).to(cuda)).sum().backward()
    r_c = float(np.asarray(out.comoving_centers).reshape(-1)[0])
    w_c = float(np.asarray(out.density_width).reshape(-1)[0])
    t_snap = float(np.asarray(out.scale_factors, dtype=np.float64).reshape(-1)[0])
    radius = dx.max_comoving_radius
    code, shift = (1, (0.0, 0.0, float(np.floor(r_c / radius) * radius))) if kind == "telephoto" else (0, (0.0,) * 3)
    assert kind == "drift" or r_c > radius
    x0n, v0n = x0.cpu().numpy().astype(np.float64), v0.cpu().numpy().astype(np.float64)
    obs = (BOX / 2,) * 3

    def loss(xx, vv):
        snap = {}

        def save(i, t, x, v):
            snap["xv"] = (x, v)

        nbo.run_nbody(o, xx, vv, A0, 1.0, DT0, np.array([t_snap], np.float32), solver="kkd", on_save=save)
        x, v = snap["xv"]
        g = so.image_positions(x, (N,) * 3, (BOX,) * 3, obs, code=code, shift=shift, vel=v,
                               drift=so.Drift(o, t_snap)).reshape((N,) * 3 + (3,))
        if target == "flat":
            zlo, zhi, scale = pkg.fields.flat_slab(dx, r_c, w_c)
            res = so.paint_flat(g, (N,) * 3, FLAT, zlo, zhi, scale, mode="absolute")
        elif target == "density":
            res = orc.cic_paint(g, (N,) * 3, mode="absolute", dtype=np.float64)
        else:
            res = g * (BOX / N)
        return float((res.reshape(W.shape) * W).sum())

    assert loss(x0n, v0n) != 0.0
    eps = 1e-6
    for which, grad in (("x", dxf.array.grad), ("v", pf.array.grad)):
        d = rng.standard_normal(x0n.shape)
        if which == "x":
            fd = (loss(x0n + eps * d, v0n) - loss(x0n - eps * d, v0n)) / (2 * eps)
        else:
            fd = (loss(x0n, v0n + eps * d) - loss(x0n, v0n - eps * d)) / (2 * eps)
        got = float((grad.double().cpu().numpy() * d).sum())
        print(f"{kind} {target} {adjoint} d/d{which}: adjoint {got:.6e} fd {fd:.6e}")
        assert abs(got - fd) / abs(fd) < 2e-3


@pytest.mark.parametrize("kind,adjoint", [("drift", "reverse"), ("drift", "checkpointed"), ("telephoto", "reverse")])
def test_cosmology_gradient_through_flat_image_snapshots(pkg, cuda, kind, adjoint):
    """d L / d (Omega_c, sigma8) through white noise -> ICs -> LPT -> N-body with one flat-sky snapshot
    painted from a drifted (DriftInterp) or rotated, shifted and drifted (TelephotoInterp, a shell beyond
    the box radius) image. Omega_c reaches the map through the ICs, the growth, the step coefficients and
    the drift factor f(chi) of the image transform (pmx_image_positions_vjp's parameter tables); sigma8
    through the IC amplitude. Against float64 central differences of the oracle's forward model at the
    snapshot's fixed geometry (scale factor, slab centre and width)."""
    N, BOX, A0, DT0, FLAT = 16, 100.0, 0.1, 0.2, (16, 16)
    KT = np.logspace(-4, 1, 128)
    c0 = pkg.Cosmology(**P)
    center, width = (30.0, 30.0) if kind == "drift" else (70.0, 30.0)
    ts = np.stack([pkg.cosmology.a_of_chi(c0, np.array([center - width / 2])),
                   pkg.cosmology.a_of_chi(c0, np.array([center + width / 2]))])
    rng = np.random.default_rng(83)
    noise = rng.standard_normal((N,) * 3).astype(np.float32)
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    s8 = torch.tensor(P["sigma8"], dtype=torch.float64, requires_grad=True)
    c = pkg.Cosmology(**{**P, "Omega_c": Oc, "sigma8": s8})
    ic = pkg.interpolate_initial_conditions(torch.from_numpy(noise).to(cuda), (N,) * 3, (BOX,) * 3, cosmo=c,
                                            flatsky_npix=FLAT)
    dx, p = pkg.lpt(c, ic, ts=A0)
    painting = pkg.PaintingOptions(target="flat")
    interp = (pkg.DriftInterp(painting=painting) if kind == "drift"
              else pkg.TelephotoInterp(painting=painting, drift_on_lightcone=True))
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    out = pkg.nbody(c, dx, p, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver, adjoint=adjoint)
    W = rng.standard_normal(tuple(out.array.shape))
    L = (out.array.double() * torch.from_numpy(W).to(cuda)).sum()
    L.backward()
    r_c = float(np.asarray(out.comoving_centers).reshape(-1)[0])
    w_c = float(np.asarray(out.density_width).reshape(-1)[0])
    t_snap = float(np.asarray(out.scale_factors, dtype=np.float64).reshape(-1)[0])
    radius = dx.max_comoving_radius
    code, shift = (1, (0.0, 0.0, float(np.floor(r_c / radius) * radius))) if kind == "telephoto" else (0, (0.0,) * 3)
    assert kind == "drift" or r_c > radius
    zlo, zhi, scale = pkg.fields.flat_slab(dx, r_c, w_c)
    obs = (BOX / 2,) * 3

    def oracle_loss(oc, sig):
        o = nbo.Cosmo(**{**P, "Omega_c": oc, "sigma8": sig})
        ic0 = nbo.color_field(noise.astype(np.float64), (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
        x0, v0 = nbo.lpt1(o, ic0, A0)
        snap = {}

        def save(i, t, x, v):
            snap["xv"] = (x, v)

        nbo.run_nbody(o, x0, v0, A0, 1.0, DT0, np.array([t_snap], np.float32), solver="kkd", on_save=save)
        x, v = snap["xv"]
        g = so.image_positions(x, (N,) * 3, (BOX,) * 3, obs, code=code, shift=shift, vel=v,
                               drift=so.Drift(o, t_snap)).reshape((N,) * 3 + (3,))
        res = so.paint_flat(g, (N,) * 3, FLAT, zlo, zhi, scale, mode="absolute")
        return float((res.reshape(W.shape) * W).sum())

    l0 = oracle_loss(P["Omega_c"], P["sigma8"])
    assert l0 != 0.0
    sc = float((out.array.detach().double().abs() * torch.from_numpy(np.abs(W)).to(cuda)).sum())
    assert abs(float(L) - l0) / sc < 1e-4
    h = 1e-6
    fd_oc = (oracle_loss(P["Omega_c"] + h, P["sigma8"]) - oracle_loss(P["Omega_c"] - h, P["sigma8"])) / (2 * h)
    fd_s8 = (oracle_loss(P["Omega_c"], P["sigma8"] + h) - oracle_loss(P["Omega_c"], P["sigma8"] - h)) / (2 * h)
    print(f"{kind} {adjoint}: d/dOmega_c adjoint {float(Oc.grad):.6e} fd {fd_oc:.6e}; "
          f"d/dsigma8 adjoint {float(s8.grad):.6e} fd {fd_s8:.6e}")
    assert abs(float(Oc.grad) - fd_oc) / abs(fd_oc) < 2e-3
    assert abs(float(s8.grad) - fd_s8) / abs(fd_s8) < 2e-3
