"""Sky projections on the GPU against the CPU oracle: HEALPix shell painting over box images (the
OnionTiler / periodic-replication paint) and its position adjoint, Born convergence and its adjoint,
LPT on a particle sub-lattice, and the gradient of a convergence-map loss through the whole forward
model (ICs -> LPT -> reversible N-body with tiled shells -> Born) against float64 finite
differences of the oracle's forward model.

Tolerances: maps and kappa 1e-5 relative L2 (float32 device arithmetic with float64 geometry and
accumulation against the float64 oracle); adjoints 1e-5 against the oracle's analytic adjoint; the
end-to-end gradient 2e-3 against finite differences (float32 forward and adjoint)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import nbody_oracle as nbo
from oracle import pm_oracle as orc
from oracle import sky_oracle as so

pytestmark = pytest.mark.gpu

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _disp(shape, scale, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(tuple(shape) + (3,)) * scale).astype(np.float32)


def _onion(box, seed=5):
    rots = np.random.default_rng(seed).integers(1, 48, size=27).astype(np.int32)
    return rots, so.onion_shifts((0.5, 0.5, 0.5), box)


@pytest.mark.parametrize("scheme", ["ngp", "bilinear"])
def test_tiled_healpix_paint_matches_oracle(pkg, cuda, scheme):
    mesh, box, obs = (16, 16, 16), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0)
    dx = _disp(mesh, 1.2, 21)
    rots, shifts = _onion(box)
    got = pkg.engine.paint_healpix(torch.from_numpy(dx).to(cuda), mesh, box, obs, 40.0, 90.0, 8, scheme=scheme,
                                   tiles=(rots, shifts)).cpu().numpy()
    ref = so.paint_healpix_tiled(dx, mesh, box, obs, 40.0, 90.0, 8, scheme=scheme, tiles=list(zip(rots, shifts)))
    assert rel(got, ref) < 1e-6


def test_identity_image_equals_plain_paint(pkg, cuda):
    mesh, box, obs = (16, 16, 16), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0)
    x = torch.from_numpy(_disp(mesh, 1.0, 22)).to(cuda)
    a = pkg.engine.paint_healpix(x, mesh, box, obs, 10.0, 45.0, 8).cpu().numpy()
    b = pkg.engine.paint_healpix(x, mesh, box, obs, 10.0, 45.0, 8, tiles=([0], [[0.0, 0.0, 0.0]])).cpu().numpy()
    assert rel(b, a) < 1e-7


def test_periodic_replication_sub_lattice(pkg, cuda):
    """512^3-on-1024^3-style set-up in miniature: a 16^3 particle lattice on a 32^3 mesh, replicated
    over [-2, 2]^3 boxes, painted to a shell beyond the box."""
    mesh, box, obs = (32, 32, 32), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0)
    dx = _disp((16, 16, 16), 1.5, 23)
    tiles = [(0, np.array(s, dtype=np.float64) * 100.0) for s in np.ndindex(5, 5, 5)]
    tiles = [(c, s - 200.0) for c, s in tiles]
    rots = np.array([c for c, _ in tiles], np.int32)
    shifts = np.stack([s for _, s in tiles])
    got = pkg.engine.paint_healpix(torch.from_numpy(dx).to(cuda), mesh, box, obs, 150.0, 210.0, 4,
                                   tiles=(rots, shifts)).cpu().numpy()
    ref = so.paint_healpix_tiled(dx, mesh, box, obs, 150.0, 210.0, 4, tiles=tiles)
    assert rel(got, ref) < 1e-6
    assert got.mean() == pytest.approx(16**3 / 100.0**3, rel=0.05)


@pytest.mark.parametrize("tiled", [False, True])
def test_healpix_vjp_matches_oracle(pkg, cuda, tiled):
    mesh, box, obs = (16, 16, 16), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0)
    dx = _disp(mesh, 1.2, 24)
    tiles = _onion(box) if tiled else None
    otiles = list(zip(*tiles)) if tiled else None
    rmin, rmax = (40.0, 90.0) if tiled else (10.0, 45.0)
    rng = np.random.default_rng(25)
    mb = rng.standard_normal(12 * 8 * 8).astype(np.float32)
    xd = torch.from_numpy(dx).to(cuda)
    got_t = pkg.engine.paint_healpix_vjp(xd, torch.from_numpy(mb).to(cuda), mesh, box, obs, rmin, rmax, 8,
                                         tiles=tiles)
    got = got_t.cpu().numpy()
    ref = so.paint_healpix_vjp(dx, mb, mesh, box, obs, rmin, rmax, 8, tiles=otiles)
    assert rel(got, ref) < 1e-5
    base = torch.ones_like(xd)
    pkg.engine.paint_healpix_vjp(xd, torch.from_numpy(mb).to(cuda), mesh, box, obs, rmin, rmax, 8, tiles=tiles,
                                 pos_bar=base, accumulate=True)
    assert torch.equal(base, 1.0 + got_t)
    z = pkg.engine.paint_healpix_vjp(xd, torch.from_numpy(mb).to(cuda), mesh, box, obs, rmin, rmax, 8,
                                     scheme="ngp", tiles=tiles)
    assert not z.any()


def _lightcone(pkg, maps_np, a, nside, box, cuda):
    S = maps_np.shape[0]
    return pkg.SphericalDensity(array=torch.from_numpy(maps_np).to(cuda), mesh_size=(16, 16, 16),
                                box_size=(box,) * 3, nside=nside, status=pkg.FieldStatus.LIGHTCONE,
                                scale_factors=a, comoving_centers=np.zeros(S), density_width=np.zeros(S))


@pytest.mark.parametrize("sources", ["distribution", "redshift"])
def test_born_matches_oracle(pkg, cuda, sources):
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    nside, box = 16, 1000.0
    rng = np.random.default_rng(26)
    a = np.array([0.55, 0.65, 0.75, 0.85, 0.95])
    maps = (rng.random((5, 12 * nside**2)) * 2.0).astype(np.float32)
    maps[3] = 0.0  # an empty shell: its mean is replaced by the float32 eps
    lc = _lightcone(pkg, maps, a, nside, box, cuda)
    if sources == "distribution":
        nz, onz = [pkg.tophat_z(0.0, 1.0), pkg.tophat_z(0.3, 0.8)], [so.tophat_pdf(0.0, 1.0), so.tophat_pdf(0.3, 0.8)]
    else:
        nz = onz = [0.5, 1.2]
    kap = pkg.born(c, lc, nz)
    r = o.chi(a)
    ref = so.born_spherical(maps.astype(np.float64), r, a, onz, lc.max_comoving_radius / 5, o.Omega_m, o.chi)
    assert kap.array.shape == (2, 12 * nside**2)
    assert kap.status == pkg.FieldStatus.KAPPA
    assert rel(kap.array.cpu().numpy(), ref) < 1e-5


def test_born_vjp_and_parameter_gradient(pkg, cuda):
    o = nbo.Cosmo(**P)
    nside, box = 8, 1000.0
    rng = np.random.default_rng(27)
    a = np.array([0.6, 0.7, 0.8, 0.9])
    maps = (rng.random((4, 12 * nside**2)) + 0.5).astype(np.float32)
    kb = rng.standard_normal((1, 12 * nside**2))
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    c = pkg.Cosmology(**{**P, "Omega_c": Oc})
    mt = torch.from_numpy(maps).to(cuda).requires_grad_(True)
    lc = _lightcone(pkg, maps, a, nside, box, cuda).replace(array=mt)
    kap = pkg.born(c, lc, [pkg.tophat_z(0.0, 1.0)])
    (kap.array.double() * torch.from_numpy(kb).to(cuda)).sum().backward()
    d_r = lc.max_comoving_radius / 4
    onz = [so.tophat_pdf(0.0, 1.0)]
    ref = so.born_vjp(maps.astype(np.float64), kb, o.chi(a), a, onz, d_r, o.Omega_m, o.chi)
    assert rel(mt.grad.cpu().numpy(), ref) < 1e-5
    h = 1e-5

    def L(oc):
        oo = nbo.Cosmo(**{**P, "Omega_c": oc})
        return float((so.born_spherical(maps.astype(np.float64), oo.chi(a), a, onz, d_r, oo.Omega_m, oo.chi)
                      * kb).sum())

    fd = (L(P["Omega_c"] + h) - L(P["Omega_c"] - h)) / (2 * h)
    assert float(Oc.grad) == pytest.approx(fd, rel=1e-4)


def test_lpt_sub_lattice_is_the_full_lattice_subsampled(pkg, cuda):
    c = pkg.Cosmology(**P)
    n, box = 32, 200.0
    ic = pkg.gaussian_initial_conditions(8, (n, n, n), (box,) * 3, cosmo=c)
    dx_full, p_full = pkg.lpt(c, ic, ts=0.1)
    lat = torch.stack(torch.meshgrid(*[torch.arange(16, dtype=torch.float32, device=cuda) * 2] * 3, indexing="ij"),
                      dim=-1)
    dx, p = pkg.lpt(c, ic, ts=0.1, initial_particles=lat)
    assert tuple(dx.array.shape) == (16, 16, 16, 3)
    assert torch.equal(dx.array, dx_full.array[::2, ::2, ::2])
    assert torch.equal(p.array, p_full.array[::2, ::2, ::2])
    # the adjoint places the cotangent on the lattice cells: same as the full lattice with zeros between
    rng = np.random.default_rng(28)
    g = torch.from_numpy(rng.standard_normal((16, 16, 16, 3)).astype(np.float32)).to(cuda)
    gf = torch.zeros((n, n, n, 3), device=cuda)
    gf[::2, ::2, ::2] = g
    a = pkg.engine.lpt1_vjp(g, 0.5 * g, (n, n, n), 0.3, 2.0)
    b = pkg.engine.lpt1_vjp(gf, 0.5 * gf, (n, n, n), 0.3, 2.0)
    assert torch.equal(a, b)


def test_convergence_loss_gradient_matches_oracle_fd(pkg, cuda):
    """d/d noise of sum(kappa^2), kappa = Born of three shells (one inside the box, two painted from
    27 rotated images) of a reversible N-body run, against float64 finite differences of the
    oracle's forward model along two random directions."""
    N, BOX, A0, DT0, NSIDE = 16, 100.0, 0.1, 0.15, 4
    KT = np.logspace(-4, 1, 128)
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    edges = np.array([20.0, 45.0, 90.0, 130.0])  # Mpc/h; box half-width 50
    r_c = 0.5 * (edges[1:] + edges[:-1])[::-1]
    widths = (edges[1:] - edges[:-1])[::-1]
    a_near = pkg.cosmology.a_of_chi(c, r_c - widths / 2)
    a_far = pkg.cosmology.a_of_chi(c, r_c + widths / 2)
    ts = np.stack([a_near, a_far])
    painting = pkg.PaintingOptions(target="spherical", scheme="bilinear")
    tiler = pkg.OnionTiler(painting=painting, drift_on_lightcone=False)
    rots = tiler.init().rotations
    rng = np.random.default_rng(29)
    noise = rng.standard_normal((N,) * 3).astype(np.float32)
    noise_t = torch.from_numpy(noise).to(cuda).requires_grad_(True)
    ic = pkg.interpolate_initial_conditions(noise_t, (N,) * 3, (BOX,) * 3, cosmo=c, nside=NSIDE)
    dx, p = pkg.lpt(c, ic, ts=A0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=tiler)
    lc = pkg.nbody(c, dx, p, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver, adjoint="reverse")
    kap = pkg.born(c, lc, [pkg.tophat_z(0.0, 0.5)])
    L = (kap.array.double() ** 2).sum()
    L.backward()
    g = noise_t.grad.double().cpu().numpy()

    t_snap = np.asarray(lc.scale_factors, dtype=np.float64)
    r_snap = np.asarray(lc.comoving_centers, dtype=np.float64)
    w_snap = np.asarray(lc.density_width, dtype=np.float64)
    order = np.argsort(t_snap)  # snapshot order of the run: ascending time
    d_r = lc.max_comoving_radius / len(t_snap)
    onz = [so.tophat_pdf(0.0, 0.5)]

    def oracle_loss(nz):
        ic0 = nbo.color_field(nz, (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
        x0, v0 = nbo.lpt1(o, ic0, A0)
        maps = {}

        def save(i, t, x, v):
            s = order[i]
            rc, wd = r_snap[s], w_snap[s]
            tiles = None if rc + wd / 2 <= lc.max_comoving_radius else list(zip(rots, so.onion_shifts(
                (0.5, 0.5, 0.5), (BOX,) * 3)))
            maps[s] = so.paint_healpix_tiled(x, (N,) * 3, (BOX,) * 3, (BOX / 2,) * 3, rc - wd / 2, rc + wd / 2, NSIDE,
                                             tiles=tiles)

        nbo.run_nbody(o, x0, v0, A0, 1.0, DT0, np.sort(t_snap).astype(np.float32), solver="kkd", on_save=save)
        m = np.stack([maps[s] for s in range(len(t_snap))])
        k = so.born_spherical(m, o.chi(t_snap), t_snap, onz, d_r, o.Omega_m, o.chi)
        return float((k**2).sum())

    L0 = oracle_loss(noise.astype(np.float64))
    assert abs(float(L) - L0) / L0 < 1e-4
    eps = 1e-6
    for j in range(2):
        d = np.random.default_rng(30 + j).standard_normal(noise.shape)
        fd = (oracle_loss(noise + eps * d) - oracle_loss(noise - eps * d)) / (2 * eps)
        got = float((g * d).sum())
        print(f"direction {j}: adjoint {got:.6e} fd {fd:.6e}")
        assert abs(got - fd) / abs(fd) < 2e-3


def test_large_structure_gradient_and_shell_membership(pkg, cuda):
    """The 512^3-on-1024^3 workload's structure at 16^3 particles on the 2-cell sub-lattice of a 32^3
    mesh (box 100 Mpc/h): ReversibleDoubleKickDrift with the reverse adjoint, 8 bilinear shells of equal
    width to 230 Mpc/h painted from the 125 periodic images (OnionTiler layers=2, rotate=False), Born for a
    top-hat n(z), loss sum(kappa^2). Against float64 central differences of the oracle's model:

    * at eps = 1e-6 (no particle-image crosses a shell edge) along two random noise directions and along
      the gradient's own direction, the adjoint agrees to 2e-3;
    * along the gradient direction at eps from 1e-2 down to 1e-5, the differences with the shell
      membership frozen at the unperturbed run (so that only the smooth part of the model moves) converge
      to the adjoint as eps shrinks (within 1e-3 at 1e-5) but are more than 10% off at eps = 1e-3: the
      loss is strongly curved along the direction that raises it fastest. The free differences add the
      membership steps (a particle-image that crosses a shell edge moves the whole particle's weight
      between two shells), which the adjoint, like autodiff, takes as zero; the table of both against
      eps, with the number of particle-images that change shell, is printed."""
    NP, NM, BOX, A0, DT0, NSIDE, NSH = 16, 32, 100.0, 0.1, 0.1, 16, 8
    KT = np.logspace(-4, 1, 128)
    c, o = pkg.Cosmology(**P), nbo.Cosmo(**P)
    chi_max = 230.0
    edges = np.linspace(0.0, chi_max, NSH + 1)
    r_c = (0.5 * (edges[1:] + edges[:-1]))[::-1]
    w = np.diff(edges)[::-1]
    ts = np.stack([pkg.cosmology.a_of_chi(c, r_c - w / 2), pkg.cosmology.a_of_chi(c, r_c + w / 2)])
    painting = pkg.PaintingOptions(target="spherical", scheme="bilinear")
    tiler = pkg.OnionTiler(painting=painting, drift_on_lightcone=False, layers=2, rotate=False)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=tiler)
    lat = torch.stack(torch.meshgrid(*[torch.arange(NP, dtype=torch.float32, device=cuda) * (NM // NP)] * 3,
                                     indexing="ij"), dim=-1)
    zmax = float(pkg.cosmology.a2z(pkg.cosmology.a_of_chi(c, chi_max)))
    rng = np.random.default_rng(77)
    noise = rng.standard_normal((NM,) * 3).astype(np.float32)
    noise_t = torch.from_numpy(noise).to(cuda).requires_grad_(True)
    ic = pkg.interpolate_initial_conditions(noise_t, (NM,) * 3, (BOX,) * 3, cosmo=c, nside=NSIDE)
    dx, p = pkg.lpt(c, ic, ts=A0, initial_particles=lat)
    lc = pkg.nbody(c, dx, p, t0=A0, t1=1.0, dt0=DT0, ts=ts, solver=solver, adjoint="reverse")
    kap = pkg.born(c, lc, [pkg.tophat_z(0.0, zmax)])
    L = (kap.array.double() ** 2).sum()
    L.backward()
    g = noise_t.grad.double().cpu().numpy()

    t_snap = np.asarray(lc.scale_factors, dtype=np.float64)
    r_snap = np.asarray(lc.comoving_centers, dtype=np.float64)
    w_snap = np.asarray(lc.density_width, dtype=np.float64)
    order = np.argsort(t_snap)
    d_r = lc.max_comoving_radius / len(t_snap)
    onz = [so.tophat_pdf(0.0, zmax)]
    shifts = so.onion_shifts((0.5, 0.5, 0.5), (BOX,) * 3, layers=2)
    tiles_all = list(zip(np.zeros(len(shifts), dtype=np.int32), shifts))
    force = lambda xx: orc.pm_forces(xx, (NM,) * 3, "relative", scale=1.5 * o.Omega_m)

    def oracle_loss(nz, frozen=None, record=None):
        ic0 = nbo.color_field(nz, (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
        F = orc.force_mesh(ic0)[::NM // NP, ::NM // NP, ::NM // NP]
        D = float(o.growth_factor(A0)[0])
        x0 = D * F
        v0 = A0**2 * float(o.growth_rate(A0)[0]) * float(o.E(A0)[0]) * x0
        maps = {}

        def save(i, t, x, v):
            s = order[i]
            rc, wd = r_snap[s], w_snap[s]
            tiles = None if rc + wd / 2 <= lc.max_comoving_radius else tiles_all
            args = (x, (NM,) * 3, (BOX,) * 3, (BOX / 2,) * 3, rc - wd / 2, rc + wd / 2)
            if record is not None:
                record[s] = so.shell_members(*args, tiles=tiles)
            maps[s] = so.paint_healpix_tiled(*args, NSIDE, tiles=tiles, members=None if frozen is None else frozen[s])

        nbo.run_nbody(o, x0, v0, A0, 1.0, DT0, np.sort(t_snap).astype(np.float32), solver="kkd", force_fn=force,
                      on_save=save)
        m = np.stack([maps[s] for s in range(len(t_snap))])
        k = so.born_spherical(m, o.chi(t_snap), t_snap, onz, d_r, o.Omega_m, o.chi)
        return float((k**2).sum())

    base = noise.astype(np.float64)
    members = {}
    L0 = oracle_loss(base, record=members)
    print(f"loss {float(L):.6e} oracle {L0:.6e}; images per far shell "
          f"{[int(sum(len(m) for m in members[s])) for s in range(NSH)]}")
    assert abs(float(L) - L0) / L0 < 1e-4
    dirs = [np.random.default_rng(80 + j).standard_normal(noise.shape) for j in range(2)]
    dirs.append(g / np.linalg.norm(g) * np.sqrt(g.size))
    for j, d in enumerate(dirs):
        eps = 1e-6
        fd = (oracle_loss(base + eps * d) - oracle_loss(base - eps * d)) / (2 * eps)
        got = float((g * d).sum())
        print(f"{'gradient' if j == 2 else 'noise'} direction {j}: adjoint {got:.6e} fd(1e-6) {fd:.6e}")
        assert abs(got - fd) / abs(fd) < 2e-3
    d = dirs[2]
    got = float((g * d).sum())
    rows = []
    for eps in (1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 1e-5):
        free = (oracle_loss(base + eps * d) - oracle_loss(base - eps * d)) / (2 * eps)
        frozen = (oracle_loss(base + eps * d, frozen=members) - oracle_loss(base - eps * d, frozen=members)) / (2 * eps)
        moved = {}
        oracle_loss(base + eps * d, record=moved)
        crossed = sum(len(np.setxor1d(a, b)) for s in range(NSH) for a, b in zip(members[s], moved[s]))
        rows.append((eps, free, frozen, crossed))
        print(f"gradient direction, eps {eps:.0e}: adjoint {got:.6e}, fd frozen membership {frozen:.6e} "
              f"(rel {abs(frozen - got) / abs(got):.1e}), fd free {free:.6e} (rel {abs(free - got) / abs(got):.1e}), "
              f"{crossed} particle-images changed shell at +eps")
    err = {eps: (abs(fr - got) / abs(got), abs(fe - got) / abs(got)) for eps, fe, fr, _ in rows}
    # the adjoint is the derivative: with the membership frozen the difference converges to it as eps -> 0
    assert err[1e-5][0] < 1e-3
    # the step of the large-configuration check (1e-3) is far outside the linear range along this direction:
    # the curvature alone (membership frozen) moves the difference by more than 10%
    assert err[1e-3][0] > 0.1
