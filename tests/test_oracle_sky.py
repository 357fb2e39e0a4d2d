"""Pins of the sky oracle (HEALPix shell painting over box images, its adjoint, Born convergence):
each restated function is checked against what mathematics fixes, chosen so that a dropped term, a
wrong sign, index or axis fails one of them. The reference commits no golden vectors for lensing or
spherical painting."""

from __future__ import annotations

import itertools

import numpy as np
import pytest
from scipy import integrate as si

from oracle import nbody_oracle as nbo
from oracle import pm_oracle as orc
from oracle import sky_oracle as so


# ---------------------------------------------------------------- JAX's generator
def test_threefry_known_answers():
    """Random123's known-answer vectors for Threefry-2x32 with 20 rounds (key, counter) -> output."""
    m = 0xFFFFFFFF
    assert so.threefry2x32(0, 0, 0, 0) == (0x6B200159, 0x99BA4EFE)
    assert so.threefry2x32(m, m, m, m) == (0x1CB996FC, 0xBB002BE7)
    assert so.threefry2x32(0x13198A2E, 0x03707344, 0x243F6A88, 0x85A308D3) == (0xC4923A9C, 0x483DF7A0)


def test_jax_split_original_matches_the_documented_keys():
    """JAX's documentation prints split(PRNGKey(0)) under the original scheme as [4146024105,
    967050713] and [2718843009, 1272950319]: pins the key words and the counter layout."""
    assert so.jax_key(0) == (0, 0) and so.jax_key(42) == (0, 42) and so.jax_key((3 << 32) + 5) == (3, 5)
    assert so.jax_split_original(so.jax_key(0), 2) == [(4146024105, 967050713), (2718843009, 1272950319)]


@pytest.mark.parametrize("x64", [False, True])
def test_jax_randint_range_and_uniformity(x64):
    n, span = 47 * 300, 47
    v = so.jax_randint(7, n, 0, span, x64=x64)
    assert v.min() >= 0 and v.max() < span
    counts = np.bincount(v, minlength=span)
    chi2 = float(((counts - n / span) ** 2 / (n / span)).sum())
    assert chi2 < 46 + 6 * np.sqrt(2 * 46)  # 46 degrees of freedom, about six sigma
    assert np.array_equal(so.jax_randint(7, 5, 10, 12, x64=x64) >= 10, np.ones(5, bool))
    # 32- and 64-bit draws differ (different words and widths)
    assert not np.array_equal(so.jax_randint(42, 27, 0, 47), so.jax_randint(42, 27, 0, 47, x64=True))


# ---------------------------------------------------------------- images
def test_cube_symmetries_are_the_48_signed_permutations():
    mats = [so.cube_symmetry(c) for c in range(48)]
    assert np.array_equal(mats[0], np.eye(3))
    keys = {m.tobytes() for m in mats}
    assert len(keys) == 48
    for m in mats:
        assert np.allclose(m @ m.T, np.eye(3))
        assert np.array_equal(np.abs(m).sum(axis=0), np.ones(3)) and np.array_equal(np.abs(m).sum(axis=1), np.ones(3))
    # enumeration order: signs outer (product([1, -1])), permutations inner (lexicographic)
    assert np.array_equal(mats[1], np.array([[1, 0, 0], [0, 0, 1], [0, 1, 0]]))
    assert np.array_equal(mats[6], np.diag([1, 1, -1]))
    assert np.array_equal(mats[6 * 7 + 5], -np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]]))


def test_onion_shifts():
    box = (100.0, 200.0, 300.0)
    sh = so.onion_shifts((0.5, 0.5, 0.5), box)
    assert sh.shape == (27, 3) and len({tuple(s) for s in sh}) == 27
    assert np.array_equal(sh[0], np.zeros(3)) and np.array_equal(sh[1], [0, 0, 300.0])
    assert np.array_equal(sh[2], [0, 0, -300.0])
    sh0 = so.onion_shifts((0.0, 1.0, 0.5), box)
    assert (sh0[:, 0] >= 0).all() and (sh0[:, 1] <= 0).all()
    assert len(so.onion_shifts((0.5, 0.5, 0.5), box, layers=2)) == 125


# ---------------------------------------------------------------- tiled shell paint
def test_single_identity_image_is_the_plain_paint():
    rng = np.random.default_rng(3)
    dx = (rng.standard_normal((8, 8, 8, 3)) * 0.8).astype(np.float32)
    args = ((8, 8, 8), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0), 10.0, 45.0, 4)
    for scheme in ("ngp", "bilinear"):
        a = so.paint_healpix_tiled(dx, *args, scheme=scheme, tiles=[(0, (0.0, 0.0, 0.0))])
        b = orc.paint_healpix(dx, *args, scheme=scheme)
        assert np.allclose(a, b.astype(np.float64), rtol=1e-6, atol=1e-9)


def test_periodic_images_fill_a_far_shell_with_the_mean_density():
    """A uniform lattice replicated over [-2, 2]^3 boxes: a shell between 1.6 and 2.2 box lengths
    holds the mean density n = N / L^3, i.e. the map's pixel mean (count per pixel volume)."""
    n, L = 8, 100.0
    tiles = [(0, np.array(s, dtype=np.float64) * L) for s in itertools.product(range(-2, 3), repeat=3)]
    m = so.paint_healpix_tiled(np.zeros((n, n, n, 3), np.float32), (n, n, n), (L, L, L), (50.0, 50.0, 50.0),
                               160.0, 220.0, 2, scheme="ngp", tiles=tiles)
    assert m.mean() == pytest.approx(n**3 / L**3, rel=0.02)


def test_rotated_image_of_a_symmetric_set_is_unchanged():
    """The lattice about a lattice-point observer is invariant under every cube symmetry."""
    n, L = 8, 80.0
    base = so.paint_healpix_tiled(np.zeros((n, n, n, 3), np.float32), (n, n, n), (L, L, L), (40.0, 40.0, 40.0),
                                  5.0, 35.0, 2, scheme="ngp")
    for code in (1, 17, 46):
        rot = so.paint_healpix_tiled(np.zeros((n, n, n, 3), np.float32), (n, n, n), (L, L, L), (40.0, 40.0, 40.0),
                                     5.0, 35.0, 2, scheme="ngp", tiles=[(code, np.zeros(3))])
        assert rot.sum() == pytest.approx(base.sum(), rel=1e-12)


def test_interp_weight_derivatives_match_differences():
    rng = np.random.default_rng(4)
    for nside in (1, 4, 16):
        for theta in list(rng.uniform(0.01, np.pi - 0.01, 12)) + [0.05, np.pi - 0.05]:
            phi = rng.uniform(0, 2 * np.pi)
            pix, w, dth, dph = so.healpix_interp_weights_grad(nside, theta, phi)
            assert np.allclose(w, orc.healpix_interp_weights(nside, theta, phi)[1], atol=1e-14)
            h = 1e-7
            for d, (t1, p1, t0, p0) in ((dth, (theta + h, phi, theta - h, phi)), (dph, (theta, phi + h, theta, phi - h))):
                p_hi, w_hi = orc.healpix_interp_weights(nside, t1, p1)
                p_lo, w_lo = orc.healpix_interp_weights(nside, t0, p0)
                if list(p_hi) != list(pix) or list(p_lo) != list(pix):
                    continue  # a ring or pixel boundary lies within h
                assert np.allclose((np.array(w_hi) - np.array(w_lo)) / (2 * h), d, atol=1e-6)


@pytest.mark.parametrize("tiles", [None, [(0, (0.0, 0.0, 0.0)), (13, (100.0, 0.0, 0.0)), (40, (0.0, -100.0, 100.0))]])
def test_healpix_vjp_matches_differences(tiles):
    rng = np.random.default_rng(5)
    args = ((8, 8, 8), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0), 20.0, 150.0, 4)
    dx = rng.standard_normal((8, 8, 8, 3)) * 0.7
    m = so.paint_healpix_tiled(dx, *args, tiles=tiles)
    mb = rng.standard_normal(m.shape)
    g = so.paint_healpix_vjp(dx, mb, *args, tiles=tiles)
    for j in range(2):
        d = rng.standard_normal(dx.shape)
        e = 1e-6
        fd = ((so.paint_healpix_tiled(dx + e * d, *args, tiles=tiles) -
               so.paint_healpix_tiled(dx - e * d, *args, tiles=tiles)) * mb).sum() / (2 * e)
        assert (g * d).sum() == pytest.approx(fd, rel=1e-6)
    assert not so.paint_healpix_vjp(dx, mb, *args, scheme="ngp", tiles=tiles).any()


# ---------------------------------------------------------------- Born
def test_simpson_weights_exact_for_cubics():
    w = so.simps_weights(0.3, 2.1, 16)
    z = np.linspace(0.3, 2.1, 17)
    f = 1 - 2 * z + 0.5 * z**2 + 0.25 * z**3
    F = lambda x: x - x**2 + x**3 / 6 + x**4 / 16
    assert (w * f).sum() == pytest.approx(F(2.1) - F(0.3), rel=1e-13)


def test_born_thin_planes_converge_to_the_born_integral():
    """delta(chi, p) = Y(p) on planes of width d_r out to the source: the plane sum is the midpoint
    rule of 3/2 Om (H0/c)^2 Y int_0^chi_s chi (chi_s - chi) / (chi_s a(chi)) dchi."""
    c = nbo.Cosmo(Omega_c=0.25, Omega_b=0.0486, h=0.6774)
    zs = 0.8
    chi_s = float(c.chi(1 / (1 + zs))[0])
    S = 400
    d_r = chi_s / S
    r = (np.arange(S) + 0.5) * d_r
    agrid = np.linspace(0.3, 1.0, 4001)
    chig = c.chi(agrid)
    a = np.interp(r, chig[::-1], agrid[::-1])
    Y = np.array([0.3, -0.3, 0.1, -0.1])
    maps = 1.0 + np.outer(np.ones(S), Y)
    k = so.born_spherical(maps, r, a, [zs], d_r, c.Omega_m, c.chi)
    a_of = lambda x: float(np.interp(x, chig[::-1], agrid[::-1]))
    integral = si.quad(lambda x: x * (chi_s - x) / (chi_s * a_of(x)), 0, chi_s, limit=200)[0]
    expect = 1.5 * c.Omega_m * so.H0_OVER_C**2 * integral * Y
    assert np.allclose(k[0], expect, rtol=2e-5)


def test_born_is_linear_in_delta_and_ignores_the_mean():
    c = nbo.Cosmo()
    rng = np.random.default_rng(6)
    Y = rng.standard_normal((3, 40))
    Y -= Y.mean(axis=1, keepdims=True)
    r, a = np.array([200.0, 400.0, 600.0]), np.array([0.93, 0.87, 0.81])
    src = [so.tophat_pdf(0.0, 1.0)]
    k1 = so.born_spherical(1.0 + 0.1 * Y, r, a, src, 100.0, c.Omega_m, c.chi)
    k2 = so.born_spherical(5.0 * (1.0 + 0.2 * Y), r, a, src, 100.0, c.Omega_m, c.chi)
    assert np.allclose(k2, 2 * k1, rtol=1e-12, atol=1e-18)
    assert so.born_spherical(np.ones((3, 40)), r, a, src, 100.0, c.Omega_m, c.chi) == pytest.approx(0.0, abs=1e-18)


def test_born_vjp_matches_differences():
    c = nbo.Cosmo(Omega_c=0.25, Omega_b=0.0486, h=0.6774)
    rng = np.random.default_rng(7)
    maps = rng.random((4, 48)) + 0.5
    maps[2] = 0.0  # an empty shell: the eps mean is a constant
    r = np.array([100.0, 200.0, 300.0, 400.0])
    a = 1 / (1 + np.array([0.03, 0.07, 0.1, 0.14]))
    for src in ([so.tophat_pdf(0, 1), so.tophat_pdf(0.2, 0.6)], [0.5, 1.0]):
        k = so.born_spherical(maps, r, a, src, 62.5, c.Omega_m, c.chi)
        kb = rng.standard_normal(k.shape)
        g = so.born_vjp(maps, kb, r, a, src, 62.5, c.Omega_m, c.chi)
        d = rng.standard_normal(maps.shape)
        d[2] = 0.0
        e = 1e-6
        fd = ((so.born_spherical(maps + e * d, r, a, src, 62.5, c.Omega_m, c.chi) -
               so.born_spherical(maps - e * d, r, a, src, 62.5, c.Omega_m, c.chi)) * kb).sum() / (2 * e)
        assert (g * d).sum() == pytest.approx(fd, rel=1e-7)


def test_tophat_normalisation():
    nz = so.tophat_pdf(0.0, 1.0)
    z = np.linspace(0, 10, 257)
    assert (so.simps_weights(0, 10, 256) * nz(z)).sum() == pytest.approx(1.0, rel=1e-12)
    assert nz(np.array([0.5]))[0] == pytest.approx(nz(np.array([0.9]))[0]) and nz(np.array([1.01]))[0] == 0


def test_drift_factor_closed_form_in_eds():
    """Omega_m = 1: Gp = a, dGfa = 1.5 a^(1/2), chi = 2 c/H0 (1 - a^(1/2))."""
    c = nbo.Cosmo(Omega_c=1.0 - 0.049, Omega_b=0.049, h=0.7)
    d = so.Drift(c, 0.8)
    r = np.array([10.0, 300.0, 900.0, 1500.0])
    at = (1.0 - r / (2 * 2997.92458)) ** 2
    expect = (at - 0.8) / (1.5 * (0.8 * at) ** 0.25)
    assert np.allclose(d.factor(r), expect, rtol=1e-6, atol=1e-9)
    r0 = 2 * 2997.92458 * (1 - np.sqrt(0.8))
    assert abs(d.factor(np.array([r0]))[0]) < 1e-8


# ---------------------------------------------------------------- flat-sky Born
def test_flat_resampling_is_scipy_grid_wrap():
    from scipy import ndimage

    rng = np.random.default_rng(8)
    plane = rng.standard_normal((12, 20))
    coords = np.stack([rng.uniform(-30, 40, (5, 7)), rng.uniform(-45, 70, (5, 7))])
    ref = ndimage.map_coordinates(plane, coords.reshape(2, -1), order=1, mode="grid-wrap").reshape(5, 7)
    assert np.allclose(so.map_coordinates_wrap(plane, coords), ref, rtol=1e-12, atol=1e-12)
    ints = np.stack(np.meshgrid(np.arange(-12, 0), np.arange(20, 40), indexing="ij")).astype(np.float64)
    assert np.array_equal(so.map_coordinates_wrap(plane, ints), plane)


def test_flat_coords_put_x_angles_on_rows():
    """x angles (linspace over nx) index rows over pixel_size[0] = box_y / ny, y angles columns."""
    c = so.flat_coords((4.0, 2.0), (10, 20), (300.0, 100.0, 50.0), 250.0, (10, 20))
    assert c.shape == (2, 10, 20)
    j = np.arange(20)
    assert np.allclose(c[0], (j * 4.0 / 20 * np.pi / 180 * 250.0 / (100.0 / 10) - 0.5)[None, :])
    i = np.arange(10)
    assert np.allclose(c[1], (i * 2.0 / 10 * np.pi / 180 * 250.0 / (300.0 / 20) - 0.5)[:, None])


def test_flat_born_impulse_lands_on_its_bilinear_footprint():
    """Unit sampling scale: output (i, j) averages rows j-1, j and columns i-1, i of the plane, so one
    raised pixel at (r, c) raises exactly the outputs i in {c, c+1}, j in {r, r+1} by a quarter."""
    c = nbo.Cosmo()
    n, chi, box, A = 8, 200.0, 160.0, 3.0
    fsz = (box / n) * n * 180 / (np.pi * chi)  # (f / n) (pi / 180) chi / pixel = 1
    geom = ((fsz, fsz), (n, n), (box, box, box))
    plane = np.ones((1, n, n))
    plane[0, 2, 5] += A
    a = np.array([0.95])
    k = so.born_flat(plane, np.array([chi]), a, [1.0], 50.0, c.Omega_m, c.chi, geom)[0]
    C = 1.5 * c.Omega_m * so.H0_OVER_C**2 * 50.0 * chi / a[0] * (1 - chi / c.chi(np.array([0.5]))[0])
    m = 1 + A / n**2
    hit = np.zeros((n, n))
    for i in (5, 6):
        for j in (2, 3):
            hit[i, j] = 1
    assert np.allclose(k, C * ((1 + A / 4 * hit) / m - 1), rtol=1e-12, atol=1e-15)


def test_flat_born_linear_blind_to_mean_and_vjp():
    c = nbo.Cosmo(Omega_c=0.25, Omega_b=0.0486, h=0.6774)
    rng = np.random.default_rng(9)
    S, ny, nx = 3, 10, 12
    maps = rng.random((S, ny, nx)) + 0.5
    r = np.array([150.0, 300.0, 450.0])
    a = np.array([0.95, 0.9, 0.86])
    geom = ((5.0, 4.0), (ny, nx), (400.0, 350.0, 400.0))
    for src in ([so.tophat_pdf(0, 1)], [0.6, 1.2]):
        k1 = so.born_flat(maps, r, a, src, 100.0, c.Omega_m, c.chi, geom)
        assert np.allclose(so.born_flat(4.0 * maps, r, a, src, 100.0, c.Omega_m, c.chi, geom), k1, rtol=1e-12)
        kb = rng.standard_normal(k1.shape)
        g = so.born_flat_vjp(maps, kb, r, a, src, 100.0, c.Omega_m, c.chi, geom)
        d = rng.standard_normal(maps.shape)
        e = 1e-6
        fd = ((so.born_flat(maps + e * d, r, a, src, 100.0, c.Omega_m, c.chi, geom) -
               so.born_flat(maps - e * d, r, a, src, 100.0, c.Omega_m, c.chi, geom)) * kb).sum() / (2 * e)
        assert (g * d).sum() == pytest.approx(fd, rel=1e-7)


def test_drifted_healpix_vjp_matches_differences():
    c = nbo.Cosmo(Omega_c=0.25, Omega_b=0.0486, h=0.6774)
    d = so.Drift(c, 0.9)
    rng = np.random.default_rng(10)
    args = ((8, 8, 8), (100.0, 100.0, 100.0), (50.0, 50.0, 50.0), 20.0, 150.0, 4)
    tiles = [(0, (0.0, 0.0, 0.0)), (13, (100.0, 0.0, 0.0)), (40, (0.0, -100.0, 100.0))]
    dx = rng.standard_normal((8, 8, 8, 3)) * 0.7
    v = rng.standard_normal((8, 8, 8, 3)) * 30.0
    m = so.paint_healpix_tiled(dx, *args, tiles=tiles, vel=v, drift=d)
    mb = rng.standard_normal(m.shape)
    gx, gv = so.paint_healpix_vjp(dx, mb, *args, tiles=tiles, vel=v, drift=d)
    assert np.abs(gv).sum() > 0
    for which in ("x", "v"):
        dd = rng.standard_normal(dx.shape)
        e = 1e-6
        if which == "x":
            hi = so.paint_healpix_tiled(dx + e * dd, *args, tiles=tiles, vel=v, drift=d)
            lo = so.paint_healpix_tiled(dx - e * dd, *args, tiles=tiles, vel=v, drift=d)
            got = (gx * dd).sum()
        else:
            hi = so.paint_healpix_tiled(dx, *args, tiles=tiles, vel=v + e * dd, drift=d)
            lo = so.paint_healpix_tiled(dx, *args, tiles=tiles, vel=v - e * dd, drift=d)
            got = (gv * dd).sum()
        fd = ((hi - lo) * mb).sum() / (2 * e)
        assert got == pytest.approx(fd, rel=1e-6), which


# ---------------------------------------------------------------- flat-sky slab paint
def test_flat_paint_of_whole_box_is_projected_cic():
    """A slab holding the whole box on a flat grid of the mesh's (nx, ny): every z weight sums to one,
    so the flat map times the pixel volume is the (pinned) 3D CIC paint summed over z."""
    rng = np.random.default_rng(3)
    mesh = (8, 6, 4)
    dx = rng.normal(0.0, 1.5, (4, 3, 2, 3))
    w = rng.uniform(0.5, 2.0, 4 * 3 * 2)
    for mode, pos in (("relative", dx), ("absolute", rng.uniform(0, 1, dx.shape) * np.array(mesh))):
        flat = so.paint_flat(pos, mesh, (8, 6), -1e9, 1e9, 1.0, mode=mode, weight=w)
        ref = orc.cic_paint(pos, mesh, mode=mode, weight=w, dtype=np.float64).sum(axis=2)
        assert np.allclose(flat, ref, rtol=1e-12, atol=1e-12)


def test_flat_paint_pixels_slab_edges_and_scale():
    mesh = (4, 4, 8)
    pos = np.zeros((1, 1, 4, 3), dtype=np.float32)
    pos[0, 0, :, 0] = [1.0, 1.5, 3.0, 3.999]
    pos[0, 0, :, 1] = [2.0, 2.0, 0.0, 1.0]
    pos[0, 0, :, 2] = [2.0, 3.0, 5.0, 2.0]  # slab (2, 5]: the first and last are out
    m = so.paint_flat(pos, mesh, (8, 4), 2.0, 5.0, 0.5, mode="absolute")
    # x scaled by 8 / 4 = 2: 1.5 -> 3.0 and 3.0 -> 6.0, integer flat coordinates fill one pixel each
    expect = np.zeros((8, 4))
    expect[3, 2] += 0.5      # (1.5, 2.0) -> (3.0, 2.0)
    expect[6, 0] += 0.5      # (3.0, 0.0) -> (6.0, 0.0)
    assert np.allclose(m, expect)
    # a half-pixel offset splits evenly; wrapping carries the far corner to pixel 0
    pos2 = np.array([[[[3.75, 0.5, 1.0]]]], dtype=np.float32)
    m2 = so.paint_flat(pos2, mesh, (8, 4), 0.0, 8.0, 1.0, mode="absolute")
    e2 = np.zeros((8, 4))
    e2[7, 0] = e2[7, 1] = e2[0, 0] = e2[0, 1] = 0.25
    assert np.allclose(m2, e2)


# ---------------------------------------------------------------- one-image transform (Telephoto)
def test_image_positions_identity_rotation_and_drift():
    rng = np.random.default_rng(5)
    mesh, box, obs = (8, 8, 8), (80.0, 80.0, 80.0), np.array([40.0, 40.0, 0.0])
    dx = rng.normal(0.0, 2.0, (4, 4, 4, 3))
    g = dx.reshape(-1, 3) + orc._lattice((4, 4, 4), mesh)
    assert np.allclose(so.image_positions(dx, mesh, box, obs), g, rtol=0, atol=1e-12)  # unwrapped
    cell = 10.0
    X = g * cell - obs
    for code, (perm, sign) in {7: ((0, 2, 1), (1, 1, -1)), 46: ((2, 0, 1), (-1, -1, -1))}.items():
        Y = np.stack([sign[c] * X[:, perm[c]] for c in range(3)], axis=1) + np.array([0.0, 0.0, 160.0])
        got = so.image_positions(dx, mesh, box, obs, code=code, shift=(0.0, 0.0, 160.0))
        assert np.allclose(got, (Y + obs) / cell, atol=1e-12)

    class ConstDrift:
        def factor(self, r):
            return 0.25 * np.ones_like(r)

    v = rng.normal(0.0, 1.0, (4, 4, 4, 3))
    got = so.image_positions(dx, mesh, box, obs, code=7, shift=(0.0, 0.0, 160.0), vel=v, drift=ConstDrift())
    Vt = np.stack([[1, 1, -1][c] * v.reshape(-1, 3)[:, (0, 2, 1)[c]] for c in range(3)], axis=1)
    Y = np.stack([[1, 1, -1][c] * X[:, (0, 2, 1)[c]] for c in range(3)], axis=1) + np.array([0.0, 0.0, 160.0])
    assert np.allclose(got, (Y + 0.25 * Vt + obs) / cell, atol=1e-12)


def test_flat_paint_vjp_matches_differences():
    rng = np.random.default_rng(8)
    mesh, flat = (8, 8, 8), (12, 10)
    for mode, pos in (("relative", rng.normal(0.0, 1.2, (4, 4, 4, 3))),
                      ("absolute", rng.uniform(0.0, 8.0, (4, 4, 4, 3)))):
        w = rng.uniform(0.5, 1.5, 64)
        mb = rng.standard_normal(flat)
        zlo, zhi, scale = 1.3, 6.7, 0.37
        g = so.paint_flat_vjp(pos, mb, mesh, zlo, zhi, scale, mode=mode, weight=w)
        assert np.all(g[:, 2] == 0.0)
        eps = 1e-6
        for _ in range(3):
            d = rng.standard_normal(pos.shape)
            d[..., 2] = 0.0  # the slab selection is a step: keep z fixed
            fp = (so.paint_flat(pos + eps * d, mesh, flat, zlo, zhi, scale, mode=mode, weight=w) * mb).sum()
            fm = (so.paint_flat(pos - eps * d, mesh, flat, zlo, zhi, scale, mode=mode, weight=w) * mb).sum()
            fd = (fp - fm) / (2 * eps)
            assert (g * d.reshape(-1, 3)).sum() == pytest.approx(fd, rel=1e-6, abs=1e-9)
