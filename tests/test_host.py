"""Host-side logic against the oracle's independent restatement (no GPU): background cosmology,
the float32 time grid, the solvers' kick/drift factors, lightcone geometry (mirrors the
reference's tests/test_equal_vol_shells.py invariants), fields validation."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import nbody_oracle as nbo

PARAMS = [dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8),
          dict(Omega_c=0.30, Omega_b=0.05, h=0.7, w0=-0.9, wa=0.1),
          dict(Omega_c=0.95, Omega_b=0.05)]


@pytest.fixture(scope="module")
def cosmo(pkg):
    return pkg.cosmology


@pytest.mark.parametrize("p", PARAMS)
def test_growth_matches_oracle(cosmo, p):
    c = cosmo.Cosmology(**p)
    o = nbo.Cosmo(**p)
    a = np.array([0.05, 0.1, 0.2317, 0.5, 0.77, 1.0])
    assert np.allclose(cosmo.growth_factor(c, a), o.growth_factor(a), rtol=2e-9)
    assert np.allclose(cosmo.growth_rate(c, a), o.growth_rate(a), rtol=2e-9)
    assert np.allclose(cosmo.Gf(c, a), o.Gf(a), rtol=2e-9)
    assert np.allclose(cosmo.gp(c, a), o.gp(a), rtol=2e-9)
    assert np.allclose(cosmo.dGfa(c, a), o.dGfa(a), rtol=1e-6)
    assert np.allclose(cosmo.E(c, a), o.E(a), rtol=1e-14)


@pytest.mark.parametrize("p", PARAMS[:2])
def test_distances_match_oracle(cosmo, p):
    c = cosmo.Cosmology(**p)
    o = nbo.Cosmo(**p)
    a = np.array([0.3, 0.5, 0.8, 0.95, 1.0])
    chi = cosmo.radial_comoving_distance(c, a)
    assert np.allclose(chi, o.chi(a), rtol=1e-9, atol=1e-9)
    assert np.allclose(cosmo.a_of_chi(c, chi[:-1]), a[:-1], rtol=1e-10)


def test_power_spectrum_matches_oracle(cosmo):
    p = PARAMS[0]
    c = cosmo.Cosmology(**p)
    o = nbo.Cosmo(**p)
    k = np.logspace(-4, 1, 128)
    assert np.allclose(cosmo.linear_matter_power(c, k), nbo.linear_matter_power(o, k), rtol=1e-6)


@pytest.mark.parametrize("t0,ts,dt0,t1", [(0.1, [1.0], 0.09, 1.0), (0.1, [0.4, 0.55, 0.7, 0.9, 1.0], 0.05, 1.0),
                                          (0.2, [1.0], 0.8 / 79, 1.0), (0.1, [0.95, 0.3], 0.05, 1.0)])
def test_time_grid_matches_oracle(pkg, t0, ts, dt0, t1):
    a = pkg.pm.time_grid(t0, ts, dt0, t1)
    b = nbo.time_grid(t0, ts, dt0, t1)
    assert len(a) == len(b)
    for ea, eb in zip(a, b):
        assert ea[0] == eb[0]
        assert np.float32(ea[1]) == np.float32(eb[1]) and np.float32(ea[2]) == np.float32(eb[2])


@pytest.mark.parametrize("p", PARAMS[:2])
def test_solver_factors_match_oracle(pkg, cosmo, p):
    c = cosmo.Cosmology(**p)
    o = nbo.Cosmo(**p)
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="particles"))
    kdk = pkg.EfficientDriftDoubleKick(interp_kernel=interp)
    kkd = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    t0, dt0 = 0.1, 0.05
    assert kdk.init_op(t0, t0 + dt0, dt0, c).kick == pytest.approx(nbo.kdk_init_coeff(o, t0, t0 + dt0), rel=1e-6)
    for _, ta, tb in [e for e in nbo.time_grid(t0, [0.43, 1.0], dt0, 1.0) if e[0] == "step"]:
        op = kdk.step_op(ta, tb, dt0, t0, c)
        cd, ck = nbo.kdk_coeffs(o, ta, tb, dt0)
        assert op.pre == pytest.approx(cd, rel=1e-6) and op.kick == pytest.approx(ck, rel=1e-6) and op.post is None
        op = kkd.step_op(ta, tb, dt0, t0, c)
        ck, cd = nbo.kkd_coeffs(o, ta, tb, dt0, t0)
        assert op.kick == pytest.approx(ck, rel=1e-6) and op.post == pytest.approx(cd, rel=1e-6) and op.pre is None
        r = kkd.reverse_op(ta, tb, dt0, t0, c)
        assert r.pre == -op.post and r.kick == -op.kick


# ---------------------------------------------------------------- lightcone geometry
# mirrors tests/test_equal_vol_shells.py:37-150 of the reference
@pytest.mark.parametrize("R,N,wmin", [(500.0, 10, 20.0), (1000.0, 8, 50.0), (2000.0, 20, 50.0)])
def test_equal_vol_shell_invariants(pkg, R, N, wmin):
    e = pkg.utils._compute_equal_vol_shells(R, N, wmin)
    assert e.shape == (N + 1,) and e[0] == 0.0 and e[-1] == pytest.approx(R)
    w = np.diff(e)
    assert np.all(w > 0) and np.all(w >= wmin - 1e-9)
    inner = w[: np.argmax(np.isclose(w, wmin))] if np.any(np.isclose(w, wmin)) else w
    vols = np.diff(e[: len(inner) + 1] ** 3)
    if len(vols) > 1:
        assert np.allclose(vols, vols[0], rtol=1e-9)
    assert np.sum(np.diff(e**3)) == pytest.approx(R**3)


def test_equal_vol_rejects_impossible(pkg):
    with pytest.raises(ValueError):
        pkg.utils._compute_equal_vol_shells(100.0, 10, 20.0)


def test_lightcone_shells_far_to_near(pkg, cosmo):
    import torch

    c = cosmo.Cosmology(Omega_c=0.25, Omega_b=0.0486)
    f = pkg.DensityField(array=torch.zeros(8, 8, 8), mesh_size=(8, 8, 8), box_size=(1000.0,) * 3)
    r, a = pkg.compute_lightcone_shells(c, f, nb_shells=5)
    assert np.all(np.diff(r) < 0) and np.all(np.diff(a) > 0)
    assert r[0] + (r[0] - r[1]) / 2 == pytest.approx(500.0)
    w = pkg.distances(r, 0.0)
    assert np.allclose(w, 100.0)
    assert np.allclose(pkg.centers(w[::-1], 0.0), r[::-1])


def test_field_validation(pkg):
    import torch

    with pytest.raises(ValueError):
        pkg.ParticleField(array=torch.zeros(4, 4, 4), mesh_size=(4, 4, 4), box_size=(1.0,) * 3,
                          unit=pkg.PositionUnit.GRID_RELATIVE)
    with pytest.raises(ValueError):
        pkg.DensityField(array=torch.zeros(4, 4, 4), mesh_size=(4, 4, 4), box_size=(1.0,) * 3,
                         observer_position=(1.5, 0.5, 0.5))
    with pytest.raises(TypeError):
        pkg.ParticleField(array=torch.zeros(4, 4, 4, 3), mesh_size=(4, 4, 4), box_size=(1.0,) * 3)
    f = pkg.DensityField(array=torch.zeros(4, 4, 4), mesh_size=(4, 4, 4), box_size=(100.0, 200.0, 300.0),
                         observer_position=(0.5, 0.5, 0.0))
    assert f.max_comoving_radius == pytest.approx(50.0)
    assert f.observer_position_mpc == (50.0, 100.0, 0.0)


def test_vectorised_kkd_coefficients_equal_step_op(pkg):
    """step_coeffs (all steps at once, used by the differentiable path) reproduces step_op exactly."""
    cosmo = pkg.Cosmology(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")))
    grid = [(ev[1], ev[2]) for ev in pkg.pm.time_grid(0.1, [0.37, 1.0], 0.09, 1.0) if ev[0] == "step"]
    vec = solver.step_coeffs(grid, 0.09, 0.1, cosmo)
    raw = solver.step_coeffs(grid, 0.09, 0.1, cosmo, raw=True)
    for i, (a, b) in enumerate(grid):
        op = solver.step_op(a, b, 0.09, 0.1, cosmo)
        opr = solver.step_op(a, b, 0.09, 0.1, cosmo, raw=True)
        assert vec[i, 0] == op.kick and vec[i, 1] == op.post
        # float64 values (used only for parameter derivatives) agree to the last bits
        assert abs(raw[i, 0] - opr.kick) <= 4e-16 * abs(opr.kick) and abs(raw[i, 1] - opr.post) <= 4e-16 * abs(opr.post)


def test_vectorised_kdk_ops_equal_step_op(pkg):
    cosmo = pkg.Cosmology(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
    solver = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")))
    grid = [(ev[1], ev[2]) for ev in pkg.pm.time_grid(0.1, [0.41, 0.73, 1.0], 0.05, 1.0) if ev[0] == "step"]
    for (a, b), op in zip(grid, solver.step_ops(grid, 0.05, 0.1, cosmo)):
        ref = solver.step_op(a, b, 0.05, 0.1, cosmo)
        assert op.pre == ref.pre and op.kick == ref.kick and op.post == ref.post


def test_native_growth_table_equals_python_loop(pkg):
    """The C++ growth-table loop (libpmx_host.so) reproduces the Python RK4 loop bit for bit."""
    c = pkg.cosmology
    if c._host_lib() is None:
        pytest.skip("libpmx_host.so not built")
    for p in ({}, {"Omega_k": 0.02, "w0": -0.9, "wa": 0.1}):
        bg = c._background(c.Cosmology(Omega_c=0.27, **p))
        xs_n, D_n, G_n, dG_n = c._growth_table_bg(bg, True)
        xs_p, D_p, G_p, dG_p = c._growth_table_bg(bg, False)
        assert np.array_equal(D_n, D_p) and np.array_equal(G_n, G_p) and np.array_equal(dG_n, dG_p)


def test_telephoto_rotations_and_shifts(pkg):
    """TelephotoInterp (src/jax_fli/pm/interp.py:507-580): the shell index advances at snapshot
    times; rotation_idx counts the shells beyond the box radius (from -1); the image is symmetry
    rotation_idx mod 47 of the 47 non-identity ones, shifted along z by the radius times
    floor(r_center / radius); shells whose centre is within the radius are painted from the box."""
    ts = np.array([0.3, 0.5, 0.7, 0.9], dtype=np.float32)
    r_c = np.array([260.0, 160.0, 70.0, 40.0])
    tel = pkg.TelephotoInterp(painting=pkg.PaintingOptions(target="flat")).update_geometry(
        ts=ts, r_centers=r_c, density_widths=np.array([60.0, 60.0, 60.0, 30.0]), max_comoving_distance=50.0)
    st = tel.init()
    assert st.rotation_idx == -1 and st.shell_idx == -1 and list(st.rotations) == list(range(1, 48))
    seen = []
    for t in ts:
        st = tel.advance(tel.advance(st, float(t) - 0.05), float(t))  # a non-snapshot time changes nothing
        seen.append((st.shell_idx, st.rotation_idx, tel.images(st, None)))
    assert [s[:2] for s in seen] == [(0, 0), (1, 1), (2, 2), (3, 2)]
    codes = [None if s[2] is None else int(s[2][0][0]) for s in seen]
    assert codes == [1, 2, 3, None]  # rotation_idx 0, 1, 2 -> entries 1, 2, 3 of the 48-entry list
    assert np.array_equal(seen[0][2][1], [[0.0, 0.0, 250.0]]) and np.array_equal(seen[1][2][1], [[0.0, 0.0, 150.0]])
    assert np.array_equal(seen[2][2][1], [[0.0, 0.0, 50.0]])
    for shell in range(3):
        assert tel.rotation_code(shell) == codes[shell]
    for t in ts[::-1]:  # rewinding undoes every advance
        st = tel.rewind(st, float(t))
    assert st.shell_idx == -1 and st.rotation_idx == -1
    many = tel.update_geometry(ts=ts, r_centers=np.full(50, 200.0), density_widths=np.full(50, 10.0),
                               max_comoving_distance=50.0)
    assert many.rotation_code(46) == 47 and many.rotation_code(47) == 1  # the index wraps mod 47
    with pytest.raises(ValueError):
        pkg.TelephotoInterp(painting=pkg.PaintingOptions(target="density")).init()
    with pytest.warns(UserWarning):
        pkg.TelephotoInterp(painting=pkg.PaintingOptions(target="spherical")).init()


def test_resolve_ts_geometry_cases(pkg, cosmo):
    """The four input forms of the lightcone geometry and its refusals (pm/_resolve_geometry.py:12-114
    of the reference), with distances from the oracle's independent quadrature."""
    import torch

    from jax_fli_amd.pm._resolve_geometry import resolve_ts_geometry
    from jax_fli_amd.painting import PaintingOptions

    p = dict(Omega_c=0.25, Omega_b=0.0486)
    c, o = cosmo.Cosmology(**p), nbo.Cosmo(**p)
    f = pkg.DensityField(array=torch.zeros(8, 8, 8), mesh_size=(8, 8, 8), box_size=(1000.0,) * 3)
    dens, sph = PaintingOptions(target="density"), PaintingOptions(target="spherical")
    # shell count: far to near, widths = spacing, float32 times
    ts, r, w, lc = resolve_ts_geometry(c, f, dens, nb_shells=4)
    assert lc and ts.dtype == np.float32 and np.allclose(r, [437.5, 312.5, 187.5, 62.5]) and np.allclose(w, 125.0)
    assert np.allclose(o.chi(ts.astype(np.float64)), r, rtol=2e-6)
    _, _, w2, _ = resolve_ts_geometry(c, f, dens, nb_shells=4, density_widths=30.0)
    assert np.array_equal(w2, np.full(4, 30.0))
    # one time: a snapshot, width 0, or the given scalar width
    ts, r, w, lc = resolve_ts_geometry(c, f, dens, ts=0.5)
    assert not lc and ts.shape == (1,) and np.allclose(r, o.chi(np.array([0.5])), rtol=1e-8) and w.tolist() == [0.0]
    ts, r, w, lc = resolve_ts_geometry(c, f, sph, ts=[0.5], density_widths=[[80.0]])
    assert not lc and w.tolist() == [80.0]
    # several times: shell centres
    a = np.array([0.6, 0.8, 0.95])
    ts, r, w, lc = resolve_ts_geometry(c, f, dens, ts=a)
    assert lc and np.allclose(r, o.chi(a), rtol=1e-8) and np.allclose(w, pkg.distances(r, 0.0))
    # (2, n): near/far edges, centre in distance, time back from the centre
    e = np.array([[0.9, 0.7], [0.8, 0.6]])
    ts, r, w, lc = resolve_ts_geometry(c, f, sph, ts=e)
    rn, rf = o.chi(e[0]), o.chi(e[1])
    assert lc and np.allclose(r, 0.5 * (rn + rf), rtol=1e-8) and np.allclose(w, np.abs(rf - rn), rtol=1e-8)
    assert np.allclose(o.chi(ts.astype(np.float64)), r, rtol=2e-5)
    # refusals
    for kw in (dict(), dict(ts=0.5, nb_shells=3), dict(ts=0.5, density_widths=[1.0, 2.0]),
               dict(ts=[0.5, 0.6], density_widths=[1.0, 2.0, 3.0]), dict(nb_shells=3, density_widths=[1.0, 2.0]),
               dict(ts=np.ones((3, 2)))):
        with pytest.raises(ValueError):
            resolve_ts_geometry(c, f, dens, **kw)
    with pytest.raises(ValueError):
        resolve_ts_geometry(c, f, sph, ts=0.5)


@pytest.mark.parametrize("oc,s8,n", [(0.25, 0.8, 64), (0.3137, 0.7123, 96), (0.21, 0.88, 192)])
def test_native_request_plan_equals_python(pkg, oc, s8, n):
    """The request server's native host preparation (pmx_host_request_plan: P(k) table, LPT factors, the
    fused KDK coefficients) against the Python layer's forward_ops: coefficients and LPT factors bitwise
    (float32, as the device takes them), the P(k) table to float64 rounding."""
    import ctypes
    import os

    path = os.path.join(os.path.dirname(pkg.__file__), "libpmx_host.so")
    if not os.path.exists(path):
        pytest.skip("libpmx_host.so not built")
    lib = ctypes.CDLL(path)
    d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    lib.pmx_host_request_plan.argtypes = [vp, i, d, d, d, d, vp, vp, vp, vp, i, vp, vp]
    req = pkg.serving.SimRequest(rid=0, mesh=n, seed=1, Omega_c=oc, sigma8=s8)
    c = req.cosmology()
    params = np.array([c.Omega_c, c.Omega_b, c.h, c.n_s, c.sigma8, c.Omega_k, c.w0, c.wa])
    kt, pt = np.zeros(128), np.zeros(128)
    lpt, ops = np.zeros(2, np.float32), np.zeros((64, 3), np.float32)
    nops, fs = ctypes.c_int(), ctypes.c_double()
    rc = lib.pmx_host_request_plan(params.ctypes.data, n, req.box_size(), req.a0, req.a1, req.dt0, kt.ctypes.data,
                                   pt.ctypes.data, lpt.ctypes.data, ops.ctypes.data, 64, ctypes.byref(nops),
                                   ctypes.byref(fs))
    assert rc == 0
    D, vf, k, ptab, pops, fsp = pkg.serving.forward_ops(req)
    np.testing.assert_allclose(kt, k, rtol=1e-15)
    np.testing.assert_allclose(pt, ptab, rtol=2e-15)
    assert lpt[0] == np.float32(D) and lpt[1] == np.float32(vf) and fs.value == fsp
    ref = np.array([[np.nan if o.pre is None else o.pre, o.kick, np.nan if o.post is None else o.post] for o in pops],
                   np.float32)
    assert nops.value == len(pops) and np.array_equal(ops[: nops.value], ref, equal_nan=True)


@pytest.mark.parametrize("x64", [False, True])
def test_jax_prng_matches_oracle(pkg, x64):
    """jax_fli_amd.prng (vectorised uint32 NumPy, JAX's multiplier form of randint's reduction) against
    the oracle's scalar restatement (the 2w-bit number reduced directly)."""
    from oracle import sky_oracle as so

    prng = pkg.prng
    rng = np.random.default_rng(5)
    k = rng.integers(0, 2**32, size=(4, 2), dtype=np.uint64)
    c = rng.integers(0, 2**32, size=(4, 2), dtype=np.uint64)
    for kk, cc in zip(k, c):
        y0, y1 = prng.threefry2x32((kk[0], kk[1]), np.array([cc[0]]), np.array([cc[1]]))
        assert (int(y0[0]), int(y1[0])) == so.threefry2x32(int(kk[0]), int(kk[1]), int(cc[0]), int(cc[1]))
    for seed in (0, 42, (5 << 32) + 9):
        key = prng.prng_key(seed)
        assert tuple(int(w) for w in key) == so.jax_key(seed)
        assert [tuple(int(w) for w in r) for r in prng.split(key, 3)] == so.jax_split(so.jax_key(seed), 3)
        for n, lo, hi in ((27, 0, 47), (125, 0, 47), (40, -3, 1000)):
            got = prng.randint(key, n, lo, hi, x64=x64)
            assert got.dtype == (np.int64 if x64 else np.int32)
            assert np.array_equal(got, so.jax_randint(seed, n, lo, hi, x64=x64))


def test_onion_tiler_rotations_are_the_reference_draw(pkg):
    """OnionTiler's 27 symmetries: jax.random.randint(PRNGKey(42), (27,), 0, 47) into the 47
    non-identity entries (src/jax_fli/pm/interp.py:333-336); layers=2 draws 125."""
    from oracle import sky_oracle as so

    sph = pkg.PaintingOptions(target="spherical")
    assert np.array_equal(pkg.OnionTiler(painting=sph).init().rotations, so.jax_randint(42, 27, 0, 47) + 1)
    assert np.array_equal(pkg.OnionTiler(painting=sph, x64=True).init().rotations,
                          so.jax_randint(42, 27, 0, 47, x64=True) + 1)
    assert np.array_equal(pkg.OnionTiler(painting=sph, layers=2).init().rotations,
                          so.jax_randint(42, 125, 0, 47) + 1)
    assert not pkg.OnionTiler(painting=sph, rotate=False).init().rotations.any()
