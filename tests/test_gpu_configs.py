"""The benchmark workloads at (or near) their real sizes, against the float64 oracle, through the
public API. Every chain starts from the seed: the oracle makes its own white noise (its Philox
restatement), its own EH98 P(k) with its own sigma8 quadrature, its own growth ODE and coefficients; only
the seed, the cosmology and the shells' float32 times are shared. Each result is printed beside the
oracle's own float32-versus-float64 error. Bounds: particle positions <= 1e-4 (relative L2 and largest
displacement error over the grid), meshes <= 1e-5, P(k) within 0.1% below k_Nyq/2.

* Single-GPU forward (256^3 in the benchmark; here 128^3 particles on a 128^3 mesh in the same 1000 Mpc/h
  box, so the oracle's float64 run finishes in about a minute): LPT1 at a = 0.1, EfficientDriftDoubleKick
  dt0 = 0.05 to a = 1, 4 HEALPix shells at nside 256 (bilinear, equal comoving width to z = 0.5, ts given
  as near/far scale-factor pairs), final positions, density and P(k).
* Serving: a Poisson mix over {64, 96, 128, 192}^3 through the continuous-batching server; the 64^3,
  96^3 and 192^3 results (96 and 192 run the radix-3 FFT passes) against the oracle, and every size
  bitwise the same alone and in the mix.
* Differentiable (128^3, box 500 Mpc/h, ReversibleDoubleKickDrift 10 steps, L2 density loss against a
  second seed): the reverse-adjoint gradient along 4 noise directions and in (Omega_c, sigma8) against
  float64 central differences of the oracle's model, committed in tests/golden/config3_gradient_fd.json
  by tests/golden/make_config3_fd.py (the oracle needs minutes for them).
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest
import torch

from oracle import nbody_oracle as nbo
from oracle import pm_oracle as orc
from oracle import sky_oracle as so

pytestmark = pytest.mark.gpu

KT = np.logspace(-4, 1, 128)  # the colouring's P(k) table (initial.py:171-173)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_ics(o, n, box, seed):
    """(dx, p) of the oracle from the seed: Philox noise, its own EH98 P(k), LPT1 at a = 0.1."""
    noise = orc.white_noise((n,) * 3, seed)
    ic = nbo.color_field(noise, (box,) * 3, KT, nbo.linear_matter_power(o, KT))
    return nbo.lpt1(o, ic, 0.1)


# ------------------------------------------------------------------------------ single-GPU forward
def test_forward_workload_128_matches_oracle(pkg, cuda):
    P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
    n, box, a0, dt0, nside, seed = 128, 1000.0, 0.1, 0.05, 256, 0
    c, o = pkg.Cosmology(**P), nbo.Cosmo(**P)
    cos = pkg.cosmology
    rmax = float(cos.radial_comoving_distance(c, cos.z2a(0.5)))
    edges = np.linspace(0.0, rmax, 5)
    # shells far -> near, ascending scale factor; near/far edge pairs as the reference's (2, N) ts
    near, far = edges[:-1][::-1], edges[1:][::-1]
    ts2 = np.stack([cos.a_of_chi(c, near), cos.a_of_chi(c, far)])
    ic = pkg.gaussian_initial_conditions(seed, (n,) * 3, (box,) * 3, cosmo=c, nside=nside)
    dx, p = pkg.lpt(c, ic, ts=a0)
    sph = pkg.EfficientDriftDoubleKick(
        interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="spherical", scheme="bilinear")))
    lc = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=dt0, ts=ts2, solver=sph)
    maps_gpu = lc.array.cpu().numpy()
    a_gpu = np.asarray(lc.scale_factors, dtype=np.float32).reshape(-1)
    ts_res, r_c, w_c, _ = pkg.pm._resolve_geometry.resolve_ts_geometry(c, dx, sph.interp_kernel.painting, ts=ts2)
    order = np.argsort(ts_res)  # the oracle saves in ascending time
    ts_save = np.concatenate([ts_res[order], np.float32([1.0])])
    # final positions on the same step grid (it restarts at every snapshot time): particle snapshots at
    # the shell times and a = 1, the last one compared
    part = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="particles")))
    snaps = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=dt0, ts=ts_save, solver=part)
    last = int(np.nonzero(np.asarray(snaps.scale_factors, dtype=np.float32).reshape(-1) == np.float32(1.0))[0][0])
    x_fin = snaps.array[last].contiguous()
    x_gpu = x_fin.cpu().numpy()
    dx0, p0 = oracle_ics(o, n, box, seed)
    obs = (box / 2,) * 3
    results = {}
    for dtype in (np.float64, np.float32):
        maps = {}

        def save(i, t, x, v, maps=maps):
            if i < 4:
                s = order[i]
                lo, hi = r_c[s] - w_c[s] / 2, r_c[s] + w_c[s] / 2
                maps[s] = so.paint_healpix_tiled(x, (n,) * 3, (box,) * 3, obs, lo, hi, nside, scheme="bilinear")

        x, _ = nbo.run_nbody(o, dx0, p0, a0, 1.0, dt0, ts_save, solver="kdk", dtype=dtype, on_save=save)
        results[dtype] = (x, maps)
    x64, maps64 = results[np.float64]
    x32, maps32 = results[np.float32]
    # the shells: the product returns them in the order of its scale factors
    for s in range(4):
        j = int(np.nonzero(a_gpu == ts_res[s])[0][0])
        ref, got = maps64[s], maps_gpu[j]
        if not np.any(ref):
            print(f"shell {s} ({r_c[s]:.0f} Mpc/h): empty (beyond the box), product max {np.abs(got).max():.1e}")
            assert not np.any(got)
            continue
        e, e32 = rel(got, ref), rel(maps32[s], ref)
        print(f"shell {s} ({r_c[s]:.0f} +- {w_c[s] / 2:.0f} Mpc/h): map rel L2 {e:.2e} (oracle f32 {e32:.2e})")
        assert e < 1e-5
    e_pos, e_pos32 = rel(x_gpu, x64), rel(x32, x64)
    abs_err = float(np.abs(x_gpu - x64).max())
    m_gpu = pkg.engine.paint(x_fin, (n,) * 3).cpu().numpy()
    m64 = orc.cic_paint(x64, (n,) * 3, dtype=np.float64)
    e_mesh, e_mesh32 = rel(m_gpu, m64), rel(orc.cic_paint(x32, (n,) * 3, dtype=np.float64), m64)
    print(f"final displacements rel L2 {e_pos:.2e} (oracle f32 {e_pos32:.2e}), max err {abs_err:.2e} cells; "
          f"density rel L2 {e_mesh:.2e} (oracle f32 {e_mesh32:.2e})")
    assert e_pos < 1e-4 and abs_err / n < 1e-4 and e_mesh < 1e-5
    kedges = np.geomspace(2 * np.pi / box, np.pi * n / box, 33)
    k, pk = pkg.power(torch.from_numpy(m_gpu).to(cuda), box_shape=(box,) * 3, kedges=kedges)
    k_ref, pk_ref = orc.power(m64, (box,) * 3, kedges)
    sel = (k_ref < 0.5 * np.pi * n / box) & (pk_ref > 0)
    err_pk = float(np.max(np.abs(pk[sel] / pk_ref[sel] - 1)))
    print(f"P(k) below k_Nyq/2: max rel err {err_pk:.2e} over {int(sel.sum())} bins")
    assert err_pk < 1e-3


# ------------------------------------------------------------------------------ serving
def test_serving_mix_matches_oracle(pkg, cuda):
    reqs = pkg.serving.poisson_requests(10, rate_per_s=300.0, seed=17, meshes=(64, 96, 128, 192))
    sizes = {r.mesh for r in reqs}
    for m in (64, 96, 128, 192):  # make sure every size is in the mix
        if m not in sizes:
            reqs.append(pkg.serving.SimRequest(rid=len(reqs), mesh=m, seed=900 + m, Omega_c=0.27, sigma8=0.82,
                                               arrival=reqs[-1].arrival + 1e-3))
    srv = pkg.serving.Server(device=cuda, meshes=(64, 96, 128, 192), slots=4)
    mixed, stats = srv.serve(reqs)
    assert stats["requests"] == len(reqs)
    print(f"served {len(reqs)} requests, peak {stats['peak_active']} in flight, p50 {stats['p50_ms']:.1f} ms "
          f"p99 {stats['p99_ms']:.1f} ms")
    first = {}
    for r in reqs:
        first.setdefault(r.mesh, r)
    for m, r in sorted(first.items()):
        alone, _ = srv.serve([r])
        assert torch.equal(alone[r.rid].displacement, mixed[r.rid].displacement), m
        assert torch.equal(alone[r.rid].density, mixed[r.rid].density), m
    for m in (64, 96, 192):
        r = first[m]
        o = nbo.Cosmo(Omega_c=r.Omega_c, Omega_b=r.Omega_b, h=r.h, n_s=r.n_s, sigma8=r.sigma8)
        dx0, p0 = oracle_ics(o, m, r.box_size(), r.seed)
        x64, _ = nbo.run_nbody(o, dx0, p0, r.a0, r.a1, r.dt0, [r.a1], solver="kdk")
        x_gpu = mixed[r.rid].displacement.cpu().numpy()
        e_pos = rel(x_gpu, x64)
        abs_err = float(np.abs(x_gpu - x64).max())
        m64 = orc.cic_paint(x64, (m,) * 3, dtype=np.float64)
        e_mesh = rel(mixed[r.rid].density.cpu().numpy(), m64)
        print(f"{m}^3 request (Omega_c {r.Omega_c:.3f}, sigma8 {r.sigma8:.3f}): displacements rel L2 {e_pos:.2e}, "
              f"max err {abs_err:.2e} cells, density rel L2 {e_mesh:.2e}")
        assert e_pos < 1e-4 and abs_err / m < 1e-4 and e_mesh < 1e-5


# ------------------------------------------------------------------------------ differentiable
def test_differentiable_workload_gradient_matches_oracle_fd(pkg, cuda):
    with open(os.path.join(GOLDEN, "config3_gradient_fd.json")) as f:
        g = json.load(f)
    n, box, a0, t1, dt0, P = g["N"], g["box"], g["a0"], g["t1"], g["dt0"], g["params"]
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")))

    def delta(noise_t, c):
        ic = pkg.interpolate_initial_conditions(noise_t, (n,) * 3, (box,) * 3, cosmo=c)
        dx, p = pkg.lpt(c, ic, ts=a0)
        rho = pkg.nbody(c, dx, p, t0=a0, t1=t1, dt0=dt0, ts=[t1], solver=solver, adjoint="reverse").array.double()
        return rho / rho.mean() - 1.0

    def noise_of(seed):
        return torch.from_numpy(np.random.default_rng(seed).standard_normal((n,) * 3).astype(np.float32)).to(cuda)

    with torch.no_grad():
        target = delta(noise_of(g["seed_target"]), pkg.Cosmology(**P))
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    s8 = torch.tensor(P["sigma8"], dtype=torch.float64, requires_grad=True)
    noise = noise_of(g["seed_noise"]).requires_grad_(True)
    L = ((delta(noise, pkg.Cosmology(**{**P, "Omega_c": Oc, "sigma8": s8})) - target) ** 2).sum()
    L.backward()
    eL = abs(float(L) - g["loss"]) / g["loss"]
    print(f"loss {float(L):.8e} vs oracle {g['loss']:.8e} (rel {eL:.1e})")
    assert eL < 1e-4
    grad = noise.grad.double().cpu().numpy()
    # a random unit-variance direction's derivative has rms |grad|: errors are measured against the larger
    # of that and the derivative itself (a direction nearly orthogonal to grad gives a small value)
    gnorm = float(np.linalg.norm(grad))
    check = g["noise_fd_check"]
    print(f"oracle fd direction 0: eps {g['eps']:.0e} {g['noise_fd'][0]:.6e}, eps {check['eps']:.0e} {check['fd']:.6e}")
    for j, fd in enumerate(g["noise_fd"]):
        d = np.random.default_rng(g["seed_dir"] + j).standard_normal(grad.shape)
        got = float((grad * d).sum())
        scale = max(abs(fd), gnorm)
        print(f"noise direction {j}: adjoint {got:.6e} oracle fd {fd:.6e} error {abs(got - fd) / scale:.1e} of "
              f"max(|fd|, |grad| = {gnorm:.3e})")
        assert abs(got - fd) / scale < 2e-3
    for name, t in (("Omega_c", Oc), ("sigma8", s8)):
        fd = g[name + "_fd"]
        got = float(t.grad)
        print(f"{name}: adjoint {got:.6e} oracle fd {fd:.6e} rel {abs(got - fd) / abs(fd):.1e}")
        assert abs(got - fd) / abs(fd) < 2e-3
