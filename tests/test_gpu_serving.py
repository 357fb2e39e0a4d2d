"""The native continuous-batching server (pmx_server_*): every result is the same bit for bit alone and
in a mix of sizes and cosmologies (batched steps replayed from HIP graphs, slots of a class at
different steps), equals the same simulation through the public API (the per-request eager path,
whose host tables come from the Python layer) to float32 rounding of the P(k) table, admission is
bounded by the slots the HBM budget gives and follows arrival order per size, and latencies are
reported. The serving workload's sizes against the float64 oracle are in test_gpu_configs.py."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _reqs(pkg, meshes, arrival=0.0, seed0=100):
    rng = np.random.default_rng(70)
    return [pkg.serving.SimRequest(rid=i, mesh=int(n), seed=seed0 + i, Omega_c=float(rng.uniform(0.2, 0.35)),
                                   sigma8=float(rng.uniform(0.7, 0.9)), arrival=arrival) for i, n in enumerate(meshes)]


def test_results_same_alone_and_in_the_mix(pkg, cuda):
    meshes = [32, 64, 32, 64, 32, 64, 32, 96]
    reqs = _reqs(pkg, meshes)
    srv = pkg.serving.Server(device=cuda, meshes=(32, 64, 96), slots={32: 3, 64: 2, 96: 1})
    try:
        alone = {}
        for r in reqs:
            res, _ = srv.serve([r])
            alone[r.rid] = res[r.rid]
        # staggered arrivals: slots join a class while the others are mid-run
        for j, r in enumerate(reqs):
            r.arrival = 0.0003 * j
        mixed, stats = srv.serve(reqs)
        assert stats["requests"] == len(reqs) and stats["peak_active"] > 1
        for r in reqs:
            assert torch.equal(mixed[r.rid].displacement, alone[r.rid].displacement), r.rid
            assert torch.equal(mixed[r.rid].density, alone[r.rid].density), r.rid
            assert float(mixed[r.rid].density.double().sum()) == pytest.approx(r.mesh**3, rel=1e-6)
    finally:
        srv.close()


@pytest.mark.parametrize("n", [64, 96])
def test_result_equals_public_api(pkg, cuda, n):
    r = _reqs(pkg, [n], seed0=7 + n)[0]
    srv = pkg.serving.Server(device=cuda, meshes=(n,), slots=2)
    try:
        res, _ = srv.serve([r])
    finally:
        srv.close()
    c = r.cosmology()
    box = r.box_size()
    ic = pkg.gaussian_initial_conditions(r.seed, (n,) * 3, (box,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=r.a0)
    solver = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="particles")))
    out = pkg.nbody(c, dx, p, t0=r.a0, t1=r.a1, dt0=r.dt0, ts=[r.a1], solver=solver)
    got, ref = res[r.rid].displacement.double(), out.array.double()
    err = float((got - ref).norm() / ref.norm())
    same = bool(torch.equal(res[r.rid].displacement, out.array))
    print(f"{n}^3: server vs public API displacement rel L2 {err:.2e} (bitwise equal: {same})")
    assert err < 1e-6
    rho = pkg.engine.paint(out.array, (n,) * 3)
    assert float((res[r.rid].density.double() - rho.double()).norm() / rho.double().norm()) < 1e-6


def test_slots_bound_admission(pkg, cuda):
    srv = pkg.serving.Server(device=cuda, meshes=(64,), slots=2)
    try:
        assert srv.slots(64) == 2 and srv.slot_bytes(64) > 64**3 * 60
        reqs = _reqs(pkg, [64] * 6)
        res, stats = srv.serve(reqs)
        assert stats["requests"] == 6 and stats["peak_active"] <= 2
        order = sorted(res.values(), key=lambda x: srv.times(x.request_id)[1])
        assert [x.rid for x in order] == list(range(6))  # admitted in arrival order
    finally:
        srv.close()
    # the budget fixes the slots: room for about three 64^3 slots
    one = pkg.serving.Server(device=cuda, meshes=(64,), slots=1).slot_bytes(64)
    srv = pkg.serving.Server(device=cuda, meshes=(64,), hbm_budget=int(3.5 * one))
    try:
        assert srv.slots(64) == 3
    finally:
        srv.close()


def test_poisson_stream_reports_latency(pkg, cuda):
    reqs = pkg.serving.poisson_requests(24, rate_per_s=400.0, seed=3, meshes=(32, 64))
    srv = pkg.serving.Server(device=cuda, meshes=(32, 64), slots=4)
    try:
        res, stats = srv.serve(reqs)
    finally:
        srv.close()
    assert stats["requests"] == 24 and stats["p99_ms"] >= stats["p50_ms"] > 0
    for r in reqs:
        assert res[r.rid].finished >= r.arrival and res[r.rid].admitted >= r.arrival - 1e-9


def test_bad_requests_are_refused(pkg, cuda):
    srv = pkg.serving.Server(device=cuda, meshes=(32,), slots=1)
    try:
        with pytest.raises(pkg._lib.PmxError):
            srv.submit(_reqs(pkg, [64])[0])  # no class for 64^3
        bad = _reqs(pkg, [32])[0]
        bad.dt0 = -1.0
        with pytest.raises(pkg._lib.PmxError):
            srv.submit(bad)
    finally:
        srv.close()
