"""Continuous-batching server: every result equals the same simulation run alone and the same
simulation through the public API (bit for bit: the kernels are deterministic), admission respects
the HBM budget, and arrivals are served in order with latencies reported."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _reqs(pkg, meshes, arrival=0.0):
    rng = np.random.default_rng(70)
    return [pkg.serving.SimRequest(rid=i, mesh=int(n), seed=100 + i, Omega_c=float(rng.uniform(0.2, 0.35)),
                                   sigma8=float(rng.uniform(0.7, 0.9)), arrival=arrival) for i, n in enumerate(meshes)]


def test_results_same_alone_and_in_the_mix(pkg, cuda):
    meshes = [32, 64, 32, 64, 32, 64, 32]
    reqs = _reqs(pkg, meshes)
    srv = pkg.serving.Server(device=cuda, max_active=8)
    alone = {}
    for r in reqs:
        res, _ = srv.serve([r])
        alone[r.rid] = res[r.rid]
    mixed, stats = srv.serve(reqs)
    assert stats["requests"] == len(reqs) and stats["peak_active"] > 1
    for r in reqs:
        assert torch.equal(mixed[r.rid].displacement, alone[r.rid].displacement)
        assert torch.equal(mixed[r.rid].density, alone[r.rid].density)
        assert float(mixed[r.rid].density.double().sum()) == pytest.approx(r.mesh**3, rel=1e-6)


def test_result_equals_public_api(pkg, cuda):
    r = _reqs(pkg, [64])[0]
    res, _ = pkg.serving.Server(device=cuda).serve([r])
    c = r.cosmology()
    n, box = r.mesh, r.box_size()
    ic = pkg.gaussian_initial_conditions(r.seed, (n,) * 3, (box,) * 3, cosmo=c)
    dx, p = pkg.lpt(c, ic, ts=r.a0)
    solver = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="particles")))
    out = pkg.nbody(c, dx, p, t0=r.a0, t1=r.a1, dt0=r.dt0, ts=[r.a1], solver=solver)
    assert torch.equal(res[r.rid].displacement, out.array)


def test_hbm_budget_limits_admission(pkg, cuda):
    srv = pkg.serving.Server(device=cuda, max_active=16)
    srv.prewarm({64: 1})
    one = srv._slot_size(64)
    srv.hbm_budget = int(2.5 * one)
    reqs = _reqs(pkg, [64] * 6)
    res, stats = srv.serve(reqs)
    assert stats["requests"] == 6 and stats["peak_active"] <= 2
    order = sorted(res.values(), key=lambda x: x.admitted)
    assert [x.rid for x in order] == list(range(6))  # admitted in arrival order


def test_poisson_stream_reports_latency(pkg, cuda):
    reqs = pkg.serving.poisson_requests(12, rate_per_s=400.0, seed=3, meshes=(32, 64))
    res, stats = pkg.serving.Server(device=cuda).serve(reqs)
    assert stats["requests"] == 12 and stats["p99_ms"] >= stats["p50_ms"] > 0
    for r in reqs:
        assert res[r.rid].finished >= r.arrival and res[r.rid].admitted >= r.arrival - 1e-9
