"""Plan buffers and launch coverage, through the C ABI.

* pmx_plan_reserve: a plan made with PMX_PLAN_NO_RESERVE refuses a paint (code 4, before any launch)
  until its sort buffers are reserved, reports the reserved bytes in pmx_plan_workspace_bytes, and the
  map accumulator of the HEALPix paint behaves the same way.
* pmx_set_grid_cap: with the grid-stride kernels capped at a few workgroups (so every workgroup loops
  over many particles) the image transform and its adjoint, white noise and LPT give bitwise the
  results of the default launch, Born to float64 rounding of its partial means.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_reserve_before_paint(pkg, cuda):
    lib = pkg._lib.load()
    n = 32
    plan = pkg._lib.Plan((n, n, n), cuda, pkg._lib.PLAN_NO_RESERVE)
    base = plan.workspace_bytes
    x = torch.zeros((n, n, n, 3), dtype=torch.float32, device=cuda)
    rho = torch.empty((n, n, n), dtype=torch.float32, device=cuda)
    rc = lib.pmx_paint_cic(plan.handle, x.data_ptr(), n, n, n, 0, None, 1.0, rho.data_ptr(), 0,
                           pkg._lib.stream_ptr())
    assert rc == 4
    assert b"pmx_plan_reserve" in lib.pmx_last_error()
    pkg._lib.check(lib.pmx_plan_reserve(plan.handle, n, n, n, 0), "reserve")
    grown = plan.workspace_bytes - base
    assert grown >= 2 * 4 * n**3
    pkg._lib.check(lib.pmx_paint_cic(plan.handle, x.data_ptr(), n, n, n, 0, None, 1.0, rho.data_ptr(), 0,
                                     pkg._lib.stream_ptr()), "paint")
    torch.cuda.synchronize()
    assert torch.equal(rho, torch.ones_like(rho))
    # a smaller grid fits the reservation; a reserve that needs nothing new allocates nothing
    pkg._lib.check(lib.pmx_plan_reserve(plan.handle, n // 2, n // 2, n // 2, 0), "reserve smaller")
    assert plan.workspace_bytes - base == grown
    xs = torch.zeros((n // 2, n // 2, n // 2, 3), dtype=torch.float32, device=cuda)
    pkg._lib.check(lib.pmx_paint_cic(plan.handle, xs.data_ptr(), n // 2, n // 2, n // 2, 0, None, 1.0,
                                     rho.data_ptr(), 0, pkg._lib.stream_ptr()), "paint sub-lattice")
    torch.cuda.synchronize()
    assert float(rho.sum()) == (n // 2) ** 3
    # HEALPix: the accumulator must be reserved for 12 nside^2 + 4 nside values
    nside = 8
    m = torch.empty(12 * nside * nside, dtype=torch.float32, device=cuda)
    args = (plan.handle, x.data_ptr(), n, n, n, 0, None, 1.0, pkg._lib.dbl3((100.0,) * 3),
            pkg._lib.dbl3((50.0,) * 3), 10.0, 40.0, nside, 1, 1, None, None, m.data_ptr(), 0, pkg._lib.stream_ptr())
    assert lib.pmx_paint_healpix_tiled(*args) == 4
    pkg._lib.check(lib.pmx_plan_reserve(plan.handle, 0, 0, 0, 12 * nside * nside + 4 * nside), "reserve acc")
    assert plan.workspace_bytes - base == grown + 8 * (12 * nside * nside + 4 * nside)
    pkg._lib.check(lib.pmx_paint_healpix_tiled(*args), "healpix")
    torch.cuda.synchronize()
    assert float(m.sum()) > 0.0


def _run_capped_kernels(pkg, cuda):
    eng = pkg.engine

    n = 32
    rng = np.random.default_rng(7)
    pos = torch.from_numpy((rng.standard_normal((n, n, n, 3)) * 0.8).astype(np.float32)).to(cuda)
    gbar = torch.from_numpy(rng.standard_normal((n, n, n, 3)).astype(np.float32)).to(cuda)
    out = {}
    box, obs = (200.0,) * 3, (100.0, 90.0, 110.0)
    out["img"] = eng.image_positions(pos, (n,) * 3, box, obs, rot=13, shift=(200.0, 0.0, -200.0)).clone()
    out["img_bar"] = eng.image_positions_vjp(pos, gbar, (n,) * 3, box, obs, rot=13,
                                             shift=(200.0, 0.0, -200.0))[0].clone()
    out["noise"] = eng.white_noise((n,) * 3, 1234, device=cuda).clone()
    dx, p = eng.lpt1(out["noise"], 0.3, 0.7)
    out["dx"], out["p"] = dx.clone(), p.clone()
    maps = torch.from_numpy(rng.random((3, 12 * 16 * 16)).astype(np.float32) + 0.5).to(cuda)
    coef = torch.from_numpy(rng.random((2, 3))).to(cuda)
    out["born"] = eng.born(maps, coef).clone()
    torch.cuda.synchronize()
    return out


def test_grid_stride_kernels_cover_capped_grids(pkg, cuda):
    lib = pkg._lib.load()
    ref = _run_capped_kernels(pkg, cuda)
    try:
        pkg._lib.check(lib.pmx_set_grid_cap(3), "grid cap")
        got = _run_capped_kernels(pkg, cuda)
    finally:
        pkg._lib.check(lib.pmx_set_grid_cap(0), "grid cap reset")
    for k in ref:
        if k == "born":  # per-workgroup float64 partial means: the grid changes their summation order
            torch.testing.assert_close(got[k], ref[k], rtol=1e-6, atol=1e-6)
        else:
            assert torch.equal(ref[k], got[k]), k
    assert lib.pmx_set_grid_cap(1 << 21) == 1
