"""pmx_paint_healpix_noalloc (the HEALPix shell paint without a plan, on a static device scratch
array) against pmx_paint_healpix_tiled on a plan, through the C ABI: two calls in a row, the second
with fewer particles and a smaller nside, plain and over rotated box images; and its refusal of an
nside beyond the scratch array (a host check, nothing launched)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _paint_both(pkg, cuda, n, nside, tiles, seed):
    lib = pkg._lib.load()
    mesh = (64, 64, 64)
    box, obs = (400.0,) * 3, (200.0, 190.0, 210.0)
    rng = np.random.default_rng(seed)
    pos = torch.from_numpy((rng.standard_normal((n, n, n, 3)) * 0.6).astype(np.float32)).to(cuda)
    if tiles:
        rot = torch.tensor([0, 5, 17, 33], dtype=torch.int32, device=cuda)
        shift = torch.tensor([[0.0, 0.0, 0.0], [400.0, 0.0, 0.0], [0.0, -400.0, 0.0], [0.0, 0.0, 400.0]],
                             dtype=torch.float64, device=cuda)
        nt, rp, sp = 4, rot.data_ptr(), shift.data_ptr()
        rmin, rmax = 150.0, 420.0
    else:
        nt, rp, sp = 1, None, None
        rmin, rmax = 60.0, 180.0
    ref = pkg.engine.paint_healpix(pos, mesh, box, obs, rmin, rmax, nside, scheme="bilinear",
                                   tiles=None if not tiles else (rot, shift))
    got = torch.full((12 * nside * nside,), -7.0, dtype=torch.float32, device=cuda)
    pkg._lib.check(lib.pmx_paint_healpix_noalloc(*mesh, pos.data_ptr(), n, n, n, 0, None, 1.0, pkg._lib.dbl3(box),
                                                 pkg._lib.dbl3(obs), rmin, rmax, nside, 1, nt, rp, sp, got.data_ptr(),
                                                 0, pkg._lib.stream_ptr()), "pmx_paint_healpix_noalloc")
    torch.cuda.synchronize()
    return ref, got


@pytest.mark.parametrize("tiles", [False, True])
def test_noalloc_shell_paint_matches_plan_paint(pkg, cuda, tiles):
    for n, nside, seed in ((64, 128, 1), (32, 32, 2)):
        ref, got = _paint_both(pkg, cuda, n, nside, tiles, seed)
        assert float(ref.abs().sum()) > 0.0
        # float64 global atomics: the summation order of a pixel's deposits may differ between calls
        torch.testing.assert_close(got, ref, rtol=1e-6, atol=1e-6 * float(ref.abs().max()))


def test_noalloc_rejects_large_nside(pkg, cuda):
    lib = pkg._lib.load()
    pos = torch.zeros((8, 8, 8, 3), dtype=torch.float32, device=cuda)
    out = torch.empty(12, dtype=torch.float32, device=cuda)
    rc = lib.pmx_paint_healpix_noalloc(64, 64, 64, pos.data_ptr(), 8, 8, 8, 0, None, 1.0, pkg._lib.dbl3((1.0,) * 3),
                                       pkg._lib.dbl3((0.5,) * 3), 0.1, 0.4, 1024, 1, 1, None, None, out.data_ptr(), 0,
                                       pkg._lib.stream_ptr())
    assert rc == 2 and b"nside" in lib.pmx_last_error()
