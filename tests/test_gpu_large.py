"""Full-size checks of the large single-GPU configuration (512^3 particles on the regular sub-lattice
of a 1024^3 mesh, box 1000 Mpc/h, HEALPix shells at nside 512 painted from periodic images of the
box) through size-independent properties, since the oracle cannot run at this size: paint mass and
determinism, momentum conservation of the forces, the reversible step's round trip, and the particle
count of a far shell. The same code paths are checked against the oracle at small sizes in
test_gpu_kernels.py, test_gpu_nbody.py and test_gpu_sky.py."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
N, NP, BOX = 1024, 512, 1000.0


@pytest.fixture(scope="module")
def large(pkg, cuda):
    c = pkg.Cosmology(**P)
    ic = pkg.gaussian_initial_conditions(3, (N,) * 3, (BOX,) * 3, cosmo=c, nside=512)
    lat = torch.stack(torch.meshgrid(*[torch.arange(NP, dtype=torch.float32, device=cuda) * (N // NP)] * 3,
                                     indexing="ij"), dim=-1)
    dx, p = pkg.lpt(c, ic, ts=0.1, initial_particles=lat)
    del ic, lat
    yield pkg, c, dx, p
    pkg._lib.release_plan((N,) * 3, cuda)
    torch.cuda.empty_cache()


def test_paint_mass_and_determinism(large):
    pkg, c, dx, p = large
    rho = pkg.engine.paint(dx.array, (N,) * 3)
    total = float(rho.sum(dtype=torch.float64))
    assert abs(total - NP**3) / NP**3 < 1e-6
    assert torch.equal(rho, pkg.engine.paint(dx.array, (N,) * 3))


def test_momentum_conservation(large):
    pkg, c, dx, p = large
    f = pkg.engine.forces(dx.array, (N,) * 3).double()
    net = f.sum(dim=(0, 1, 2)).cpu().numpy()
    assert np.abs(net).max() < 1e-3 * f.abs().sum().item() / f.numel() * np.sqrt(f.numel())


def test_reversible_step_roundtrip(large):
    pkg, c, dx, p = large
    interp = pkg.NoInterp(painting=pkg.PaintingOptions(target="particles")).update_geometry(
        ts=np.array([1.0], np.float32), r_centers=np.zeros(1), density_widths=np.ones(1), max_comoving_distance=1.0)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=interp)
    d0, p0, st = solver.init(dx, p, 0.1, 0.13, 0.03, c)
    d1, p1, st1 = solver.step(d0, p0, 0.1, 0.13, 0.03, st, c)
    d2, p2, _ = solver.reverse(d1, p1, 0.1, 0.13, 0.03, st1, c)
    rel = lambda a, b: float((a.double() - b.double()).norm() / b.double().norm())
    assert rel(d2.array, d0.array) < 1e-6
    assert rel(p2.array, p0.array) < 1e-5


def test_far_shell_counts_particle_images(large):
    """A shell at 1.5-1.8 box lengths painted from the 125 images of the box holds N_p V_shell / L^3
    particle-images (structure moves it by far less than 1%), and ngp and bilinear deposit the same
    total weight."""
    pkg, c, dx, p = large
    order = [0, 1, -1, 2, -2]
    shifts = np.array([(a, b, d) for a in order for b in order for d in order], dtype=np.float64) * BOX
    tiles = (np.zeros(len(shifts), np.int32), shifts)
    rmin, rmax, nside = 1500.0, 1800.0, 512
    vol = 4 * np.pi / (12 * nside**2) * (rmax**3 - rmin**3) / 3
    tot = {}
    for scheme in ("ngp", "bilinear"):
        m = pkg.engine.paint_healpix(dx.array, (N,) * 3, (BOX,) * 3, (BOX / 2,) * 3, rmin, rmax, nside, scheme=scheme,
                                     tiles=tiles)
        tot[scheme] = float(m.sum(dtype=torch.float64)) * vol
    expect = NP**3 * (4 * np.pi / 3) * (rmax**3 - rmin**3) / BOX**3
    assert tot["bilinear"] == pytest.approx(expect, rel=1e-2)
    assert tot["ngp"] == pytest.approx(tot["bilinear"], rel=1e-6)
