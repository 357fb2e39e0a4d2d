"""Second-order and lightcone LPT (pmx_lpt) against the oracle's restatement of jaxpm.pm.lpt
(src/jax_fli/pm/lpt.py:19-214): 2LPT displacements and momenta at one scale factor on cubic and
non-cubic meshes and on a particle sub-lattice, the second-order part on its own, per-particle
scale factors of the lightcone mode, and its flat and HEALPix shell maps.

Tolerances (relative L2): total displacements and momenta 1e-5; the second-order part alone 2e-4
(its source is a product of two float32 Hessians, against float64); shell maps 1e-5 against the
oracle's paint of the device's own displacements; per-particle scale factors 1e-6."""

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


def _ic(pkg, mesh, box, seed, **kw):
    return pkg.gaussian_initial_conditions(seed, mesh, box, cosmo=pkg.Cosmology(**P), **kw)


@pytest.mark.parametrize("mesh,box,a", [((32, 32, 32), (100.0,) * 3, 0.5), ((64, 32, 16), (256.0, 128.0, 64.0), 0.2)])
def test_lpt2_matches_oracle(pkg, cuda, mesh, box, a):
    c = pkg.Cosmology(**P)
    ic = _ic(pkg, mesh, box, 31)
    dx1, p1 = pkg.lpt(c, ic, ts=a, order=1)
    dx2, p2 = pkg.lpt(c, ic, ts=a, order=2)
    assert dx2.status == pkg.FieldStatus.LPT2
    o = nbo.Cosmo(**P)
    icn = ic.array.cpu().numpy()
    r1, q1 = nbo.lpt(o, icn, a, order=1)
    r2, q2 = nbo.lpt(o, icn, a, order=2)
    g1, g2 = dx1.array.cpu().numpy(), dx2.array.cpu().numpy()
    print(f"{mesh}: |dx2 - dx1| / |dx1| = {np.linalg.norm(r2 - r1) / np.linalg.norm(r1):.3e}")
    assert rel(g2, r2) < 1e-5 and rel(p2.array.cpu().numpy(), q2) < 1e-5
    assert rel(g2 - g1, r2 - r1) < 2e-4
    assert rel(p2.array.cpu().numpy() - p1.array.cpu().numpy(), q2 - q1) < 2e-4


def test_lpt2_on_a_sub_lattice(pkg, cuda):
    c = pkg.Cosmology(**P)
    mesh, box, a = (32, 32, 32), (128.0,) * 3, 0.4
    ic = _ic(pkg, mesh, box, 32)
    lat = torch.stack(torch.meshgrid(*[torch.arange(16, dtype=torch.float32) * 2] * 3, indexing="ij"), -1)
    dx, p = pkg.lpt(c, ic, ts=a, order=2, initial_particles=lat)
    assert tuple(dx.array.shape) == (16, 16, 16, 3)
    full, _ = nbo.lpt(nbo.Cosmo(**P), ic.array.cpu().numpy(), a, order=2)
    assert rel(dx.array.cpu().numpy(), full[::2, ::2, ::2]) < 1e-5


@pytest.mark.parametrize("order", [1, 2])
def test_lightcone_lpt_per_particle_scale_factors(pkg, cuda, order):
    c = pkg.Cosmology(**P)
    mesh, box = (32, 32, 32), (400.0,) * 3
    ic = _ic(pkg, mesh, box, 33, observer_position=(0.5, 0.5, 0.0))
    dx, p = pkg.lpt(c, ic, nb_shells=4, order=order, painting=pkg.PaintingOptions(target="particles"))
    assert dx.status == pkg.FieldStatus.LIGHTCONE
    o = nbo.Cosmo(**P)
    a = nbo.lightcone_scale_factors(o, mesh, box, ic.observer_position_mpc)
    assert np.abs(np.asarray(dx.scale_factors, np.float64) - a).max() < 1e-6
    r, q = nbo.lpt(o, ic.array.cpu().numpy(), a, order=order)
    assert rel(dx.array.cpu().numpy(), r) < 1e-5 and rel(p.array.cpu().numpy(), q) < 1e-5
    # the nearest and farthest particles grow differently: a single-a LPT differs
    near = nbo.lpt(o, ic.array.cpu().numpy(), float(a.max()), order=order)[0]
    assert rel(dx.array.cpu().numpy(), near) > 1e-2


def test_lightcone_lpt_shell_maps(pkg, cuda):
    c = pkg.Cosmology(**P)
    mesh, box = (32, 32, 32), (400.0,) * 3
    ic = _ic(pkg, mesh, box, 34, observer_position=(0.5, 0.5, 0.0), nside=8, flatsky_npix=(16, 16))
    dxp, _ = pkg.lpt(c, ic, nb_shells=3, order=2, painting=pkg.PaintingOptions(target="particles"))
    disp = dxp.array.cpu().numpy()
    sph, _ = pkg.lpt(c, ic, nb_shells=3, order=2, painting=pkg.PaintingOptions(target="spherical", scheme="bilinear"))
    flat, _ = pkg.lpt(c, ic, nb_shells=3, order=2, painting=pkg.PaintingOptions(target="flat"))
    # centres come back far -> near reversed to near -> far, with scale factors a(centre)
    centers = np.asarray(sph.comoving_centers)
    widths = np.asarray(sph.density_width)
    ns = centers.size
    assert ns >= 2 and sph.array.shape == (ns, 12 * 8 * 8) and flat.array.shape == (ns, 16, 16)
    assert np.all(np.diff(centers) > 0)
    obs = ic.observer_position_mpc
    o = nbo.Cosmo(**P)
    assert np.allclose(np.asarray(sph.scale_factors), nbo.a_of_chi(o, centers), rtol=1e-7)
    for b in range(ns):
        ref = orc.paint_healpix(disp, mesh, box, obs, centers[b] - widths[b] / 2, centers[b] + widths[b] / 2, 8,
                                scheme="bilinear")
        assert rel(sph.array[b].cpu().numpy(), ref) < 1e-5
        fc, fw = np.asarray(flat.comoving_centers)[b], np.asarray(flat.density_width)[b]
        zlo, zhi, scale = pkg.fields.flat_slab(flat, fc, fw)
        ref = so.paint_flat(disp, mesh, (16, 16), zlo, zhi, scale)
        assert rel(flat.array[b].cpu().numpy(), ref) < 1e-5


def test_second_order_lpt_refuses_gradients(pkg, cuda):
    c = pkg.Cosmology(**P)
    ic = _ic(pkg, (16, 16, 16), (64.0,) * 3, 35)
    ic = ic.replace(array=ic.array.clone().requires_grad_(True))
    with pytest.raises(NotImplementedError):
        pkg.lpt(c, ic, ts=0.5, order=2)
