"""Flat-sky Born convergence on the GPU against the CPU oracle (_born_core_impl's flat branch,
src/jax_fli/_src/lensing/_born.py:53-98, and _born_flat, :131-190): each plane resampled by periodic
bilinear interpolation at its chi-scaled angular grid, then weighted by the lensing efficiency and
the n(z) quadrature. Tolerances: kappa and the map adjoint 1e-5 relative L2 (float32 maps, float64
geometry and sums); the Omega_c gradient (through the coefficients and the sampling positions)
1e-4 against float64 central differences of the oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import nbody_oracle as nbo
from oracle import sky_oracle as so

pytestmark = pytest.mark.gpu

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _lightcone(pkg, maps, a, field_size, box, cuda):
    S, ny, nx = maps.shape
    return pkg.FlatDensity(array=torch.from_numpy(maps).to(cuda), mesh_size=(16, 16, 16), box_size=box,
                           flatsky_npix=(ny, nx), field_size=field_size, status=pkg.FieldStatus.LIGHTCONE,
                           scale_factors=a, comoving_centers=np.zeros(S), density_width=np.zeros(S))


@pytest.mark.parametrize("sources", ["distribution", "redshift"])
@pytest.mark.parametrize("shape,fsz", [((64, 64), 10.0), ((48, 80), (6.0, 9.0))])
def test_flat_born_matches_oracle(pkg, cuda, sources, shape, fsz):
    c = pkg.Cosmology(**P)
    o = nbo.Cosmo(**P)
    box = (500.0, 400.0, 600.0)
    rng = np.random.default_rng(50)
    a = np.array([0.6, 0.7, 0.8, 0.9])
    maps = (rng.random((4,) + shape) * 2.0).astype(np.float32)
    maps[2] = 0.0  # an empty plane: its mean is replaced by the float32 eps
    lc = _lightcone(pkg, maps, a, fsz, box, cuda)
    if sources == "distribution":
        nz, onz = [pkg.tophat_z(0.0, 1.0), pkg.tophat_z(0.3, 0.8)], [so.tophat_pdf(0.0, 1.0), so.tophat_pdf(0.3, 0.8)]
    else:
        nz = onz = [0.5, 1.2]
    kap = pkg.born(c, lc, nz)
    geom = (fsz if isinstance(fsz, tuple) else (fsz, fsz), shape, box)
    ref = so.born_flat(maps.astype(np.float64), o.chi(a), a, onz, lc.max_comoving_radius / 4, o.Omega_m, o.chi, geom)
    assert kap.array.shape == (2,) + shape
    assert isinstance(kap, pkg.FlatKappaField) and kap.status == pkg.FieldStatus.KAPPA
    assert rel(kap.array.cpu().numpy(), ref) < 1e-5


def test_flat_born_vjp_and_parameter_gradient(pkg, cuda):
    o = nbo.Cosmo(**P)
    box, fsz, shape = (500.0, 500.0, 500.0), (8.0, 8.0), (40, 48)
    rng = np.random.default_rng(51)
    a = np.array([0.6, 0.7, 0.8, 0.9])
    maps = (rng.random((4,) + shape) + 0.5).astype(np.float32)
    kb = rng.standard_normal((1,) + shape)
    Oc = torch.tensor(P["Omega_c"], dtype=torch.float64, requires_grad=True)
    c = pkg.Cosmology(**{**P, "Omega_c": Oc})
    mt = torch.from_numpy(maps).to(cuda).requires_grad_(True)
    lc = _lightcone(pkg, maps, a, fsz, box, cuda).replace(array=mt)
    kap = pkg.born(c, lc, [pkg.tophat_z(0.0, 1.0)])
    (kap.array.double() * torch.from_numpy(kb).to(cuda)).sum().backward()
    d_r = lc.max_comoving_radius / 4
    onz = [so.tophat_pdf(0.0, 1.0)]
    geom = (fsz, shape, box)
    ref = so.born_flat_vjp(maps.astype(np.float64), kb, o.chi(a), a, onz, d_r, o.Omega_m, o.chi, geom)
    assert rel(mt.grad.cpu().numpy(), ref) < 1e-5
    h = 1e-5

    def L(oc):
        oo = nbo.Cosmo(**{**P, "Omega_c": oc})
        return float((so.born_flat(maps.astype(np.float64), oo.chi(a), a, onz, d_r, oo.Omega_m, oo.chi, geom)
                      * kb).sum())

    fd = (L(P["Omega_c"] + h) - L(P["Omega_c"] - h)) / (2 * h)
    assert float(Oc.grad) == pytest.approx(fd, rel=1e-4)


def test_flat_born_geometry_errors(pkg, cuda):
    c = pkg.Cosmology(**P)
    maps = np.ones((2, 8, 8), np.float32)
    lc = _lightcone(pkg, maps, np.array([0.8, 0.9]), 5.0, (100.0,) * 3, cuda)
    with pytest.raises(ValueError):
        pkg.born(c, lc.replace(field_size=None), [0.5])
