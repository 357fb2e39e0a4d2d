"""Power spectra through pmx_power_multipoles against the oracle's estimator (_power,
src/jax_fli/_src/power/_compute.py:5-119): auto and cross spectra, Legendre multipoles about an oblique
line of sight on cubic and non-cubic meshes (the x/y Nyquist lines, whose mirror mode is not -k), the
pole groups of many-bin requests, transfer and coherence, and the analytic plane-wave ratios.

Tolerance: per pole, max |P_gpu - P_oracle| <= 2e-6 max |P_0| (float32 FFT of the same float32 mesh
against a float64 FFT; float64 binning on both sides); k averages to 1e-12 relative.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import pm_oracle as orc

pytestmark = pytest.mark.gpu


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape).astype(np.float32)
    b = (0.6 * a + 0.8 * rng.standard_normal(shape)).astype(np.float32)
    return a, b


def _close(got, ref, scale, tol=2e-6):
    got, ref = np.atleast_2d(got), np.atleast_2d(ref)
    assert got.shape == ref.shape
    err = np.abs(got - ref).max(axis=1)
    assert np.all(err <= tol * scale), (err / scale)


@pytest.mark.parametrize("shape,box", [((32, 32, 32), (100.0,) * 3), ((64, 32, 16), (200.0, 100.0, 50.0))])
@pytest.mark.parametrize("cross", [False, True])
def test_multipoles_match_oracle(pkg, cuda, shape, box, cross):
    a, b = _fields(shape, 5)
    poles, los = (0, 1, 2, 3, 4), (0.3, -0.5, 0.8)
    kedges = np.linspace(0.02, np.pi * min(np.asarray(shape) / np.asarray(box)), 21)
    m2 = torch.from_numpy(b).to(cuda) if cross else None
    k, pk = pkg.power(torch.from_numpy(a).to(cuda), m2, box_shape=box, kedges=kedges, multipoles=poles, los=los)
    k_ref, pk_ref = orc.power(a, box, kedges, mesh2=b if cross else None, multipoles=poles, los=los)
    assert np.allclose(k, k_ref, rtol=1e-12, atol=0)
    _close(pk, pk_ref, np.abs(pk_ref[0]).max())
    if not cross:  # odd poles of an auto spectrum vanish
        assert np.abs(pk[[1, 3]]).max() < 1e-6 * np.abs(pk[0]).max()


def test_monopole_paths_agree(pkg, cuda):
    """multipoles=0 (no line of sight), [0] with one, and the plain pmx_power sums give one monopole."""
    shape, box = (32, 16, 64), (90.0, 40.0, 160.0)
    a, _ = _fields(shape, 8)
    t = torch.from_numpy(a).to(cuda)
    k0, p0 = pkg.power(t, box_shape=box)
    _, p1 = pkg.power(t, box_shape=box, multipoles=[0], los=(1.0, 2.0, 3.0))
    k_ref, p_ref = orc.power(a, box)
    assert np.allclose(k0, k_ref, rtol=1e-12, atol=0)
    _close(p0, p_ref, np.abs(p_ref).max())
    assert np.allclose(p0, p1, rtol=1e-12, atol=0)  # float64 atomics: summation order varies


def test_many_bins_and_poles_split_into_launches(pkg, cuda):
    """512 bins and 16 poles exceed one launch's LDS histograms: the poles go in several launches."""
    shape, box = (32, 32, 32), (64.0, 64.0, 64.0)
    a, b = _fields(shape, 9)
    poles = tuple(range(16))
    kedges = np.linspace(0.05, 3.0, 512)
    k, pk = pkg.power(torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda), box_shape=box, kedges=kedges,
                      multipoles=poles, los=(0.0, 1.0, 1.0))
    k_ref, pk_ref = orc.power(a, box, kedges, mesh2=b, multipoles=poles, los=(0.0, 1.0, 1.0))
    assert pk.shape == (16, 511)
    sel = ~np.isnan(k_ref)
    assert np.allclose(k[sel], k_ref[sel], rtol=1e-12, atol=0)
    _close(pk, pk_ref, np.abs(pk_ref[0]).max(), tol=5e-6)


def test_plane_wave_ratios(pkg, cuda):
    """cos(2 pi 4 x / n): P_l / P_0 = 2l+1 (even l) along x, (2l+1) P_l(0) across; against cos(. + phi)
    the cross monopole is cos(phi) P_0 and the x dipole 3 sin(phi) P_0."""
    n, L, phi = 32, 100.0, 0.7
    x = 2 * np.pi * 4 * np.arange(n) / n
    w = np.broadcast_to(np.cos(x)[:, None, None], (n, 64, 16)).astype(np.float32).copy()
    w2 = np.broadcast_to(np.cos(x + phi)[:, None, None], (n, 64, 16)).astype(np.float32).copy()
    box = (L, 1.5 * L, 0.5 * L)
    kedges = np.linspace(0.01, 1.5, 40)
    t, t2 = torch.from_numpy(w).to(cuda), torch.from_numpy(w2).to(cuda)
    _, p0 = pkg.power(t, box_shape=box, kedges=kedges)
    bi = int(np.argmax(p0))
    _, along = pkg.power(t, box_shape=box, kedges=kedges, multipoles=(0, 1, 2, 3, 4), los=(1.0, 0.0, 0.0))
    assert np.allclose(along[:, bi] / p0[bi], [1, 0, 5, 0, 9], atol=1e-6)
    _, across = pkg.power(t, box_shape=box, kedges=kedges, multipoles=(0, 2, 4), los=(0.0, 0.0, 3.0))
    assert np.allclose(across[:, bi] / p0[bi], [1, -2.5, 27 / 8], atol=1e-6)
    _, c = pkg.power(t, t2, box_shape=box, kedges=kedges, multipoles=(0, 1), los=(1.0, 0.0, 0.0))
    assert c[0][bi] / p0[bi] == pytest.approx(np.cos(phi), abs=1e-6)
    assert c[1][bi] / p0[bi] == pytest.approx(3 * np.sin(phi), abs=1e-6)


def test_transfer_coherence_and_cross_power(pkg, cuda):
    shape, box = (32, 32, 32), (80.0,) * 3
    a, b = _fields(shape, 12)
    ta, tb = torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda)
    k, tr = pkg.transfer(ta, tb, box_shape=box)
    k_ref, tr_ref = orc.transfer(a, b, box)
    assert np.allclose(tr, tr_ref, rtol=2e-6)
    _, coh = pkg.coherence(ta, tb, box_shape=box)
    _, coh_ref = orc.coherence(a, b, box)
    assert np.allclose(coh, coh_ref, rtol=2e-6)
    # batched field: B(B+1)/2 spectra in upper-triangular order
    c = (0.3 * a - b).astype(np.float32)
    stack = torch.from_numpy(np.stack([a, b, c])).to(cuda)
    fld = pkg.DensityField(array=stack, mesh_size=shape, box_size=box)
    cp = fld.cross_power()
    assert cp.spectra.shape[0] == 6
    mats = [a, b, c]
    pairs = [(i, j) for i in range(3) for j in range(i, 3)]
    for s, (i, j) in zip(cp.spectra, pairs):
        _, ref = orc.power(mats[i], box, mesh2=mats[j])
        _close(s, ref, np.abs(ref).max())
    tf = fld.transfer(pkg.DensityField(array=2 * stack, mesh_size=shape, box_size=box))
    assert np.allclose(tf.spectra, 2.0, rtol=1e-6)


def test_bad_arguments_are_refused(pkg, cuda):
    t = torch.zeros((16, 16, 16), device=cuda)
    with pytest.raises(pkg._lib.PmxError):
        pkg.power(t, box_shape=(10.0,) * 3, multipoles=(0, 17))
    with pytest.raises(pkg._lib.PmxError):
        pkg.power(t, box_shape=(10.0,) * 3, multipoles=(2,), los=(0.0, 0.0, 0.0))
    with pytest.raises(ValueError):
        pkg.power(t, torch.zeros((16, 16, 8), device=cuda), box_shape=(10.0,) * 3)
