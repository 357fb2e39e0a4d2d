"""Source redshift distributions for lensing: ``redshift_distribution`` (the jax_cosmo base class the
reference's ``nz_shear`` entries are, src/jax_fli/_src/lensing/_normalize_nz.py:22-30) and
``tophat_z`` (src/jax_fli/power/theory.py:28-47).

A distribution is called on redshifts and returns pz(z) / norm, norm = Simpson's rule of pz over
[0, zmax] with 256 intervals (jax_cosmo's normalisation, zmax = 10 by default). Host float64.
"""

from __future__ import annotations

import numpy as np


def simps_weights(a: float, b: float, n: int) -> np.ndarray:
    """Composite Simpson weights dx/3 [1, 4, 2, ..., 4, 1] for n (even) intervals
    (_simps_weights, src/jax_fli/_src/lensing/_born.py:16-26)."""
    if n % 2:
        raise ValueError("Simpson's rule needs an even number of intervals")
    dx = (b - a) / n
    w = np.ones(n + 1)
    w[1:-1:2] = 4.0
    w[2:-2:2] = 2.0
    return w * (dx / 3.0)


class redshift_distribution:
    """Base class: subclasses define pz_fn(z) with their parameters in self.params."""

    def __init__(self, *params, gals_per_arcmin2: float = 1.0, zmax: float = 10.0):
        self.params = params
        self.gals_per_arcmin2 = gals_per_arcmin2
        self.zmax = zmax
        z = np.linspace(0.0, zmax, 257)
        self._norm = float((simps_weights(0.0, zmax, 256) * np.asarray(self.pz_fn(z), dtype=np.float64)).sum())

    def pz_fn(self, z):
        raise NotImplementedError

    def __call__(self, z):
        return np.asarray(self.pz_fn(np.asarray(z, dtype=np.float64)), dtype=np.float64) / self._norm

    @property
    def gals_per_steradian(self) -> float:
        return self.gals_per_arcmin2 * (180.0 * 60.0 / np.pi) ** 2


class tophat_z(redshift_distribution):
    """Top-hat n(z) on [zmin, zmax] (both edges included)."""

    def pz_fn(self, z):
        zmin, zmax = self.params
        return np.where((z >= zmin) & (z <= zmax), 1.0, 0.0)

