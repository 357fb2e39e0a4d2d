"""ts / nb_shells / density_widths -> (ts, r_centers, widths, is_lightcone)
(src/jax_fli/pm/_resolve_geometry.py:12-114). Host float64, ts returned as float32."""

from __future__ import annotations

import numpy as np

from .. import cosmology as _c
from ..utils import compute_lightcone_shells, distances


def resolve_ts_geometry(cosmo, field, painting, ts=None, nb_shells=None, density_widths=None, equal_vol=False,
                        min_width=50.0):
    if ts is None and nb_shells is None:
        raise ValueError("Cannot specify both `ts` and `nb_shells`.")
    is_projecting = painting.target in ("flat", "spherical")
    if ts is not None and nb_shells is not None:
        raise ValueError("Cannot specify both `ts` and `nb_shells`.")
    f32 = lambda a: np.asarray(a, dtype=np.float32)
    if nb_shells is not None:
        r_centers, ts_resolved = compute_lightcone_shells(cosmo, field, nb_shells=nb_shells, equal_vol=equal_vol,
                                                          min_width=min_width)
        if density_widths is not None:
            try:
                density_widths = np.broadcast_to(np.asarray(density_widths, dtype=np.float64), r_centers.shape)
            except ValueError:
                raise ValueError("density_widths must be broadcastable to the shape of r_centers.")
        else:
            density_widths = distances(r_centers, 0.0)
        return f32(ts_resolved), r_centers, np.asarray(density_widths, dtype=np.float64), True
    ts = np.asarray(ts, dtype=np.float64)
    if ts.ndim == 0 or (ts.ndim == 1 and ts.size == 1):
        ts = np.atleast_1d(ts).reshape(())
        r_center = _c.radial_comoving_distance(cosmo, ts)
        if density_widths is not None:
            dw = np.asarray(density_widths).squeeze()
            if dw.ndim != 0:
                raise ValueError("For single-shell mode, density_widths must be a scalar.")
            density_widths = np.atleast_1d(dw).astype(np.float64)
        elif is_projecting:
            raise ValueError("Projection painting requires explicit density_widths for single-shell mode.")
        else:
            density_widths = np.array([0.0])
        return f32(np.atleast_1d(ts)), np.atleast_1d(r_center), density_widths, False
    if ts.ndim == 1:
        r_centers = _c.radial_comoving_distance(cosmo, ts)
        if density_widths is not None:
            try:
                density_widths = np.broadcast_to(np.asarray(density_widths, dtype=np.float64), r_centers.shape)
            except ValueError:
                raise ValueError("density_widths must be broadcastable to the shape of ts.")
        else:
            density_widths = distances(r_centers, 0.0)
        return f32(ts), r_centers, np.asarray(density_widths, dtype=np.float64), True
    if ts.ndim == 2 and ts.shape[0] == 2:
        r_near = _c.radial_comoving_distance(cosmo, ts[0])
        r_far = _c.radial_comoving_distance(cosmo, ts[1])
        r_centers = 0.5 * (r_near + r_far)
        return f32(_c.a_of_chi(cosmo, r_centers)), r_centers, np.abs(r_far - r_near), True
    raise ValueError(f"ts has unsupported shape {ts.shape}. Expected scalar, 1-D, or (2, N).")
