"""Lightcone geometry of an `nbody` / `lpt` call.

The caller names its output times in one of four ways, and each maps to shells of comoving distance
(the contract of resolve_ts_geometry, src/jax_fli/pm/_resolve_geometry.py:12-114):

==========================  =========================================  ============
input                        shells                                     lightcone?
==========================  =========================================  ============
``nb_shells=n``              ``compute_lightcone_shells`` of the box    yes
``ts`` one scale factor      one shell at chi(a); width 0 unless given  no
``ts`` 1-D, several values   shells centred at chi(a_i)                 yes
``ts`` (2, n)                near/far scale factors per shell           yes
==========================  =========================================  ============

Widths default to the spacing of the centres (`utils.distances`); explicit widths broadcast to the
shells. Everything is host float64; scale factors come back as float32, the time type of the loop.
Errors are ValueErrors: both or neither of ``ts`` / ``nb_shells``, widths that do not broadcast,
a non-scalar width for a single time, a projected target (flat, spherical) at a single time without
a width, and any other shape of ``ts``.
"""

from __future__ import annotations

import numpy as np

from .. import cosmology as _c
from ..utils import compute_lightcone_shells, distances


def _shell_widths(centres: np.ndarray, widths, what: str) -> np.ndarray:
    if widths is None:
        return np.asarray(distances(centres, 0.0), dtype=np.float64)
    try:
        return np.array(np.broadcast_to(np.asarray(widths, dtype=np.float64), centres.shape))
    except ValueError:
        raise ValueError(f"density_widths must be broadcastable to the shape of {what}.") from None


def _as_times(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float32)


def _by_count(cosmo, field, count, widths, equal_vol, min_width):
    centres, times = compute_lightcone_shells(cosmo, field, nb_shells=count, equal_vol=equal_vol,
                                              min_width=min_width)
    return _as_times(times), centres, _shell_widths(centres, widths, "r_centers"), True


def _single_time(cosmo, a: np.ndarray, widths, projected: bool):
    a = a.reshape(1)
    centre = np.atleast_1d(_c.radial_comoving_distance(cosmo, a[0]))
    if widths is None:
        if projected:
            raise ValueError("Projection painting requires explicit density_widths for single-shell mode.")
        return _as_times(a), centre, np.zeros(1), False
    w = np.asarray(widths, dtype=np.float64).squeeze()
    if w.ndim != 0:
        raise ValueError("For single-shell mode, density_widths must be a scalar.")
    return _as_times(a), centre, w.reshape(1), False


def _by_centres(cosmo, a: np.ndarray, widths):
    centres = _c.radial_comoving_distance(cosmo, a)
    return _as_times(a), centres, _shell_widths(centres, widths, "ts"), True


def _by_edges(cosmo, a: np.ndarray):
    near = _c.radial_comoving_distance(cosmo, a[0])
    far = _c.radial_comoving_distance(cosmo, a[1])
    centres = 0.5 * (near + far)
    return _as_times(_c.a_of_chi(cosmo, centres)), centres, np.abs(far - near), True


def resolve_ts_geometry(cosmo, field, painting, ts=None, nb_shells=None, density_widths=None, equal_vol=False,
                        min_width=50.0):
    """(ts float32, r_centers, density_widths, is_lightcone) for one of the inputs in the table above."""
    if (ts is None) == (nb_shells is None):
        raise ValueError("Give exactly one of `ts` and `nb_shells`.")
    if nb_shells is not None:
        return _by_count(cosmo, field, nb_shells, density_widths, equal_vol, min_width)
    a = np.asarray(ts, dtype=np.float64)
    if a.size == 1 and a.ndim <= 1:
        return _single_time(cosmo, a, density_widths, painting.target in ("flat", "spherical"))
    if a.ndim == 1:
        return _by_centres(cosmo, a, density_widths)
    if a.ndim == 2 and a.shape[0] == 2:
        return _by_edges(cosmo, a)
    raise ValueError(f"ts has unsupported shape {a.shape}. Expected scalar, 1-D, or (2, N).")
