"""Lightcone shell geometry on the host (src/jax_fli/utils.py:22-312): equal-width or equal-volume
shells, shell edges/widths/centres. Pure float64 NumPy; no device work."""

from __future__ import annotations

import numpy as np

from . import cosmology as _c


def _compute_equal_vol_shells(R_last: float, N_total: int, min_width: float) -> np.ndarray:
    """Equal-volume inner shells with at most M outer shells of exactly min_width (utils.py:22-71)."""
    R_start = 0.0
    if N_total * min_width > R_last - R_start:
        raise ValueError(f"Cannot fit {N_total} shells of min_width={min_width} Mpc/h in max_radius={R_last:.2f} Mpc/h")
    for M in range(N_total):
        R_split = R_last - M * min_width
        N_inner = N_total - M
        k_inner = np.arange(N_inner + 1)
        R_inner = (R_start**3 + (k_inner / N_inner) * (R_split**3 - R_start**3)) ** (1 / 3)
        if R_inner[-1] - R_inner[-2] >= min_width - 1e-10:
            if M > 0:
                return np.concatenate((R_inner, R_split + np.arange(1, M + 1) * min_width))
            return R_inner
    raise ValueError(f"No valid equal-volume + min_width partition found for nb_shells={N_total}, "
                     f"min_width={min_width}, max_radius={R_last:.2f}")


def compute_lightcone_shells(cosmo, field, nb_shells: int, equal_vol: bool = False, min_width: float = 50.0):
    """(r_center, a_center), far to near (utils.py:176-230)."""
    max_radius = field.max_comoving_radius
    if equal_vol:
        r_edges = _compute_equal_vol_shells(float(max_radius), nb_shells, float(min_width))[::-1]
    else:
        w = max_radius / nb_shells
        n_lens = int(max_radius // float(w))
        r_edges = np.linspace(0.0, float(n_lens) * float(w), n_lens + 1)[::-1]
    r_center = 0.5 * (r_edges[1:] + r_edges[:-1])
    return r_center, _c.a_of_chi(cosmo, r_center)


def edges(centers, r_left=None) -> np.ndarray:
    """(2, N) left/right edges from centres, identical for ascending/descending input (utils.py:247-281)."""
    centers = np.asarray(centers, dtype=np.float64)
    desc = centers[0] > centers[-1]
    c = centers[::-1] if desc else centers
    if r_left is None:
        r_left = 1.5 * c[0] - 0.5 * c[1]
    e = [float(r_left)]
    for ci in c:
        e.append(2 * ci - e[-1])
    e = np.asarray(e)
    left, right = e[:-1], e[1:]
    if desc:
        left, right = left[::-1], right[::-1]
    return np.stack([left, right])


def distances(centers, r_left=None) -> np.ndarray:
    e = edges(centers, r_left)
    return e[1] - e[0]


def centers(distances, r_left, is_descending=False) -> np.ndarray:
    d = np.asarray(distances, dtype=np.float64)
    d = d[::-1] if is_descending else d
    r = np.concatenate([[r_left], r_left + np.cumsum(d)])
    c = 0.5 * (r[1:] + r[:-1])
    return c[::-1] if is_descending else c
