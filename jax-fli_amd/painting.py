"""PaintingOptions: what nbody/lpt produce at each snapshot (src/jax_fli/fields/painting.py:15-96)."""

from __future__ import annotations

import dataclasses
from typing import Any, Literal

PaintingTarget = Literal["spherical", "flat", "density", "particles"]
DEFAULT_CHUNK_SIZE = 2**24


@dataclasses.dataclass(frozen=True)
class PaintingOptions:
    target: PaintingTarget
    scheme: Literal["ngp", "bilinear", "rbf_neighbor"] = "bilinear"
    kernel_width_arcmin: float | None = None
    smoothing_interpretation: str = "fwhm"
    paint_nside: int | None = None
    ud_grade_power: float = 0.0
    ud_grade_order_in: str = "RING"
    ud_grade_order_out: str = "RING"
    ud_grade_pess: bool = False
    chunk_size: int = DEFAULT_CHUNK_SIZE
    weights: Any = 1.0
    batch_size: int | None = None

    def __post_init__(self):
        if self.target not in ("spherical", "flat", "density", "particles"):
            raise ValueError(f"Unknown painting target {self.target}.")
