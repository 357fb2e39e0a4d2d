"""Field containers mirroring the reference's PyTrees (``AbstractField``, ``ParticleField``,
``DensityField``, ``FlatDensity``, ``SphericalDensity``; src/jax_fli/_src/base/_core.py:175-246,
src/jax_fli/fields/particles.py:33-442, fields/density.py:24-47, fields/lightcone.py:446-471).

Same field names, enums, static metadata and methods; the ``array`` is a torch tensor on the GPU.
Fields are immutable: every method returns a new field (``replace``), as in the reference.
Painting and readout dispatch to the pmx kernels.
"""

from __future__ import annotations

import dataclasses
from enum import Enum, auto
from typing import Any

import numpy as np
import torch

from . import engine


class FieldStatus(str, Enum):
    UNKNOWN = auto()
    INITIAL_FIELD = auto()
    LPT1 = auto()
    LPT2 = auto()
    DENSITY_FIELD = auto()
    PARTICLES = auto()
    PROJECTED_DENSITY = auto()
    LIGHTCONE = auto()
    KAPPA = auto()
    GAMMA = auto()


class PhysicalUnit(str, Enum):
    INVALID_UNIT = auto()


class PositionUnit(str, Enum):
    GRID_RELATIVE = auto()
    GRID_ABSOLUTE = auto()
    MPC_H = auto()


class DensityUnit(str, Enum):
    OVERDENSITY = auto()
    DENSITY = auto()
    COUNTS = auto()
    MSUN_H_PER_MPC3 = auto()


class ConvergenceUnit(str, Enum):
    DIMENSIONLESS = auto()
    EFFECTIVE_DENSITY = auto()


def uniform_particles(mesh_size, device=None) -> torch.Tensor:
    """Lattice positions (X, Y, Z, 3) in grid units (jaxpm.distributed.uniform_particles)."""
    axes = [torch.arange(n, dtype=torch.float32, device=device) for n in mesh_size]
    return torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1).contiguous()


@dataclasses.dataclass(frozen=True, eq=False)
class AbstractField:
    array: Any
    mesh_size: tuple
    box_size: tuple
    z_sources: Any = None
    scale_factors: Any = None
    comoving_centers: Any = None
    density_width: Any = None
    observer_position: tuple = (0.5, 0.5, 0.5)
    sharding: Any = None
    halo_size: tuple = (0, 0)
    nside: int | None = None
    flatsky_npix: tuple | None = None
    field_size: float | None = None
    status: FieldStatus = FieldStatus.UNKNOWN
    unit: Any = PhysicalUnit.INVALID_UNIT

    def __post_init__(self):
        # validation of _core.py:201-229
        ms, bs = self.mesh_size, self.box_size
        if not (isinstance(ms, tuple) and len(ms) == 3):
            raise ValueError(f"mesh_size must be a tuple of 3, got {ms}")
        if not (isinstance(bs, tuple) and len(bs) == 3):
            raise ValueError(f"box_size must be a tuple of 3, got {bs}")
        op = self.observer_position
        if not (isinstance(op, tuple) and len(op) == 3):
            raise ValueError(f"observer_position must be a tuple of 3, got {op}")
        if any(not 0.0 <= f <= 1.0 for f in op):
            raise ValueError("observer_position entries must lie within [0, 1]")
        if not (isinstance(self.halo_size, tuple) and len(self.halo_size) == 2) or any(h < 0 for h in self.halo_size):
            raise ValueError("halo_size must be a tuple of 2 non-negative ints")
        if self.nside is not None and self.nside <= 0:
            raise ValueError(f"nside must be positive, got {self.nside}")
        if self.flatsky_npix is not None and not (len(self.flatsky_npix) == 2 and all(x > 0 for x in self.flatsky_npix)):
            raise ValueError(f"flatsky_npix must be 2 positive ints, got {self.flatsky_npix}")
        if self.sharding is not None:
            raise NotImplementedError("sharded fields are not supported on this single-GPU engine")

    # ------------------------------------------------------------------ pytree-like helpers
    def replace(self, **kw):
        return dataclasses.replace(self, **kw)

    @classmethod
    def FromDensityMetadata(cls, array, field: "AbstractField", **kw):
        meta = {f.name: getattr(field, f.name) for f in dataclasses.fields(AbstractField) if f.name != "array"}
        for k in ("status", "unit"):
            meta.pop(k)
        meta.update(kw)
        return cls(array=array, **meta)

    @property
    def observer_position_mpc(self) -> tuple:
        return tuple(f * l for f, l in zip(self.observer_position, self.box_size))

    @property
    def max_comoving_radius(self) -> float:
        box = np.asarray(self.box_size, dtype=np.float64)
        f = np.clip(np.asarray(self.observer_position, dtype=np.float64), 0.0, 1.0)
        return float(np.min(box / (1.0 + 2.0 * np.minimum(f, 1.0 - f))))

    @property
    def shape(self):
        return tuple(self.array.shape)

    def _base_ndim(self) -> int:
        return 3

    def is_batched(self) -> bool:
        return self.array.ndim > self._base_ndim()

    def __getitem__(self, key):
        if not self.is_batched():
            raise ValueError("indexing needs a batched field")
        n = self.array.shape[0]
        if isinstance(key, slice) and key.step is not None and key.step < 0:  # torch has no negative steps
            key = np.arange(n)[key].copy()

        def pick(v):
            if isinstance(v, (torch.Tensor, np.ndarray)) and np.ndim(v) >= 1 and v.shape[0] == n:
                if isinstance(v, torch.Tensor) and isinstance(key, np.ndarray):
                    return v[torch.as_tensor(key, device=v.device)]
                return v[key]
            return v
        return self.replace(array=pick(self.array), scale_factors=pick(self.scale_factors),
                            z_sources=pick(self.z_sources), comoving_centers=pick(self.comoving_centers),
                            density_width=pick(self.density_width))

    def apply_fn(self, fn):
        return self.replace(array=fn(self.array))

    # arithmetic on the array (_core.py:84-115)
    def _binop(self, other, op):
        o = other.array if isinstance(other, AbstractField) else other
        return self.replace(array=op(self.array, o))

    def __add__(self, o):
        return self._binop(o, torch.add)

    def __sub__(self, o):
        return self._binop(o, torch.sub)

    def __mul__(self, o):
        return self._binop(o, torch.mul)

    def __truediv__(self, o):
        return self._binop(o, torch.div)

    __radd__ = __add__
    __rmul__ = __mul__

    def __neg__(self):
        return self.replace(array=-self.array)


def flat_slab(field, center, width):
    """(zlo, zhi, scale) of the flat-sky slab paint (_painting.py:164-202): the slab
    (c - w/2, c + w/2] in grid units of z, float32 bounds, and 1 / pixel volume."""
    nz, Lz = field.mesh_size[2], field.box_size[2]
    fnx, fny = field.flatsky_npix
    wg, cg = float(width) * nz / Lz, float(center) * nz / Lz
    pixvol = (field.box_size[0] / fnx) * (field.box_size[1] / fny) * float(width)
    return float(np.float32(cg - wg / 2)), float(np.float32(cg + wg / 2)), 1.0 / pixvol


def _paint_mode(unit: PositionUnit) -> str:
    if unit == PositionUnit.MPC_H:
        raise ValueError("Cannot paint ParticleField with unit MPC_H; convert to GRID_RELATIVE or GRID_ABSOLUTE first.")
    return "relative" if unit == PositionUnit.GRID_RELATIVE else "absolute"


@dataclasses.dataclass(frozen=True, eq=False)
class ParticleField(AbstractField):
    """Particle displacements (GRID_RELATIVE), grid positions (GRID_ABSOLUTE) or Mpc/h positions."""

    def __post_init__(self):
        super().__post_init__()
        s = tuple(self.array.shape)
        if not (len(s) in (4, 5, 6) and s[-1] == 3):
            raise ValueError(f"ParticleField array must have shape (X, Y, Z, 3), (S, X, Y, Z, 3), or (N, S, X, Y, Z, 3); got shape {s}")
        if not isinstance(self.unit, PositionUnit):
            raise TypeError(f"ParticleField.unit must be a PositionUnit, got {self.unit!r}")

    def _base_ndim(self) -> int:
        return 4

    # -------------------------------------------------------------- unit conversion (units.py:78-129)
    def to(self, unit: PositionUnit) -> "ParticleField":
        if unit == self.unit:
            return self
        ms = torch.tensor(self.mesh_size, dtype=torch.float32, device=self.array.device)
        bs = torch.tensor(self.box_size, dtype=torch.float32, device=self.array.device)
        arr = self.array
        lattice = None
        if PositionUnit.GRID_RELATIVE in (self.unit, unit):
            # the particle lattice: (i, j, k) * mesh / np (one particle per cell when np = mesh)
            npart = tuple(int(n) for n in arr.shape[-4:-1])
            spacing = torch.tensor([m / n for m, n in zip(self.mesh_size, npart)], dtype=torch.float32,
                                   device=arr.device)
            lattice = uniform_particles(npart, arr.device) * spacing
        if self.unit == PositionUnit.GRID_RELATIVE:
            grid = arr + lattice
        elif self.unit == PositionUnit.MPC_H:
            grid = arr / bs * ms
        else:
            grid = arr
        if unit == PositionUnit.GRID_ABSOLUTE:
            out = grid
        elif unit == PositionUnit.GRID_RELATIVE:
            out = grid - lattice
        else:
            out = grid / ms * bs
        return self.replace(array=out.contiguous(), unit=unit)

    # -------------------------------------------------------------- 3D paint / readout
    def paint(self, *, mesh=None, weights=1.0, chunk_size: int = 2**24, batch_size=None) -> "DensityField":
        """CIC paint (particles.py:120-178); relative -> cic_paint_dx, absolute -> cic_paint."""
        mode = _paint_mode(self.unit)
        data = self.array
        batched = data.ndim == 5
        if data.ndim == 4:
            data = data[None]
        elif data.ndim != 5:
            raise ValueError(f"paint() expects 4D or 5D array, got shape {tuple(data.shape)}")
        outs = []
        for b in range(data.shape[0]):
            out = None if mesh is None else (mesh.array if isinstance(mesh, AbstractField) else mesh).clone()
            outs.append(engine.paint(data[b].contiguous(), self.mesh_size, mode=mode, weights=weights, out=out,
                                     accumulate=mesh is not None))
        painted = torch.stack(outs) if batched else outs[0]
        return DensityField(array=painted, mesh_size=self.mesh_size, box_size=self.box_size,
                            observer_position=self.observer_position, nside=self.nside,
                            flatsky_npix=self.flatsky_npix, halo_size=self.halo_size,
                            status=FieldStatus.DENSITY_FIELD, scale_factors=self.scale_factors,
                            unit=DensityUnit.DENSITY)

    def read_out(self, density_mesh: "DensityField") -> "DensityField":
        """CIC readout of a mesh at the particles (particles.py:183-229)."""
        mode = _paint_mode(self.unit)
        mesh = density_mesh.array if isinstance(density_mesh, AbstractField) else density_mesh
        vals = engine.readout(mesh.contiguous(), self.array.contiguous(), mode=mode)
        return DensityField(array=vals, mesh_size=self.mesh_size, box_size=self.box_size,
                            observer_position=self.observer_position, nside=self.nside,
                            flatsky_npix=self.flatsky_npix, halo_size=self.halo_size, status=self.status,
                            scale_factors=self.scale_factors, unit=DensityUnit.DENSITY)

    # -------------------------------------------------------------- flat-sky (particles.py:234-322)
    def _shell_sets(self, centers, widths, what):
        """Particle sets painted into every shell (particles.py:256-283, 363-392): a 5D batch paints all
        of its S sets into each of its S shells (the reference's lightcone mode); a 4D field paints its
        one set into one shell, or into several when it is a lightcone with per-particle scale factors."""
        data = self.array
        if centers.shape != widths.shape:
            raise ValueError(f"center and density_plane_width must have the same shape; got {centers.shape} and "
                             f"{widths.shape}")
        if data.ndim == 5:
            if centers.size != data.shape[0]:
                raise ValueError(f"Batched input: center must have {data.shape[0]} elements, got {centers.size}")
            return [data[s] for s in range(data.shape[0])]
        if data.ndim == 4:
            if centers.size != 1:
                sf = self.scale_factors
                if sf is None or tuple(np.shape(np.squeeze(np.asarray(sf)))) != tuple(self.mesh_size) or \
                        self.status != FieldStatus.LIGHTCONE:
                    raise ValueError("Painting with mutiple centers/widths requires batched input data ")
            return [data]
        raise ValueError(f"{what}() expects 4D or 5D array, got shape {tuple(data.shape)}")

    def paint_2d(self, center, density_plane_width, *, weights=None, batch_size=None) -> "FlatDensity":
        if self.flatsky_npix is None:
            raise ValueError("ParticleField requires `flatsky_npix` for paint_2d.")
        mode = _paint_mode(self.unit)
        centers = np.atleast_1d(np.asarray(center, dtype=np.float64))
        widths = np.atleast_1d(np.asarray(density_plane_width, dtype=np.float64))
        sets = self._shell_sets(centers, widths, "paint_2d")
        maps = []
        for b in range(centers.size):
            zlo, zhi, scale = flat_slab(self, centers[b], widths[b])
            m = None
            for k, part in enumerate(sets):
                m = engine.paint_flat(part.contiguous(), self.mesh_size, self.flatsky_npix, zlo, zhi, scale,
                                      mode=mode, weights=1.0 if weights is None else weights, out=m,
                                      accumulate=k > 0)
            maps.append(m)
        arr = torch.stack(maps) if len(maps) > 1 else maps[0]
        return FlatDensity.FromDensityMetadata(array=arr, field=self, status=FieldStatus.LIGHTCONE,
                                               comoving_centers=centers, density_width=widths,
                                               unit=DensityUnit.DENSITY)

    # -------------------------------------------------------------- HEALPix (particles.py:326-442)
    def paint_spherical(self, center, density_plane_width, *, scheme="bilinear", weights=None,
                        kernel_width_arcmin=None, smoothing_interpretation="fwhm", paint_nside=None,
                        ud_grade_power=0.0, ud_grade_order_in="RING", ud_grade_order_out="RING",
                        ud_grade_pess=False, batch_size=None, tiles=None) -> "SphericalDensity":
        """HEALPix shell maps (particles.py:326-442). `tiles` (an extension): (rotation codes, shifts
        (n, 3) Mpc/h) sums the shell paint over that table of images of the periodic box."""
        if self.nside is None:
            raise ValueError("Spherical painting requires `nside`.")
        # kernel_width_arcmin / smoothing_interpretation set the rbf_neighbor width (unused by ngp, bilinear)
        sigma = engine.rbf_sigma(kernel_width_arcmin, smoothing_interpretation) if scheme == "rbf_neighbor" else 0.0
        pn = self.nside if paint_nside is None else int(paint_nside)
        mode = _paint_mode(self.unit)
        centers = np.atleast_1d(np.asarray(center, dtype=np.float64))
        widths = np.atleast_1d(np.asarray(density_plane_width, dtype=np.float64))
        sets = self._shell_sets(centers, widths, "paint_spherical")
        maps = []
        for b in range(centers.size):
            c, wdt = centers[b], widths[b]
            m = None
            for k, part in enumerate(sets):
                m = engine.paint_healpix(part.contiguous(), self.mesh_size, self.box_size, self.observer_position_mpc,
                                         c - wdt / 2, c + wdt / 2, pn, scheme=scheme, mode=mode,
                                         weights=1.0 if weights is None else weights, tiles=tiles, out=m,
                                         accumulate=k > 0, sigma=sigma)
            if paint_nside is not None:  # painted at paint_nside, then resampled to the field's nside
                m = engine.healpix_ud_grade(m, self.nside, power=ud_grade_power, order_in=ud_grade_order_in,
                                            order_out=ud_grade_order_out, pess=ud_grade_pess)
            maps.append(m)
        arr = torch.stack(maps) if len(maps) > 1 else maps[0]
        return SphericalDensity.FromDensityMetadata(array=arr, field=self, status=FieldStatus.LIGHTCONE,
                                                    comoving_centers=centers, density_width=widths,
                                                    unit=DensityUnit.DENSITY)


@dataclasses.dataclass(frozen=True, eq=False)
class DensityField(AbstractField):
    def __post_init__(self):
        super().__post_init__()
        s = tuple(self.array.shape)
        if len(s) >= 3 and tuple(s[-3:]) != tuple(self.mesh_size) and self.status not in (
                FieldStatus.LPT1, FieldStatus.LPT2, FieldStatus.PARTICLES):
            raise ValueError(f"DensityField array shape {s} does not match mesh_size {self.mesh_size}")

    def _pairs(self, other, what):
        a = self.array
        if a.ndim not in (3, 4):
            raise ValueError(f"DensityField.{what} expects array shape (X,Y,Z) or (B,X,Y,Z)")
        if other is None:
            return [(x, None) for x in ([a] if a.ndim == 3 else list(a))]
        b = other.array
        if tuple(b.shape) != tuple(a.shape):
            raise ValueError(f"other array must match shape for {what}")
        return [(a, b)] if a.ndim == 3 else list(zip(list(a), list(b)))

    def power(self, mesh2=None, *, kedges=None, multipoles=0, los=(0.0, 0.0, 1.0), batch_size=None):
        """3D power spectrum P(k) (fields/density.py:276-320): auto, or cross with mesh2 (pairwise over
        a batch), for the pole(s) `multipoles` about `los`."""
        from .power import PowerSpectrum, power

        ks, pks = [], []
        for x, y in self._pairs(mesh2, "power"):
            k, pk = power(x, y, box_shape=self.box_size, kedges=kedges, multipoles=multipoles, los=los)
            ks.append(k)
            pks.append(pk)
        pk = pks[0] if self.array.ndim == 3 else np.stack(pks)
        return PowerSpectrum(wavenumber=ks[0], spectra=pk, name="pk", scale_factors=self.scale_factors)

    def cross_power(self, *, kedges=None, multipoles=0, los=(0.0, 0.0, 1.0), batch_size=None):
        """All B(B+1)/2 spectra of a batch of B >= 2 fields in upper-triangular order (0,0), (0,1), ...,
        (B-1,B-1) (fields/density.py:322-402)."""
        from .power import PowerSpectrum, power

        data = self.array
        if data.ndim != 4:
            raise ValueError(f"cross_power requires batched array with shape (B, X, Y, Z), got array with "
                             f"{data.ndim} dimensions. Use power() for single fields.")
        nb = data.shape[0]
        if nb < 2:
            raise ValueError(f"cross_power requires at least 2 maps in batch dimension, got {nb}. "
                             "Use power() for single fields.")
        poles = tuple(multipoles) if isinstance(multipoles, (list, tuple)) else (multipoles,)
        k, pks = None, []
        for i in range(nb):
            for j in range(i, nb):
                k, pk = power(data[i], data[j], box_shape=self.box_size, kedges=kedges, multipoles=poles, los=los)
                pks.append(pk)
        return PowerSpectrum(wavenumber=k, spectra=np.stack(pks), name="cross_pk", scale_factors=self.scale_factors)

    def transfer(self, other, *, kedges=None, batch_size=None):
        """Monopole transfer function sqrt(P_other / P_self) (fields/density.py:404-438)."""
        from .power import PowerSpectrum, transfer

        out = [transfer(x, y, box_shape=self.box_size, kedges=kedges) for x, y in self._pairs(other, "transfer")]
        sp = out[0][1] if self.array.ndim == 3 else np.stack([o[1] for o in out])
        return PowerSpectrum(wavenumber=out[0][0], spectra=sp, name="transfer", scale_factors=self.scale_factors)

    def coherence(self, other, *, kedges=None, batch_size=None):
        """Monopole coherence P01 / sqrt(P0 P1) (fields/density.py:440-475)."""
        from .power import PowerSpectrum, coherence

        out = [coherence(x, y, box_shape=self.box_size, kedges=kedges) for x, y in self._pairs(other, "coherence")]
        sp = out[0][1] if self.array.ndim == 3 else np.stack([o[1] for o in out])
        return PowerSpectrum(wavenumber=out[0][0], spectra=sp, name="coherence", scale_factors=self.scale_factors)


@dataclasses.dataclass(frozen=True, eq=False)
class FlatDensity(AbstractField):
    def _base_ndim(self) -> int:
        return 2


@dataclasses.dataclass(frozen=True, eq=False)
class SphericalDensity(AbstractField):
    def __post_init__(self):
        super().__post_init__()
        if self.nside is None:
            raise ValueError("SphericalDensity requires `nside`.")
        s = tuple(self.array.shape)
        if s != () and (len(s) not in (1, 2, 3) or s[-1] != 12 * self.nside**2):
            raise ValueError(f"SphericalDensity array shape {s} does not match npix {12 * self.nside**2}")

    def _base_ndim(self) -> int:
        return 1


@dataclasses.dataclass(frozen=True, eq=False)
class FlatKappaField(FlatDensity):
    """Convergence map, flat sky (fields/lensing_maps.py:19): (ny, nx) or (n_sources, ny, nx)."""


@dataclasses.dataclass(frozen=True, eq=False)
class SphericalKappaField(SphericalDensity):
    """Convergence map, HEALPix RING (fields/lensing_maps.py:72): (npix,) or (n_sources, npix)."""
