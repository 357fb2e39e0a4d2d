"""Snapshot interpolation kernels (src/jax_fli/pm/interp.py:46-699): the abstract interface,
``NoInterp`` (paint the particles as they are at each snapshot), ``DriftInterp`` (spherical shells
with every particle drifted to its own lightcone crossing) and ``OnionTiler`` (spherical shells
beyond the box painted from rotated periodic images of it, optionally drifted). The drift runs inside
the HEALPix paint kernel from host-built float64 tables of a(chi), Gp(a) and dGfa(a); and
``TelephotoInterp`` (one rotated, shifted image of the box per shell, for flat or spherical
targets)."""

from __future__ import annotations

import dataclasses
import itertools
import warnings
from typing import Any

import numpy as np
import torch

from .. import cosmology as _c
from .. import engine
from .. import prng
from ..fields import DensityUnit, FieldStatus, PositionUnit, SphericalDensity, _paint_mode, flat_slab
from ..painting import PaintingOptions

DRIFT_TABLE = 4096


def drift_tables(cosmo, a_now: float, chi_max: float, device) -> dict:
    """Tables of _drift_to_time (interp.py:110-121) for particles at time a_now: a(chi) on
    [0, chi_max], Gp(a) and dGfa(a) on [a(chi_max), 1] (float64, linear interpolation on the device),
    and a bound on |(Gp(a_t) - Gp(a_now)) / dGfa(sqrt(a_now a_t))| over that range."""
    chi = np.linspace(0.0, chi_max, DRIFT_TABLE)
    a_chi = np.asarray(_c.a_of_chi(cosmo, chi), dtype=np.float64)
    a_lo, a_hi = float(min(a_chi[-1], a_now)), 1.0
    ag = np.linspace(a_lo, a_hi, DRIFT_TABLE)
    gp_t = np.asarray(_c.Gp(cosmo, ag), dtype=np.float64)
    dg_t = np.asarray(_c.dGfa(cosmo, ag), dtype=np.float64)
    gp_now = float(np.asarray(_c.Gp(cosmo, a_now)).reshape(-1)[0])
    f = (gp_t - gp_now) / np.asarray(_c.dGfa(cosmo, np.sqrt(a_now * ag)), dtype=np.float64)
    tab = engine.to_device(np.stack([a_chi, gp_t, dg_t]), device)
    return {"tables": tab.contiguous(), "chi_max": chi_max, "a_lo": a_lo, "a_hi": a_hi, "a_now": float(a_now),
            "gp_now": gp_now, "bound": float(np.abs(f).max()) * 1.01 + 1e-12}


def paint_nside_of(field, painting) -> int:
    """nside a shell is painted at: the painting's paint_nside, else the field's."""
    pn = getattr(painting, "paint_nside", None)
    return int(field.nside if pn is None else pn)


def ud_grade_to(m, field, painting):
    """A map painted at paint_nside brought to the field's nside (healpy.ud_grade with the painting's
    power, orders and pess); unchanged when no paint_nside is set."""
    if getattr(painting, "paint_nside", None) is None:
        return m
    return engine.healpix_ud_grade(m, field.nside, power=painting.ud_grade_power, order_in=painting.ud_grade_order_in,
                                   order_out=painting.ud_grade_order_out, pess=painting.ud_grade_pess)


def ud_grade_adjoint(map_bar, field, painting):
    """Transpose of ud_grade_to for maps without bad pixels: ud_grade back to paint_nside with power
    -power - 2 and the orders swapped (degrading averages the children, upgrading copies the parent)."""
    if getattr(painting, "paint_nside", None) is None:
        return map_bar
    return engine.healpix_ud_grade(map_bar, paint_nside_of(field, painting),
                                   power=-float(painting.ud_grade_power) - 2.0,
                                   order_in=painting.ud_grade_order_out, order_out=painting.ud_grade_order_in)


def paint_drifted(dx, p, t, cosmo, r_center, width, painting, tiles=None):
    """Spherical shell map of the particles each drifted to the lightcone (DriftInterp.paint and
    OnionTiler with drift_on_lightcone, interp.py:258-319, 375-419, 476-493)."""
    if dx.nside is None:
        raise ValueError("Spherical painting requires `nside`.")
    rmin, rmax = r_center - width / 2, r_center + width / 2
    drift = drift_tables(cosmo, float(t), 2.0 * rmax + 100.0, dx.array.device)
    arr = engine.paint_healpix_drift(dx.array.contiguous(), p.array.contiguous(), dx.mesh_size, dx.box_size,
                                     dx.observer_position_mpc, rmin, rmax, paint_nside_of(dx, painting), drift,
                                     scheme=painting.scheme, mode=_paint_mode(dx.unit),
                                     weights=1.0 if painting.weights is None else painting.weights, tiles=tiles,
                                     sigma=engine.rbf_sigma(painting.kernel_width_arcmin,
                                                            painting.smoothing_interpretation)
                                     if painting.scheme == "rbf_neighbor" else 0.0)
    arr = ud_grade_to(arr, dx, painting)
    return SphericalDensity.FromDensityMetadata(array=arr, field=dx, status=FieldStatus.LIGHTCONE,
                                                comoving_centers=np.atleast_1d(r_center),
                                                density_width=np.atleast_1d(width), unit=DensityUnit.DENSITY)


def drift_factor_table(cosmo, a_now: float, chi_max: float) -> np.ndarray:
    """f(chi) = (Gp(a(chi)) - Gp(a_now)) / dGfa(sqrt(a_now a(chi))) on the drift tables' chi grid."""
    chi = np.linspace(0.0, chi_max, DRIFT_TABLE)
    a_chi = np.asarray(_c.a_of_chi(cosmo, chi), dtype=np.float64)
    gp_now = float(np.asarray(_c.Gp(cosmo, a_now)).reshape(-1)[0])
    return (np.asarray(_c.Gp(cosmo, a_chi), dtype=np.float64) - gp_now) / np.asarray(
        _c.dGfa(cosmo, np.sqrt(a_now * a_chi)), dtype=np.float64)


def paint_drifted_vjp(dx, p, t, cosmo, r_center, width, painting, map_bar, tiles=None, param_names=()):
    """(position, velocity, parameter dots or None) cotangents of paint_drifted for the map cotangent
    map_bar; param_names: cosmological parameters differentiated through the drift factor (float64
    differences of its table)."""
    from .. import autodiff

    rmin, rmax = r_center - width / 2, r_center + width / 2
    chi_max = 2.0 * rmax + 100.0
    dev = dx.array.device
    drift = drift_tables(cosmo, float(t), chi_max, dev)
    ptab = pdots = None
    if param_names:
        ptab = engine.to_device(np.stack([autodiff.fd(lambda c: drift_factor_table(c, float(t), chi_max), cosmo, n)
                                          for n in param_names]), dev)
        pdots = torch.zeros(len(param_names), dtype=torch.float64, device=dev)
    map_bar = ud_grade_adjoint(map_bar.contiguous(), dx, painting)
    gx, gv = engine.paint_healpix_drift_vjp(dx.array.contiguous(), p.array.contiguous(), map_bar.contiguous(),
                                            dx.mesh_size, dx.box_size, dx.observer_position_mpc, rmin, rmax,
                                            paint_nside_of(dx, painting), drift, scheme=painting.scheme,
                                            mode=_paint_mode(dx.unit),
                                            weights=1.0 if painting.weights is None else painting.weights,
                                            tiles=tiles, param_tables=ptab, param_dots=pdots)
    return gx, gv, pdots


def _image_spec(interp, state, dx, t, cosmo):
    """(rotation code, shift, drift tables or None) of the image a flat, density or particle snapshot of
    DriftInterp / TelephotoInterp is painted from."""
    tiles = interp.images(state, dx) if isinstance(interp, TelephotoInterp) else None
    code, shift = (0, np.zeros(3)) if tiles is None else (int(tiles[0][0]), np.asarray(tiles[1][0], np.float64))
    drift = None
    if isinstance(interp, DriftInterp) or getattr(interp, "drift_on_lightcone", False):
        reach = float(np.linalg.norm(shift)) + 2.0 * float(np.linalg.norm(dx.box_size)) + 100.0
        drift = drift_tables(cosmo, float(t), reach, dx.array.device)
    return code, shift, drift


def _image_field(interp, state, dx, p, t, cosmo):
    """The snapshot's particles as one (rotated, shifted, drifted) image in absolute grid units."""
    code, shift, drift = _image_spec(interp, state, dx, t, cosmo)
    g = engine.image_positions(dx.array.contiguous(), dx.mesh_size, dx.box_size, dx.observer_position_mpc, rot=code,
                               shift=shift, mode=_paint_mode(dx.unit), vel=p.array.contiguous() if drift else None,
                               drift=drift)
    return dx.replace(array=g, unit=PositionUnit.GRID_ABSOLUTE)


def image_snapshot_vjp(interp, state, dx, p, t, cosmo, g, param_names=()):
    """(position, velocity or None, parameter dots or None) cotangents of a flat, density or particle
    snapshot painted from an image (DriftInterp, TelephotoInterp) for its cotangent g: the target's
    paint adjoint at the image positions, then pmx_image_positions_vjp through the transform and drift
    (and the drift factor's dependence on the cosmological parameters param_names)."""
    from .. import autodiff

    code, shift, drift = _image_spec(interp, state, dx, t, cosmo)
    vel = p.array.contiguous() if drift else None
    img = engine.image_positions(dx.array.contiguous(), dx.mesh_size, dx.box_size, dx.observer_position_mpc,
                                 rot=code, shift=shift, mode=_paint_mode(dx.unit), vel=vel, drift=drift)
    pt = interp.painting
    weights = 1.0 if pt.weights is None else pt.weights
    if pt.target == "flat":
        r_center, width = interp.shell(state)
        zlo, zhi, scale = flat_slab(dx, r_center, width)
        gi = engine.paint_flat_vjp(img, g.contiguous(), dx.mesh_size, zlo, zhi, scale, mode="absolute",
                                   weights=weights)
    elif pt.target == "density":
        gi = engine.paint_vjp(img, g.contiguous(), mode="absolute", weights=weights)[0]
    else:  # particles, returned in Mpc/h
        cell = torch.tensor([b / m for b, m in zip(dx.box_size, dx.mesh_size)], dtype=torch.float32, device=img.device)
        gi = (g * cell).contiguous()
    ptab = pdots = None
    if drift is not None and param_names:
        chi_max = drift["chi_max"]
        ptab = engine.to_device(np.stack([autodiff.fd(lambda c: drift_factor_table(c, float(t), chi_max), cosmo, n)
                                          for n in param_names]), dx.array.device)
        pdots = torch.zeros(len(param_names), dtype=torch.float64, device=dx.array.device)
    gx, gv = engine.image_positions_vjp(dx.array.contiguous(), gi, dx.mesh_size, dx.box_size,
                                        dx.observer_position_mpc, rot=code, shift=shift, mode=_paint_mode(dx.unit),
                                        vel=vel, drift=drift, param_tables=ptab, param_dots=pdots)
    return gx, gv, pdots


@dataclasses.dataclass(frozen=True)
class InterpTilerState:
    rotation_idx: int = -1
    shell_idx: int = -1
    rotations: Any = None


@dataclasses.dataclass(frozen=True)
class AbstractInterp:
    painting: PaintingOptions
    ts: Any = None
    r_centers: Any = None
    density_widths: Any = None
    max_comoving_distance: float | None = None

    def update_geometry(self, ts, r_centers, density_widths, max_comoving_distance):
        return dataclasses.replace(self, ts=ts, r_centers=r_centers, density_widths=density_widths,
                                   max_comoving_distance=max_comoving_distance)

    def init(self) -> InterpTilerState:
        raise NotImplementedError

    def advance(self, state: InterpTilerState, t) -> InterpTilerState:
        raise NotImplementedError

    def rewind(self, state: InterpTilerState, t) -> InterpTilerState:
        raise NotImplementedError

    def paint(self, state: InterpTilerState, t, y, cosmo):
        raise NotImplementedError

    def shell(self, state: InterpTilerState):
        """(r_center, width) of the current shell (Mpc/h)."""
        return (float(np.asarray(self.r_centers).reshape(-1)[state.shell_idx]),
                float(np.asarray(self.density_widths).reshape(-1)[state.shell_idx]))

    def images(self, state: InterpTilerState, field):
        """Image table (rotation codes, shifts (n, 3) Mpc/h) the current shell is painted from, or None
        for the box alone."""
        return None


@dataclasses.dataclass(frozen=True)
class NoInterp(AbstractInterp):
    def init(self) -> InterpTilerState:
        if self.r_centers is not None and self.density_widths is not None and self.max_comoving_distance is not None:
            far = float(np.asarray(self.r_centers).reshape(-1)[0] + np.asarray(self.density_widths).reshape(-1)[0] / 2)
            if far > self.max_comoving_distance:
                warnings.warn(
                    f"NoInterp does not support tiling. Your furthest shell is at {far:.2f} Mpc/h, but your box only "
                    f"extends to {self.max_comoving_distance:.2f} Mpc/h. Your simulations will run but will have "
                    f"artifacts for the far shells.")
        return InterpTilerState(rotation_idx=-1, shell_idx=-1, rotations=None)

    def advance(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx + 1) if hit else state

    def rewind(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx - 1) if hit else state

    def paint(self, state, t, y, cosmo):
        dx, p = y
        r_center = float(np.asarray(self.r_centers).reshape(-1)[state.shell_idx])
        width = float(np.asarray(self.density_widths).reshape(-1)[state.shell_idx])
        tgt = None if self.painting is None else self.painting.target
        if tgt is None or tgt == "particles":
            result = dx.replace(array=dx.array.clone())
        elif tgt == "spherical":
            pt = self.painting
            result = dx.paint_spherical(center=r_center, density_plane_width=width, scheme=pt.scheme,
                                        weights=pt.weights, kernel_width_arcmin=pt.kernel_width_arcmin,
                                        smoothing_interpretation=pt.smoothing_interpretation,
                                        paint_nside=pt.paint_nside, ud_grade_power=pt.ud_grade_power,
                                        ud_grade_order_in=pt.ud_grade_order_in,
                                        ud_grade_order_out=pt.ud_grade_order_out,
                                        ud_grade_pess=pt.ud_grade_pess, batch_size=pt.batch_size)
        elif tgt == "flat":
            result = dx.paint_2d(center=r_center, density_plane_width=width, weights=self.painting.weights,
                                 batch_size=self.painting.batch_size)
        elif tgt == "density":
            result = dx.paint(weights=self.painting.weights, chunk_size=self.painting.chunk_size,
                              batch_size=self.painting.batch_size)
        else:
            result = dx
        return result.replace(scale_factors=float(t), comoving_centers=r_center, density_width=width)


@dataclasses.dataclass(frozen=True)
class DriftInterp(AbstractInterp):
    """No tiling; each particle is drifted from the snapshot time to the time its distance from the
    observer corresponds to before painting (interp.py:212-320): spherical shells drift inside the
    HEALPix paint kernel; flat, density and particle targets take the drifted positions from
    pmx_image_positions (particles are returned in Mpc/h, as the reference returns them)."""

    def init(self) -> InterpTilerState:
        if self.r_centers is not None and self.density_widths is not None and self.max_comoving_distance is not None:
            far = float(np.asarray(self.r_centers).reshape(-1)[0] + np.asarray(self.density_widths).reshape(-1)[0] / 2)
            if far > self.max_comoving_distance:
                warnings.warn(
                    f"DriftInterp does not support tiling. Your furthest shell is at {far:.2f} Mpc/h, but your box "
                    f"only extends to {self.max_comoving_distance:.2f} Mpc/h. Your simulations will run but will "
                    f"have artifacts for the far shells.")
        return InterpTilerState(rotation_idx=-1, shell_idx=-1, rotations=None)

    def advance(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx + 1) if hit else state

    def rewind(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx - 1) if hit else state

    def paint(self, state, t, y, cosmo):
        dx, p = y
        r_center, width = self.shell(state)
        pt = self.painting
        tgt = None if pt is None else pt.target
        if tgt == "spherical":
            out = paint_drifted(dx, p, t, cosmo, r_center, width, pt)
            return out.replace(scale_factors=float(t), comoving_centers=r_center, density_width=width)
        # the drifted particles as absolute grid positions (interp.py:277-289), then the target's paint
        drifted = _image_field(self, state, dx, p, t, cosmo)
        if tgt == "flat":
            out = drifted.paint_2d(center=r_center, density_plane_width=width, weights=pt.weights,
                                   batch_size=pt.batch_size)
        elif tgt == "density":
            out = drifted.paint(weights=1.0 if pt.weights is None else pt.weights, chunk_size=pt.chunk_size,
                                batch_size=pt.batch_size)
        else:
            out = drifted.to(PositionUnit.MPC_H)
        return out.replace(scale_factors=float(t), comoving_centers=r_center, density_width=width)


@dataclasses.dataclass(frozen=True)
class OnionTiler(AbstractInterp):
    """Spherical shells from images of the periodic box (src/jax_fli/pm/interp.py:323-504): a shell
    inside the box (r_center + width/2 <= max_comoving_distance) is painted from the box alone;
    a farther one from the 27 images at offsets product([0, 1, -1], repeat=3) x box (shifted +1 / -1
    on axes where the observer sits on the face at 0 / 1), each rotated by a random cube symmetry,
    all summed into one map by a single pmx_paint_healpix_tiled launch.

    The rotations are the reference's draw, jax.random.randint(PRNGKey(seed), (n,), 0, 47) with
    seed 42 (interp.py:334-335), reproduced on the host by jax_fli_amd.prng (Threefry-2x32 with JAX's
    partitionable key conventions; int32 as under jax_enable_x64=False, int64 with x64=True); index i
    of the reference's 47-entry table is entry i + 1 of the 48-entry one here.

    Differences from the reference, by design: particles are wrapped into the box before imaging (the
    reference images unwrapped positions); `layers` > 1 extends the offsets to [-layers, layers]^3 (and
    the draw to (2 layers + 1)^3 rotations) so shells beyond 1.5 boxes are covered, and `rotate=False`
    gives plain periodic replication. With drift_on_lightcone (the reference default) every image is
    drifted to its own lightcone time inside the same paint kernel."""

    drift_on_lightcone: bool = True
    layers: int = 1
    rotate: bool = True
    seed: int = 42
    x64: bool = False

    def init(self) -> InterpTilerState:
        if self.painting.target != "spherical":
            raise ValueError("OnionTiler only supports spherical painting.")
        if self.layers < 1:
            raise ValueError("layers must be >= 1")
        n = (2 * self.layers + 1) ** 3
        if self.rotate:  # entries 1..47 of the symmetry table (0 is the identity)
            rots = prng.randint(prng.prng_key(self.seed), n, 0, 47, x64=self.x64).astype(np.int64) + 1
        else:
            rots = np.zeros(n, dtype=np.int64)
        return InterpTilerState(rotation_idx=-1, shell_idx=-1, rotations=rots.astype(np.int32))

    def advance(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx + 1) if hit else state

    def rewind(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        return dataclasses.replace(state, shell_idx=state.shell_idx - 1) if hit else state

    def images(self, state, field):
        r_center, width = self.shell(state)
        if self.max_comoving_distance is not None and r_center + width / 2 <= self.max_comoving_distance:
            return None
        order = [0] + [s * k for k in range(1, self.layers + 1) for s in (1, -1)]
        shifts = np.array(list(itertools.product(order, repeat=3)), dtype=np.float64)
        f = np.asarray(field.observer_position, dtype=np.float64)
        shifts = np.where(f == 0, shifts + 1, shifts)
        shifts = np.where(f == 1, shifts - 1, shifts)
        return np.asarray(state.rotations, dtype=np.int32), shifts * np.asarray(field.box_size, dtype=np.float64)

    def paint(self, state, t, y, cosmo):
        dx, p = y
        r_center, width = self.shell(state)
        pt = self.painting
        if self.drift_on_lightcone:
            out = paint_drifted(dx, p, t, cosmo, r_center, width, pt, tiles=self.images(state, dx))
            return out.replace(scale_factors=float(t), comoving_centers=r_center, density_width=width)
        result = dx.paint_spherical(center=r_center, density_plane_width=width, scheme=pt.scheme, weights=pt.weights,
                                    kernel_width_arcmin=pt.kernel_width_arcmin,
                                    smoothing_interpretation=pt.smoothing_interpretation,
                                    paint_nside=pt.paint_nside, ud_grade_power=pt.ud_grade_power,
                                    ud_grade_order_in=pt.ud_grade_order_in, ud_grade_order_out=pt.ud_grade_order_out,
                                    ud_grade_pess=pt.ud_grade_pess, batch_size=pt.batch_size,
                                    tiles=self.images(state, dx))
        return result.replace(scale_factors=float(t), comoving_centers=r_center, density_width=width)


@dataclasses.dataclass(frozen=True)
class TelephotoInterp(AbstractInterp):
    """One rotated, shifted copy of the box per shell, for narrow fields of view (src/jax_fli/pm/
    interp.py:507-699). A shell whose centre is inside the box radius (r_center <=
    max_comoving_distance) is painted from the box itself; a farther one from the box rotated about
    the observer by symmetry (rotation_idx mod 47) of get_all_47_symmetries and moved along the line
    of sight (z) by floor(r_center / max_comoving_distance) * max_comoving_distance, where
    rotation_idx counts the shells painted so far beyond the box radius (from -1). With
    drift_on_lightcone (the default) every particle is drifted to its own lightcone time first.
    Flat targets: pmx_image_positions builds the image's absolute grid positions and the flat slab
    paint takes them; spherical targets: one image of the HEALPix shell paint.

    Differences from the reference, by design: the drift's lightcone time is taken at the image's
    own distance from the observer (the reference measures it from the image minus the observer
    position a second time, interp.py:647, which agrees only for an observer at the origin); rewind
    undoes advance with the same centre test (the reference's rewind tests r_center + width/2,
    interp.py:556-558), so the adjoint paints each shell with the forward's rotation; spherical
    images are painted from box-wrapped positions, as OnionTiler's are."""

    drift_on_lightcone: bool = True

    def init(self) -> InterpTilerState:
        tgt = None if self.painting is None else self.painting.target
        if tgt not in ("spherical", "flat"):
            raise ValueError("TelephotoInterp only supports projections (flat or spherical).")
        if tgt == "spherical":
            warnings.warn("TelephotoInterp is meant for narrow-field flat-sky maps: full-sky HEALPix shells painted "
                          "from one image of the box show strong artifacts.")
        return InterpTilerState(rotation_idx=-1, shell_idx=-1, rotations=np.arange(1, 48, dtype=np.int32))

    def _r_center(self, idx):
        return float(np.asarray(self.r_centers).reshape(-1)[idx])

    def _beyond(self, idx):
        return self._r_center(idx) > self.max_comoving_distance

    def advance(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        if not hit:
            return state
        nxt = state.shell_idx + 1
        return dataclasses.replace(state, shell_idx=nxt, rotation_idx=state.rotation_idx + int(self._beyond(nxt)))

    def rewind(self, state, t):
        hit = self.ts is not None and bool(np.isin(np.float32(t), np.asarray(self.ts, dtype=np.float32)))
        if not hit:
            return state
        return dataclasses.replace(state, shell_idx=state.shell_idx - 1,
                                   rotation_idx=state.rotation_idx - int(self._beyond(state.shell_idx)))

    def rotation_code(self, shell_idx: int) -> int:
        """Cube-symmetry code (1..47) of the shell's image: entry rotation_idx mod 47 of the 47
        non-identity symmetries, rotation_idx = (shells 0..shell_idx beyond the box radius) - 1."""
        rc = np.asarray(self.r_centers, dtype=np.float64).reshape(-1)[:shell_idx + 1]
        return int((int((rc > self.max_comoving_distance).sum()) - 1) % 47) + 1

    def images(self, state, field):
        r_center, _ = self.shell(state)
        if self.max_comoving_distance is None or r_center <= self.max_comoving_distance:
            return None
        d = float(np.floor(r_center / self.max_comoving_distance) * self.max_comoving_distance)
        return (np.array([self.rotation_code(state.shell_idx)], dtype=np.int32),
                np.array([[0.0, 0.0, d]], dtype=np.float64))

    def paint(self, state, t, y, cosmo):
        dx, p = y
        r_center, width = self.shell(state)
        pt = self.painting
        tiles = self.images(state, dx)
        if pt.target == "spherical":
            if self.drift_on_lightcone:
                out = paint_drifted(dx, p, t, cosmo, r_center, width, pt, tiles=tiles)
            else:
                out = dx.paint_spherical(center=r_center, density_plane_width=width, scheme=pt.scheme,
                                         weights=pt.weights, kernel_width_arcmin=pt.kernel_width_arcmin,
                                         smoothing_interpretation=pt.smoothing_interpretation,
                                         paint_nside=pt.paint_nside, ud_grade_power=pt.ud_grade_power,
                                         ud_grade_order_in=pt.ud_grade_order_in,
                                        ud_grade_order_out=pt.ud_grade_order_out,
                                         ud_grade_pess=pt.ud_grade_pess, batch_size=pt.batch_size, tiles=tiles)
        else:
            src = dx
            if tiles is not None or self.drift_on_lightcone:
                src = _image_field(self, state, dx, p, t, cosmo)
            out = src.paint_2d(center=r_center, density_plane_width=width, weights=pt.weights,
                               batch_size=pt.batch_size)
        return out.replace(scale_factors=float(t), comoving_centers=r_center, density_width=width)
