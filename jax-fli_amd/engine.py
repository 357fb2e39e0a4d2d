"""Tensor-level wrappers of the pmx C ABI (device tensors in, device tensors out).

These are the building blocks the reference-shaped API (fields, solvers, lpt, nbody) calls.
All work is enqueued on torch's current stream; nothing synchronises.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, dbl3, get_plan, load, ptr, stream_ptr

MODES = {"relative": 0, "absolute": 1}


def _f32(t: torch.Tensor, name: str) -> torch.Tensor:
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if not t.is_cuda:
        raise _lib.PmxError(f"{name} must live on the GPU")
    return t.contiguous()


def to_device(a, device, dtype=torch.float64) -> torch.Tensor:
    """Host array or tensor -> contiguous tensor on `device`. From host memory to a GPU the copy goes
    through a pinned staging buffer and is asynchronous on the current stream, so the host does not
    wait for the work already queued there (a pageable copy would)."""
    t = torch.as_tensor(a, dtype=dtype)
    device = torch.device(device)
    if t.device == device:
        return t.contiguous()
    if t.device.type == "cpu" and device.type == "cuda":
        return t.contiguous().pin_memory().to(device, non_blocking=True)
    return t.to(device).contiguous()


def _pgrid(pos: torch.Tensor):
    if pos.ndim != 4 or pos.shape[-1] != 3:
        raise ValueError(f"particle array must have shape (X, Y, Z, 3), got {tuple(pos.shape)}")
    return int(pos.shape[0]), int(pos.shape[1]), int(pos.shape[2])


def _weights(weights, npart: int):
    if weights is None:
        return None, 1.0
    if isinstance(weights, (int, float)):
        return None, float(weights)
    w = _f32(torch.as_tensor(weights), "weights")
    if w.numel() == 1:
        return None, float(w.item())
    if w.numel() != npart:
        raise ValueError("weights must be a scalar or one value per particle")
    return w, 1.0


def paint(pos: torch.Tensor, mesh_shape, mode="relative", weights=1.0, out: torch.Tensor | None = None,
          accumulate=False, plan=None) -> torch.Tensor:
    pos = _f32(pos, "pos")
    npx, npy, npz = _pgrid(pos)
    plan = plan or get_plan(mesh_shape, pos.device)
    if out is None:
        out = torch.empty(tuple(mesh_shape), dtype=torch.float32, device=pos.device)
    w, w0 = _weights(weights, npx * npy * npz)
    plan.reserve(npx, npy, npz)
    check(load().pmx_paint_cic(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], ptr(w), w0, ptr(out),
                               int(accumulate), stream_ptr()), "pmx_paint_cic")
    return out


def readout(mesh: torch.Tensor, pos: torch.Tensor, mode="relative") -> torch.Tensor:
    mesh = _f32(mesh, "mesh")
    pos = _f32(pos, "pos")
    npx, npy, npz = _pgrid(pos)
    ncomp = 1 if mesh.ndim == 3 else int(mesh.shape[3])
    plan = get_plan(tuple(mesh.shape[:3]), pos.device)
    out = torch.empty((npx, npy, npz) + ((ncomp,) if mesh.ndim == 4 else ()), dtype=torch.float32, device=pos.device)
    check(load().pmx_readout_cic(plan.handle, ptr(mesh), ncomp, ptr(pos), npx, npy, npz, MODES[mode], ptr(out),
                                 stream_ptr()), "pmx_readout_cic")
    return out


def rfftn(mesh: torch.Tensor) -> torch.Tensor:
    """Unnormalised r2c 3D FFT. Returns complex64 laid out [kz][ky][kx] (nz/2+1, ny, nx)."""
    mesh = _f32(mesh, "mesh")
    nx, ny, nz = mesh.shape
    plan = get_plan(mesh.shape, mesh.device)
    spec = torch.empty((nz // 2 + 1, ny, nx), dtype=torch.complex64, device=mesh.device)
    check(load().pmx_fft_r2c(plan.handle, ptr(mesh), ptr(spec), stream_ptr()), "pmx_fft_r2c")
    return spec


def irfftn(spec: torch.Tensor, mesh_shape) -> torch.Tensor:
    """Unnormalised c2r inverse of rfftn (result = N * x)."""
    if spec.dtype != torch.complex64:
        raise TypeError("spectrum must be complex64")
    spec = spec.contiguous()
    plan = get_plan(mesh_shape, spec.device)
    out = torch.empty(tuple(mesh_shape), dtype=torch.float32, device=spec.device)
    check(load().pmx_fft_c2r(plan.handle, ptr(spec), ptr(out), stream_ptr()), "pmx_fft_c2r")
    return out


def force_mesh(delta: torch.Tensor, scale=1.0, deconv=False) -> torch.Tensor:
    delta = _f32(delta, "density")
    plan = get_plan(delta.shape, delta.device)
    out = torch.empty(tuple(delta.shape) + (3,), dtype=torch.float32, device=delta.device)
    check(load().pmx_force_mesh(plan.handle, ptr(delta), float(scale), _lib.DECONV if deconv else 0, ptr(out),
                                stream_ptr()), "pmx_force_mesh")
    return out


def forces(pos: torch.Tensor, mesh_shape, mode="relative", scale=1.0, deconv=False, atomic_paint=False) -> torch.Tensor:
    pos = _f32(pos, "pos")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    out = torch.empty_like(pos)
    flags = (_lib.DECONV if deconv else 0) | (_lib.PAINT_ATOMIC if atomic_paint else 0)
    plan.reserve(npx, npy, npz)
    check(load().pmx_forces(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], float(scale), flags, ptr(out),
                            stream_ptr()), "pmx_forces")
    return out


def step(x: torch.Tensor, v: torch.Tensor, mesh_shape, mode="relative", force_scale=1.0, c_kick=0.0,
         c_drift_pre=None, c_drift_post=None, deconv=False, atomic_paint=False, plan=None,
         force_mesh: torch.Tensor | None = None) -> None:
    """In place: [x += v*c_drift_pre]; v += F(x)*force_scale*c_kick; [x += v*c_drift_post]. With
    force_mesh (float32 (nx, ny, nz, 3)), the step's force mesh is written there and kept."""
    if not (x.is_contiguous() and v.is_contiguous()):
        raise _lib.PmxError("state tensors must be contiguous")
    _f32(x, "x")
    _f32(v, "v")
    npx, npy, npz = _pgrid(x)
    if v.shape != x.shape:
        raise ValueError("x and v must have the same shape")
    plan = plan or get_plan(mesh_shape, x.device)
    flags = (_lib.DECONV if deconv else 0) | (_lib.PAINT_ATOMIC if atomic_paint else 0)
    flags |= _lib.DRIFT_PRE if c_drift_pre is not None else 0
    flags |= _lib.DRIFT_POST if c_drift_post is not None else 0
    if force_mesh is not None:
        _force_mesh_arg(force_mesh, mesh_shape)
        plan.reserve(npx, npy, npz)
        check(load().pmx_step_keep(plan.handle, ptr(x), ptr(v), npx, npy, npz, MODES[mode], float(force_scale),
                                   float(c_drift_pre or 0.0), float(c_kick), float(c_drift_post or 0.0), flags,
                                   ptr(force_mesh), stream_ptr()), "pmx_step_keep")
        return
    plan.reserve(npx, npy, npz)
    check(load().pmx_step(plan.handle, ptr(x), ptr(v), npx, npy, npz, MODES[mode], float(force_scale),
                          float(c_drift_pre or 0.0), float(c_kick), float(c_drift_post or 0.0), flags,
                          stream_ptr()), "pmx_step")


def _force_mesh_arg(force_mesh: torch.Tensor, mesh_shape) -> None:
    _f32(force_mesh, "force_mesh")
    if tuple(force_mesh.shape) != tuple(int(n) for n in mesh_shape) + (3,) or not force_mesh.is_contiguous():
        raise ValueError("force_mesh must be a contiguous float32 (nx, ny, nz, 3) tensor")


def pgd(x: torch.Tensor, v, mesh_shape, mode="relative", alpha=0.2, kl=0.3, ks=1.0, target="x", v_scale=1.0,
        plan=None):
    """In place: x += S(x) (target "x") or v += S(x) * v_scale (target "v"), S = alpha * PGD force."""
    x = _f32(x, "x")
    npx, npy, npz = _pgrid(x)
    plan = plan or get_plan(mesh_shape, x.device)
    tv = {"x": 0, "v": 1}[target]
    if tv == 1:
        v = _f32(v, "v")
    plan.reserve(npx, npy, npz)
    check(load().pmx_pgd(plan.handle, ptr(x), ptr(v) if tv == 1 else None, npx, npy, npz, MODES[mode], float(alpha),
                         float(kl), float(ks), tv, float(v_scale), stream_ptr()), "pmx_pgd")
    return x, v


def power_sums(mesh: torch.Tensor, box, kedges: torch.Tensor):
    """Device sums (pk_sum, k_sum, k_cnt) over nedges+1 bins, float64."""
    mesh = _f32(mesh, "mesh")
    plan = get_plan(mesh.shape, mesh.device)
    plan.set_box(box)
    kedges = to_device(kedges, mesh.device)
    nb = kedges.numel() + 1
    pk = torch.empty(nb, dtype=torch.float64, device=mesh.device)
    ks = torch.empty_like(pk)
    kc = torch.empty_like(pk)
    check(load().pmx_power(plan.handle, ptr(mesh), ptr(kedges), kedges.numel(), ptr(pk), ptr(ks), ptr(kc),
                           stream_ptr()), "pmx_power")
    return pk, ks, kc


def power_multipole_sums(mesh: torch.Tensor, mesh2: torch.Tensor | None, box, kedges: torch.Tensor, poles, los):
    """Device sums (re, im, k_sum, k_cnt): re/im of shape (npoles, nedges+1) (im None for an auto
    spectrum), float64 (pmx_power_multipoles)."""
    mesh = _f32(mesh, "mesh")
    if mesh2 is not None:
        mesh2 = _f32(mesh2, "mesh2")
        if tuple(mesh2.shape) != tuple(mesh.shape):
            raise ValueError("mesh2 must have the mesh's shape")
    plan = get_plan(mesh.shape, mesh.device)
    plan.set_box(box)
    kedges = to_device(kedges, mesh.device)
    nb = kedges.numel() + 1
    poles = [int(l) for l in poles]
    re = torch.empty((len(poles), nb), dtype=torch.float64, device=mesh.device)
    im = torch.empty_like(re) if mesh2 is not None else None
    ks = torch.empty(nb, dtype=torch.float64, device=mesh.device)
    kc = torch.empty_like(ks)
    cp = (ctypes.c_int * len(poles))(*poles)
    cl = (ctypes.c_double * 3)(*[float(v) for v in los]) if los is not None else None
    check(load().pmx_power_multipoles(plan.handle, ptr(mesh), ptr(mesh2) if mesh2 is not None else None, ptr(kedges),
                                      kedges.numel(), cp, len(poles), cl, ptr(re),
                                      ptr(im) if im is not None else None, ptr(ks), ptr(kc), stream_ptr()),
          "pmx_power_multipoles")
    return re, im, ks, kc


def white_noise(mesh_shape, seed: int, device=None, out: torch.Tensor | None = None, plan=None) -> torch.Tensor:
    plan = plan or get_plan(mesh_shape, device)
    if out is None:
        out = torch.empty(tuple(mesh_shape), dtype=torch.float32, device=plan.device)
    check(load().pmx_white_noise(plan.handle, ctypes.c_uint64(int(seed) & (2**64 - 1)), ptr(out), stream_ptr()),
          "pmx_white_noise")
    return out


def philox_raw(seed: int, nblocks: int, device=None) -> torch.Tensor:
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    out = torch.empty((nblocks, 4), dtype=torch.int32, device=dev)
    check(load().pmx_philox4x32(ctypes.c_uint64(int(seed) & (2**64 - 1)), nblocks, ptr(out), stream_ptr()),
          "pmx_philox4x32")
    return out


def color_field(field: torch.Tensor, box, ktab: torch.Tensor, ptab: torch.Tensor, out: torch.Tensor | None = None,
                plan=None) -> torch.Tensor:
    field = _f32(field, "field")
    plan = plan or get_plan(field.shape, field.device)
    plan.set_box(box)
    ktab = to_device(ktab, field.device)
    ptab = to_device(ptab, field.device)
    if out is None:
        out = torch.empty_like(field)
    check(load().pmx_color_field(plan.handle, ptr(field), ptr(ktab), ptr(ptab), ktab.numel(), ptr(out),
                                 stream_ptr()), "pmx_color_field")
    return out


def lpt1(ic: torch.Tensor, growth: float, vel_factor: float, out=None, plan=None, particles=None):
    """dx = growth F(ic), p = vel_factor dx at the particle lattice; `particles` = (npx, npy, npz), a grid
    dividing the mesh (default: one particle per cell)."""
    ic = _f32(ic, "ic")
    plan = plan or get_plan(ic.shape, ic.device)
    npx, npy, npz = tuple(ic.shape) if particles is None else tuple(int(n) for n in particles)
    if out is None:
        dx = torch.empty((npx, npy, npz, 3), dtype=torch.float32, device=ic.device)
        pv = torch.empty_like(dx)
    else:
        dx, pv = out
    check(load().pmx_lpt1_lattice(plan.handle, ptr(ic), npx, npy, npz, float(growth), float(vel_factor), ptr(dx),
                                  ptr(pv), stream_ptr()), "pmx_lpt1_lattice")
    return dx, pv


def lpt(ic: torch.Tensor, order: int, coef, particles=None, r_max: float = 0.0, box=None, observer=None, plan=None):
    """LPT of order 1 or 2 at the particle lattice (pmx_lpt): coef is (4,) = (D1, c1, D2, c2) constants or
    (4, ntab) columns over the comoving distance grid [0, r_max] of a lightcone (box, observer in Mpc/h)."""
    ic = _f32(ic, "ic")
    plan = plan or get_plan(ic.shape, ic.device)
    npx, npy, npz = tuple(ic.shape) if particles is None else tuple(int(n) for n in particles)
    coef = np.asarray(coef, dtype=np.float64)
    ntab = 1 if coef.ndim == 1 else coef.shape[1]
    if coef.shape[0] != 4:
        raise ValueError("coef must hold the four columns D1, c1, D2, c2")
    cdev = torch.from_numpy(np.ascontiguousarray(coef.reshape(4, ntab))).to(ic.device)
    dx = torch.empty((npx, npy, npz, 3), dtype=torch.float32, device=ic.device)
    pv = torch.empty_like(dx)
    check(load().pmx_lpt(plan.handle, ptr(ic), npx, npy, npz, int(order), ptr(cdev), ntab, float(r_max),
                         dbl3(box) if box is not None else None, dbl3(observer) if observer is not None else None,
                         ptr(dx), ptr(pv), stream_ptr()), "pmx_lpt")
    return dx, pv


def lpt1_vjp(dx_bar, p_bar, mesh_shape, growth: float, vel_factor: float) -> torch.Tensor:
    """Cotangent of the initial field for lpt1 given (dx_bar, p_bar) (either may be None)."""
    ref = dx_bar if dx_bar is not None else p_bar
    ref = _f32(ref, "dx_bar")
    npx, npy, npz = _pgrid(ref)
    dx_bar = None if dx_bar is None else _f32(dx_bar, "dx_bar")
    p_bar = None if p_bar is None else _f32(p_bar, "p_bar")
    plan = get_plan(mesh_shape, ref.device)
    out = torch.empty(tuple(mesh_shape), dtype=torch.float32, device=ref.device)
    check(load().pmx_lpt1_vjp(plan.handle, ptr(dx_bar), ptr(p_bar), npx, npy, npz, float(growth), float(vel_factor),
                              ptr(out), stream_ptr()), "pmx_lpt1_vjp")
    return out


_SCHEMES = {"ngp": 0, "bilinear": 1, "rbf_neighbor": 2}


def _scheme(scheme):
    sch = _SCHEMES.get(scheme)
    if sch is None:
        raise NotImplementedError(f"spherical painting scheme {scheme!r} is not supported (ngp, bilinear, rbf_neighbor)")
    return sch


def _tiles(tiles, device):
    """(ntiles, rot ptr, shift ptr, keep-alive) for an image table (rot codes, shifts in Mpc/h) or None."""
    if tiles is None:
        return 1, None, None, None
    rot, shift = tiles
    rot = to_device(torch.as_tensor(rot, dtype=torch.int32).reshape(-1), device, torch.int32)
    shift = to_device(torch.as_tensor(shift, dtype=torch.float64).reshape(-1, 3), device)
    if rot.numel() != shift.shape[0] or rot.numel() < 1:
        raise ValueError("image table: one rotation code and one shift per image")
    return rot.numel(), rot, shift, (rot, shift)


_ORDERS = {"RING": 0, "NESTED": 1}


def healpix_ud_grade(m: torch.Tensor, nside_out: int, power: float = 0.0, order_in: str = "RING",
                     order_out: str = "RING", pess: bool = False) -> torch.Tensor:
    """healpy.ud_grade of a float32 HEALPix map on the device (pmx_healpix_ud_grade)."""
    m = _f32(m, "map")
    nside_in = int(round((m.numel() / 12) ** 0.5))
    if 12 * nside_in * nside_in != m.numel():
        raise ValueError("not a HEALPix map")
    oi, oo = _ORDERS.get(str(order_in).upper()), _ORDERS.get(str(order_out).upper())
    if oi is None or oo is None:
        raise ValueError("ud_grade orders are 'RING' or 'NESTED'")
    out = torch.empty(12 * int(nside_out) ** 2, dtype=torch.float32, device=m.device)
    check(load().pmx_healpix_ud_grade(ptr(m), nside_in, ptr(out), int(nside_out), float(power or 0.0), oi, oo,
                                      int(bool(pess)), stream_ptr()), "pmx_healpix_ud_grade")
    return out


def rbf_sigma(kernel_width_arcmin=None, smoothing_interpretation="fwhm") -> float:
    """Gaussian width (radians) of the rbf_neighbor paint: kernel_width_arcmin read as a FWHM or a sigma;
    0 (no width given) selects the pixel resolution as FWHM."""
    if kernel_width_arcmin is None:
        return 0.0
    w = float(kernel_width_arcmin) * math.pi / (180.0 * 60.0)
    if smoothing_interpretation == "fwhm":
        return w / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    if smoothing_interpretation == "sigma":
        return w
    raise ValueError(f"smoothing_interpretation must be 'fwhm' or 'sigma', got {smoothing_interpretation!r}")


def paint_healpix(pos: torch.Tensor, mesh_shape, box, observer, rmin, rmax, nside, scheme="bilinear",
                  mode="relative", weights=1.0, out=None, accumulate=False, tiles=None, sigma=0.0) -> torch.Tensor:
    """HEALPix shell map; `tiles` = (rotation codes, shifts (n, 3) Mpc/h) sums the box's images; `sigma`:
    the rbf_neighbor width in radians (rbf_sigma)."""
    pos = _f32(pos, "pos")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    check(load().pmx_plan_set_rbf_width(plan.handle, float(sigma)), "pmx_plan_set_rbf_width")
    if out is None:
        out = torch.empty(12 * nside * nside, dtype=torch.float32, device=pos.device)
    w, w0 = _weights(weights, npx * npy * npz)
    nt, rot, shift, _keep = _tiles(tiles, pos.device)
    plan.reserve(acc=12 * int(nside) ** 2 + 4 * int(nside))
    check(load().pmx_paint_healpix_tiled(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], ptr(w), w0, dbl3(box),
                                         dbl3(observer), float(rmin), float(rmax), int(nside), _scheme(scheme), nt,
                                         ptr(rot), ptr(shift), ptr(out), int(accumulate), stream_ptr()),
          "pmx_paint_healpix_tiled")
    return out


def paint_healpix_drift(pos: torch.Tensor, vel: torch.Tensor, mesh_shape, box, observer, rmin, rmax, nside, drift,
                        scheme="bilinear", mode="relative", weights=1.0, out=None, accumulate=False, tiles=None,
                        sigma=0.0):
    """paint_healpix with every (image) position drifted to its own lightcone time first. drift: dict with
    "tables" (device float64 (3, ntab): a(chi), Gp(a), dGfa(a)), "chi_max", "a_lo", "a_hi", "a_now",
    "gp_now", "bound"."""
    pos = _f32(pos, "pos")
    vel = _f32(vel, "vel")
    npx, npy, npz = _pgrid(pos)
    if tuple(vel.shape) != tuple(pos.shape):
        raise ValueError("vel must have the shape of pos")
    plan = get_plan(mesh_shape, pos.device)
    check(load().pmx_plan_set_rbf_width(plan.handle, float(sigma)), "pmx_plan_set_rbf_width")
    if out is None:
        out = torch.empty(12 * nside * nside, dtype=torch.float32, device=pos.device)
    w, w0 = _weights(weights, npx * npy * npz)
    nt, rot, shift, _keep = _tiles(tiles, pos.device)
    tab = drift["tables"]
    plan.reserve(acc=12 * int(nside) ** 2 + 4 * int(nside))
    check(load().pmx_paint_healpix_lightcone(
        plan.handle, ptr(pos), ptr(vel), npx, npy, npz, MODES[mode], ptr(w), w0, dbl3(box), dbl3(observer),
        float(rmin), float(rmax), int(nside), _scheme(scheme), nt, ptr(rot), ptr(shift), ptr(tab),
        int(tab.shape[-1]), float(drift["chi_max"]), float(drift["a_lo"]), float(drift["a_hi"]),
        float(drift["a_now"]), float(drift["gp_now"]), float(drift["bound"]), ptr(out), int(accumulate),
        stream_ptr()), "pmx_paint_healpix_lightcone")
    return out


def paint_healpix_drift_vjp(pos: torch.Tensor, vel: torch.Tensor, map_bar: torch.Tensor, mesh_shape, box, observer,
                            rmin, rmax, nside, drift, scheme="bilinear", mode="relative", weights=1.0, tiles=None,
                            pos_bar=None, vel_bar=None, accumulate=False, param_tables=None, param_dots=None):
    """(pos_bar, vel_bar) (+)= d<map_bar, paint_healpix_drift(pos, vel)>/d(pos, vel). param_tables
    (float64 device (nparam, ntab): df/dtheta on the drift's a(chi) grid) and param_dots (float64 device
    (nparam,), added to): the cotangents of up to 4 parameters through the drift factor."""
    pos = _f32(pos, "pos")
    vel = _f32(vel, "vel")
    map_bar = _f32(map_bar, "map_bar")
    if map_bar.numel() != 12 * nside * nside:
        raise ValueError("map_bar must have 12 nside^2 pixels")
    if tuple(vel.shape) != tuple(pos.shape):
        raise ValueError("vel must have the shape of pos")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    if pos_bar is None or vel_bar is None:
        if accumulate:
            raise ValueError("accumulate needs both pos_bar and vel_bar")
        pos_bar = torch.empty_like(pos) if pos_bar is None else pos_bar
        vel_bar = torch.empty_like(vel) if vel_bar is None else vel_bar
    w, w0 = _weights(weights, npx * npy * npz)
    nt, rot, shift, _keep = _tiles(tiles, pos.device)
    tab = drift["tables"]
    plan.reserve(acc=12 * int(nside) ** 2 + 4 * int(nside))
    check(load().pmx_paint_healpix_lightcone_vjp(
        plan.handle, ptr(pos), ptr(vel), npx, npy, npz, MODES[mode], ptr(w), w0, dbl3(box), dbl3(observer),
        float(rmin), float(rmax), int(nside), _scheme(scheme), nt, ptr(rot), ptr(shift), ptr(tab), int(tab.shape[-1]),
        float(drift["chi_max"]), float(drift["a_lo"]), float(drift["a_hi"]), float(drift["a_now"]),
        float(drift["gp_now"]), float(drift["bound"]), ptr(map_bar), ptr(pos_bar), ptr(vel_bar), int(accumulate),
        0 if param_tables is None else int(param_tables.shape[0]), ptr(param_tables), ptr(param_dots),
        stream_ptr()), "pmx_paint_healpix_lightcone_vjp")
    return pos_bar, vel_bar


def paint_healpix_vjp(pos: torch.Tensor, map_bar: torch.Tensor, mesh_shape, box, observer, rmin, rmax, nside,
                      scheme="bilinear", mode="relative", weights=1.0, tiles=None, pos_bar=None,
                      accumulate=False) -> torch.Tensor:
    """pos_bar (+)= d<map_bar, paint_healpix(pos)>/d pos."""
    pos = _f32(pos, "pos")
    map_bar = _f32(map_bar, "map_bar")
    if map_bar.numel() != 12 * nside * nside:
        raise ValueError("map_bar must have 12 nside^2 pixels")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    if pos_bar is None:
        pos_bar = torch.empty_like(pos)
        accumulate = False
    w, w0 = _weights(weights, npx * npy * npz)
    nt, rot, shift, _keep = _tiles(tiles, pos.device)
    plan.reserve(acc=12 * int(nside) ** 2 + 4 * int(nside))
    check(load().pmx_paint_healpix_vjp(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], ptr(w), w0, dbl3(box),
                                       dbl3(observer), float(rmin), float(rmax), int(nside), _scheme(scheme), nt,
                                       ptr(rot), ptr(shift), ptr(map_bar), ptr(pos_bar), int(accumulate),
                                       stream_ptr()), "pmx_paint_healpix_vjp")
    return pos_bar


def _born_args(maps, coef):
    maps = _f32(maps, "maps")
    if maps.ndim != 2:
        raise ValueError("maps must be (nshell, npix)")
    coef = to_device(coef, maps.device)
    if coef.ndim != 2 or coef.shape[1] != maps.shape[0]:
        raise ValueError("coef must be (nsrc, nshell)")
    return maps, coef


def born(maps: torch.Tensor, coef, out=None, accumulate=False) -> torch.Tensor:
    """kappa (nsrc, npix) = sum_s coef[k, s] (maps[s] / mean(maps[s]) - 1)."""
    maps, coef = _born_args(maps, coef)
    S, npix = int(maps.shape[0]), int(maps.shape[1])
    K = int(coef.shape[0])
    plan = get_plan((8, 8, 8), maps.device)
    if out is None:
        out = torch.empty((K, npix), dtype=torch.float32, device=maps.device)
    check(load().pmx_born(plan.handle, ptr(maps), S, npix, ptr(coef), K, ptr(out), int(accumulate), stream_ptr()),
          "pmx_born")
    return out


def born_vjp(maps: torch.Tensor, coef, kappa_bar: torch.Tensor, want_dots=False):
    """(maps_bar, dots | None): cotangent of born for kappa_bar; dots = (D (nsrc, nshell), E (nsrc,),
    mean (nshell,)) with D = <kappa_bar[k], maps[s]>, E = sum kappa_bar[k], the shell means as used
    (float64, device)."""
    maps, coef = _born_args(maps, coef)
    kappa_bar = _f32(kappa_bar, "kappa_bar")
    S, npix = int(maps.shape[0]), int(maps.shape[1])
    K = int(coef.shape[0])
    if tuple(kappa_bar.shape) != (K, npix):
        raise ValueError("kappa_bar must be (nsrc, npix)")
    plan = get_plan((8, 8, 8), maps.device)
    maps_bar = torch.empty_like(maps)
    dots = torch.empty(K * S + K + S, dtype=torch.float64, device=maps.device) if want_dots else None
    check(load().pmx_born_vjp(plan.handle, ptr(maps), S, npix, ptr(coef), K, ptr(kappa_bar), ptr(maps_bar), 0,
                              ptr(dots), stream_ptr()), "pmx_born_vjp")
    if dots is None:
        return maps_bar, None
    return maps_bar, (dots[:K * S].view(K, S), dots[K * S:K * S + K], dots[K * S + K:])


def _born_flat_args(maps, coef, scales):
    maps = _f32(maps, "maps")
    if maps.ndim != 3:
        raise ValueError("flat maps must be (nshell, ny, nx)")
    coef = to_device(coef, maps.device)
    if coef.ndim != 2 or coef.shape[1] != maps.shape[0]:
        raise ValueError("coef must be (nsrc, nshell)")
    scales = to_device(scales, maps.device)
    if tuple(scales.shape) != (int(maps.shape[0]), 2):
        raise ValueError("scales must be (nshell, 2): row and column coordinate per pixel index")
    return maps, coef, scales


def born_flat(maps: torch.Tensor, coef, scales, out=None, accumulate=False) -> torch.Tensor:
    """kappa (nsrc, ny, nx) = sum_s coef[k, s] (R_s / mean(maps[s]) - 1), R_s the plane resampled at
    (row j scales[s, 0] - 1/2, column i scales[s, 1] - 1/2) for output pixel (i, j)."""
    maps, coef, scales = _born_flat_args(maps, coef, scales)
    S, ny, nx = (int(n) for n in maps.shape)
    K = int(coef.shape[0])
    plan = get_plan((8, 8, 8), maps.device)
    if out is None:
        out = torch.empty((K, ny, nx), dtype=torch.float32, device=maps.device)
    check(load().pmx_born_flat(plan.handle, ptr(maps), S, ny, nx, ptr(coef), K, ptr(scales), ptr(out),
                               int(accumulate), stream_ptr()), "pmx_born_flat")
    return out


def born_flat_vjp(maps: torch.Tensor, coef, scales, kappa_bar: torch.Tensor, want_dots=False):
    """(maps_bar, dots | None) for born_flat; dots = (D (nsrc, nshell), E (nsrc,), mean (nshell,),
    P (nshell,)) with D = <kappa_bar[k], R_s>, P[s] = chi_s d<kappa_bar, kappa>/d chi_s through the
    sampling positions (float64, device)."""
    maps, coef, scales = _born_flat_args(maps, coef, scales)
    kappa_bar = _f32(kappa_bar, "kappa_bar")
    S, ny, nx = (int(n) for n in maps.shape)
    K = int(coef.shape[0])
    if tuple(kappa_bar.shape) != (K, ny, nx):
        raise ValueError("kappa_bar must be (nsrc, ny, nx)")
    plan = get_plan((8, 8, 8), maps.device)
    maps_bar = torch.empty_like(maps)
    dots = torch.empty(K * S + K + 2 * S, dtype=torch.float64, device=maps.device) if want_dots else None
    check(load().pmx_born_flat_vjp(plan.handle, ptr(maps), S, ny, nx, ptr(coef), K, ptr(scales), ptr(kappa_bar),
                                   ptr(maps_bar), 0, ptr(dots), stream_ptr()), "pmx_born_flat_vjp")
    if dots is None:
        return maps_bar, None
    return maps_bar, (dots[:K * S].view(K, S), dots[K * S:K * S + K], dots[K * S + K:K * S + K + S],
                      dots[K * S + K + S:])


def paint_flat(pos: torch.Tensor, mesh_shape, flatsky_npix, zlo, zhi, scale, mode="absolute", weights=1.0,
               out=None, accumulate=False) -> torch.Tensor:
    pos = _f32(pos, "pos")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    fnx, fny = int(flatsky_npix[0]), int(flatsky_npix[1])
    if out is None:
        out = torch.empty((fnx, fny), dtype=torch.float32, device=pos.device)
    w, w0 = _weights(weights, npx * npy * npz)
    plan.reserve(acc=fnx * fny)
    check(load().pmx_paint_flat(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], ptr(w), w0, fnx, fny,
                                float(zlo), float(zhi), float(scale), ptr(out), int(accumulate), stream_ptr()),
          "pmx_paint_flat")
    return out


def paint_flat_vjp(pos: torch.Tensor, map_bar: torch.Tensor, mesh_shape, zlo, zhi, scale, mode="absolute",
                   weights=1.0, pos_bar: torch.Tensor | None = None, accumulate=False) -> torch.Tensor:
    """pos_bar (+)= d<map_bar, paint_flat(pos, ...)>/d pos (x and y; z gets nothing)."""
    pos = _f32(pos, "pos")
    map_bar = _f32(map_bar, "map_bar")
    if map_bar.dim() != 2:
        raise ValueError("map_bar must be a (fnx, fny) map")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    fnx, fny = int(map_bar.shape[0]), int(map_bar.shape[1])
    if pos_bar is None:
        pos_bar = torch.empty_like(pos)
        accumulate = False
    w, w0 = _weights(weights, npx * npy * npz)
    plan.reserve(acc=fnx * fny)
    check(load().pmx_paint_flat_vjp(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], ptr(w), w0, fnx, fny,
                                    float(zlo), float(zhi), float(scale), ptr(map_bar), ptr(pos_bar), int(accumulate),
                                    stream_ptr()), "pmx_paint_flat_vjp")
    return pos_bar


def image_positions(pos: torch.Tensor, mesh_shape, box, observer, rot=0, shift=(0.0, 0.0, 0.0), mode="relative",
                    vel=None, drift=None, out=None) -> torch.Tensor:
    """Absolute grid positions (npx, npy, npz, 3) of one image of the particles: rotated by the cube
    symmetry `rot` about the observer, shifted by `shift` (Mpc/h) and, with vel and drift (the tables
    of paint_healpix_drift), drifted to each image's own lightcone time. Positions are not wrapped."""
    pos = _f32(pos, "pos")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    if out is None:
        out = torch.empty_like(pos)
    if vel is not None:
        vel = _f32(vel, "vel")
        if tuple(vel.shape) != tuple(pos.shape):
            raise ValueError("vel must have the shape of pos")
        if drift is None:
            raise ValueError("a drift needs its tables")
        tab = drift["tables"]
        targs = (ptr(tab), int(tab.shape[-1]), float(drift["chi_max"]), float(drift["a_lo"]), float(drift["a_hi"]),
                 float(drift["a_now"]), float(drift["gp_now"]))
    else:
        targs = (None, 0, 0.0, 0.0, 0.0, 0.0, 0.0)
    check(load().pmx_image_positions(plan.handle, ptr(pos), ptr(vel), npx, npy, npz, MODES[mode], dbl3(box),
                                     dbl3(observer), int(rot), dbl3(shift), *targs, ptr(out), stream_ptr()),
          "pmx_image_positions")
    return out


def image_positions_vjp(pos: torch.Tensor, out_bar: torch.Tensor, mesh_shape, box, observer, rot=0,
                        shift=(0.0, 0.0, 0.0), mode="relative", vel=None, drift=None, param_tables=None,
                        param_dots=None):
    """(pos_bar, vel_bar | None): cotangents of image_positions for out_bar; param_tables (float64 device
    (nparam, ntab): df/dtheta on the drift's a(chi) grid) and param_dots (float64 device, added to) give
    up to 4 parameter cotangents through the drift factor."""
    pos = _f32(pos, "pos")
    out_bar = _f32(out_bar, "out_bar")
    if out_bar.shape != pos.shape:
        raise ValueError("out_bar must have the shape of pos")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    pos_bar = torch.empty_like(pos)
    vel_bar = None
    nparam = 0
    if vel is not None:
        vel = _f32(vel, "vel")
        if drift is None:
            raise ValueError("a drift needs its tables")
        vel_bar = torch.empty_like(vel)
        tab = drift["tables"]
        targs = (ptr(tab), int(tab.shape[-1]), float(drift["chi_max"]), float(drift["a_lo"]), float(drift["a_hi"]),
                 float(drift["a_now"]), float(drift["gp_now"]))
        if param_tables is not None:
            nparam = int(param_tables.shape[0])
    else:
        targs = (None, 0, 0.0, 0.0, 0.0, 0.0, 0.0)
    check(load().pmx_image_positions_vjp(plan.handle, ptr(pos), ptr(vel), npx, npy, npz, MODES[mode], dbl3(box),
                                         dbl3(observer), int(rot), dbl3(shift), *targs, ptr(out_bar), ptr(pos_bar),
                                         ptr(vel_bar), 0, nparam, ptr(param_tables) if nparam else None,
                                         ptr(param_dots) if nparam else None, stream_ptr()), "pmx_image_positions_vjp")
    return pos_bar, vel_bar


# ---------------------------------------------------------------------------- adjoints
def _zeros_like_or(t, like):
    return torch.zeros_like(like) if t is None else t


def paint_vjp(pos: torch.Tensor, mesh_bar: torch.Tensor, mode="relative", weights=1.0, want_weights=False):
    """(pos_bar, weights_bar | None): cotangents of paint(pos, weights) for the mesh cotangent mesh_bar."""
    pos = _f32(pos, "pos")
    mesh_bar = _f32(mesh_bar, "mesh_bar")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(tuple(mesh_bar.shape), pos.device)
    w, w0 = _weights(weights, npx * npy * npz)
    pos_bar = torch.empty_like(pos)
    wbar = torch.empty((npx, npy, npz), dtype=torch.float32, device=pos.device) if want_weights else None
    plan.reserve(npx, npy, npz)
    check(load().pmx_paint_cic_vjp(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], ptr(w), w0, ptr(mesh_bar),
                                   ptr(pos_bar), ptr(wbar), 0, stream_ptr()), "pmx_paint_cic_vjp")
    return pos_bar, wbar


def readout_vjp(mesh: torch.Tensor, pos: torch.Tensor, out_bar: torch.Tensor, mode="relative", want_mesh=True,
                want_pos=True):
    """(mesh_bar | None, pos_bar | None): cotangents of readout(mesh, pos) for out_bar."""
    mesh = _f32(mesh, "mesh")
    pos = _f32(pos, "pos")
    out_bar = _f32(out_bar, "out_bar")
    npx, npy, npz = _pgrid(pos)
    ncomp = 1 if mesh.ndim == 3 else int(mesh.shape[3])
    plan = get_plan(tuple(mesh.shape[:3]), pos.device)
    mesh_bar = torch.empty_like(mesh) if want_mesh else None
    pos_bar = torch.empty_like(pos) if want_pos else None
    plan.reserve(npx, npy, npz)
    check(load().pmx_readout_cic_vjp(plan.handle, ptr(mesh), ncomp, ptr(pos), npx, npy, npz, MODES[mode],
                                     ptr(out_bar), ptr(mesh_bar), ptr(pos_bar), 0, stream_ptr()),
          "pmx_readout_cic_vjp")
    return mesh_bar, pos_bar


def force_mesh_vjp(fmesh_bar: torch.Tensor, scale=1.0, deconv=False) -> torch.Tensor:
    """Density cotangent of force_mesh(delta, scale) for the force-mesh cotangent (nx, ny, nz, 3)."""
    fmesh_bar = _f32(fmesh_bar, "fmesh_bar")
    plan = get_plan(tuple(fmesh_bar.shape[:3]), fmesh_bar.device)
    out = torch.empty(tuple(fmesh_bar.shape[:3]), dtype=torch.float32, device=fmesh_bar.device)
    check(load().pmx_force_mesh_vjp(plan.handle, ptr(fmesh_bar), float(scale), _lib.DECONV if deconv else 0, ptr(out),
                                    stream_ptr()), "pmx_force_mesh_vjp")
    return out


def forces_vjp(pos: torch.Tensor, mesh_shape, forces_bar: torch.Tensor, mode="relative", scale=1.0, deconv=False,
               pos_bar: torch.Tensor | None = None, accumulate=False, dot: torch.Tensor | None = None,
               force_mesh: torch.Tensor | None = None, bar_scale=1.0) -> torch.Tensor:
    """pos_bar (+)= J(pos)^T (bar_scale forces_bar) for forces(pos, scale); dot (float64 device scalar) +=
    forces_bar . F. force_mesh: the force mesh at pos, kept by step(..., force_mesh=) (not recomputed)."""
    pos = _f32(pos, "pos")
    forces_bar = _f32(forces_bar, "forces_bar")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    if pos_bar is None:
        pos_bar = torch.empty_like(pos)
        accumulate = False
    if force_mesh is not None:
        _force_mesh_arg(force_mesh, mesh_shape)
        plan.reserve(npx, npy, npz)
        check(load().pmx_forces_vjp_saved(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], float(scale),
                                          _lib.DECONV if deconv else 0, ptr(force_mesh), ptr(forces_bar),
                                          float(bar_scale), ptr(pos_bar), int(accumulate), ptr(dot), stream_ptr()),
              "pmx_forces_vjp_saved")
        return pos_bar
    plan.reserve(npx, npy, npz)
    check(load().pmx_forces_vjp(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], float(scale),
                                _lib.DECONV if deconv else 0, ptr(forces_bar), float(bar_scale), ptr(pos_bar),
                                int(accumulate), ptr(dot), stream_ptr()), "pmx_forces_vjp")
    return pos_bar


def axpy(x: torch.Tensor, v: torch.Tensor, c: float, out: torch.Tensor | None = None) -> torch.Tensor:
    """out = x + c v (float32, same shapes; out may be x)."""
    x = _f32(x, "x")
    v = _f32(v, "v")
    if v.shape != x.shape:
        raise ValueError("x and v must have the same shape")
    if out is None:
        out = torch.empty_like(x)
    check(load().pmx_axpy(ptr(out), ptr(x), ptr(v), float(c), x.numel(), stream_ptr()), "pmx_axpy")
    return out


def adjoint_drift(xbar: torch.Tensor, vbar: torch.Tensor, v: torch.Tensor, c: float, dot: torch.Tensor,
                  mesh_shape) -> None:
    """Transpose of the drift x <- x + c v: vbar += c xbar (in place) and dot (float64 device, added to)
    += <xbar, v>. mesh_shape names the plan whose reduction buffer the dot uses."""
    if not vbar.is_contiguous():
        raise ValueError("vbar is updated in place and must be contiguous")
    xbar = _f32(xbar, "xbar")
    vbar = _f32(vbar, "vbar")
    v = _f32(v, "v")
    if not (xbar.shape == vbar.shape == v.shape):
        raise ValueError("xbar, vbar and v must have the same shape")
    if dot.dtype != torch.float64 or dot.numel() < 1 or not dot.is_contiguous():
        raise ValueError("dot must be a contiguous float64 device tensor")
    plan = get_plan(mesh_shape, xbar.device)
    check(load().pmx_adjoint_drift(plan.handle, ptr(xbar), ptr(vbar), ptr(v), xbar.numel(), float(c), ptr(dot),
                                   stream_ptr()), "pmx_adjoint_drift")


def pgd_vjp(pos: torch.Tensor, mesh_shape, disp_bar: torch.Tensor, mode="relative", alpha=0.2, kl=0.3, ks=1.0,
            pos_bar: torch.Tensor | None = None, accumulate=False, dot: torch.Tensor | None = None) -> torch.Tensor:
    """pos_bar (+)= J_S(pos)^T disp_bar for S = alpha * PGD force (pgd); dot (float64 device) += disp_bar . S."""
    pos = _f32(pos, "pos")
    disp_bar = _f32(disp_bar, "disp_bar")
    npx, npy, npz = _pgrid(pos)
    plan = get_plan(mesh_shape, pos.device)
    if pos_bar is None:
        pos_bar = torch.empty_like(pos)
        accumulate = False
    plan.reserve(npx, npy, npz)
    check(load().pmx_pgd_vjp(plan.handle, ptr(pos), npx, npy, npz, MODES[mode], float(alpha), float(kl), float(ks),
                             ptr(disp_bar), ptr(pos_bar), int(accumulate), ptr(dot), stream_ptr()), "pmx_pgd_vjp")
    return pos_bar


def step_reverse_vjp(x, v, x_bar, v_bar, mesh_shape, mode="relative", force_scale=1.0, c_kick=0.0, c_drift=0.0,
                     deconv=False, coef_bar: torch.Tensor | None = None,
                     force_mesh: torch.Tensor | None = None, sharpen=None) -> torch.Tensor:
    """In place: undo one reversible step (v += F ck; x += v cd) on (x, v) and pull (x_bar, v_bar) back
    through it. Returns coef_bar (float64 device [dL/dc_drift, dL/dc_kick]). force_mesh: that step's
    force mesh kept by step(..., force_mesh=) (not recomputed). sharpen = (alpha, kl, ks, v_scale): the
    step also boosted v += alpha F_pgd(x) v_scale before the drift (coef_bar[1] is then
    <v_bar, F ck + S v_scale>)."""
    for t, nm in ((x, "x"), (v, "v"), (x_bar, "x_bar"), (v_bar, "v_bar")):
        _f32(t, nm)
        if not t.is_contiguous():
            raise _lib.PmxError(f"{nm} must be contiguous")
    npx, npy, npz = _pgrid(x)
    plan = get_plan(mesh_shape, x.device)
    if coef_bar is None:
        coef_bar = torch.empty(2, dtype=torch.float64, device=x.device)
    if sharpen is not None:
        alpha, kl, ks, vs = (float(q) for q in sharpen)
        plan.reserve(npx, npy, npz)
        check(load().pmx_step_reverse_vjp_sharpen(plan.handle, ptr(x), ptr(v), npx, npy, npz, MODES[mode],
                                                  float(force_scale), float(c_kick), float(c_drift),
                                                  _lib.DECONV if deconv else 0, alpha, kl, ks, vs, ptr(x_bar),
                                                  ptr(v_bar), ptr(coef_bar), stream_ptr()),
              "pmx_step_reverse_vjp_sharpen")
        return coef_bar
    if force_mesh is not None:
        _force_mesh_arg(force_mesh, mesh_shape)
        plan.reserve(npx, npy, npz)
        check(load().pmx_step_reverse_vjp_saved(plan.handle, ptr(x), ptr(v), npx, npy, npz, MODES[mode],
                                                float(force_scale), float(c_kick), float(c_drift),
                                                _lib.DECONV if deconv else 0, ptr(force_mesh), ptr(x_bar), ptr(v_bar),
                                                ptr(coef_bar), stream_ptr()), "pmx_step_reverse_vjp_saved")
        return coef_bar
    plan.reserve(npx, npy, npz)
    check(load().pmx_step_reverse_vjp(plan.handle, ptr(x), ptr(v), npx, npy, npz, MODES[mode], float(force_scale),
                                      float(c_kick), float(c_drift), _lib.DECONV if deconv else 0, ptr(x_bar),
                                      ptr(v_bar), ptr(coef_bar), stream_ptr()), "pmx_step_reverse_vjp")
    return coef_bar
