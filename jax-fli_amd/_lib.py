"""ctypes binding of libpmx.so (the C ABI declared in include/pmx.h).

The library is loaded from this package directory. There is no fallback: a missing or
unloadable library raises immediately, and every failing call raises ``PmxError`` with the
library's message.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmx.so")

# every symbol include/pmx.h declares, with its argument types
_c = ctypes
_vp = _c.c_void_p
_i = _c.c_int
_f = _c.c_float
_d = _c.c_double
_u = _c.c_uint
_dp = _c.POINTER(_c.c_double)
_sz = _c.c_size_t
SIGNATURES = {
    "pmx_abi_version": (_i, []),
    "pmx_last_error": (_c.c_char_p, []),
    "pmx_fft_supported": (_i, [_i]),
    "pmx_set_grid_cap": (_i, [_u]),
    "pmx_plan_reserve": (_i, [_vp, _i, _i, _i, _sz]),
    "pmx_plan_create": (_i, [_c.POINTER(_vp), _i, _i, _i, _u]),
    "pmx_plan_destroy": (_i, [_vp]),
    "pmx_plan_workspace_bytes": (_c.c_size_t, [_vp]),
    "pmx_plan_set_box": (_i, [_vp, _dp]),
    "pmx_plan_escapes": (_i, [_vp, _c.POINTER(_c.c_uint64)]),
    "pmx_plan_timing": (_i, [_vp, _i]),
    "pmx_plan_timing_read": (_i, [_vp, _dp, _c.POINTER(_i)]),
    "pmx_paint_cic": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _f, _vp, _i, _vp]),
    "pmx_readout_cic": (_i, [_vp, _vp, _i, _vp, _i, _i, _i, _i, _vp, _vp]),
    "pmx_fft_r2c": (_i, [_vp, _vp, _vp, _vp]),
    "pmx_fft_c2r": (_i, [_vp, _vp, _vp, _vp]),
    "pmx_force_mesh": (_i, [_vp, _vp, _f, _u, _vp, _vp]),
    "pmx_forces": (_i, [_vp, _vp, _i, _i, _i, _i, _f, _u, _vp, _vp]),
    "pmx_step": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _f, _f, _u, _vp]),
    "pmx_step_keep": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _f, _f, _u, _vp, _vp]),
    "pmx_power": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "pmx_plan_set_rbf_width": (_i, [_vp, _d]),
    "pmx_healpix_ud_grade": (_i, [_vp, _i, _vp, _i, _f, _i, _i, _i, _vp]),
    "pmx_lpt": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    "pmx_power_multipoles": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pmx_white_noise": (_i, [_vp, _c.c_uint64, _vp, _vp]),
    "pmx_philox4x32": (_i, [_c.c_uint64, _c.c_size_t, _vp, _vp]),
    "pmx_color_field": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "pmx_lpt1": (_i, [_vp, _vp, _f, _f, _vp, _vp, _vp]),
    "pmx_paint_healpix": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _f, _dp, _dp, _d, _d, _i, _i, _vp, _i, _vp]),
    "pmx_paint_flat": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _f, _i, _i, _f, _f, _d, _vp, _i, _vp]),
    "pmx_paint_flat_vjp": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _f, _i, _i, _f, _f, _d, _vp, _vp, _i, _vp]),
    "pmx_image_positions_vjp": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _dp, _dp, _i, _dp, _vp, _i, _d, _d, _d, _d, _d,
                                     _vp, _vp, _vp, _i, _i, _vp, _vp, _vp]),
    "pmx_image_positions": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _dp, _dp, _i, _dp, _vp, _i, _d, _d, _d, _d, _d, _vp,
                                 _vp]),
    "pmx_paint_cic_vjp": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _f, _vp, _vp, _vp, _i, _vp]),
    "pmx_readout_cic_vjp": (_i, [_vp, _vp, _i, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp]),
    "pmx_force_mesh_vjp": (_i, [_vp, _vp, _f, _u, _vp, _vp]),
    "pmx_forces_vjp": (_i, [_vp, _vp, _i, _i, _i, _i, _f, _u, _vp, _f, _vp, _i, _vp, _vp]),
    "pmx_step_reverse_vjp": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _f, _u, _vp, _vp, _vp, _vp]),
    "pmx_step_reverse_vjp_saved": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _f, _u, _vp, _vp, _vp, _vp, _vp]),
    "pmx_step_reverse_vjp_sharpen": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _f, _u, _f, _f, _f, _f, _vp, _vp, _vp,
                                          _vp]),
    "pmx_lpt1_lattice": (_i, [_vp, _vp, _i, _i, _i, _f, _f, _vp, _vp, _vp]),
    "pmx_lpt1_vjp": (_i, [_vp, _vp, _vp, _i, _i, _i, _f, _f, _vp, _vp]),
    "pmx_paint_healpix_tiled": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _f, _dp, _dp, _d, _d, _i, _i, _i, _vp, _vp, _vp,
                                     _i, _vp]),
    "pmx_paint_healpix_noalloc": (_i, [_i, _i, _i, _vp, _i, _i, _i, _i, _vp, _f, _dp, _dp, _d, _d, _i, _i, _i, _vp,
                                       _vp, _vp, _i, _vp]),
    "pmx_paint_healpix_vjp": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _f, _dp, _dp, _d, _d, _i, _i, _i, _vp, _vp, _vp,
                                   _vp, _i, _vp]),
    "pmx_born": (_i, [_vp, _vp, _i, _c.c_int64, _vp, _i, _vp, _i, _vp]),
    "pmx_forces_vjp_saved": (_i, [_vp, _vp, _i, _i, _i, _i, _f, _u, _vp, _vp, _f, _vp, _i, _vp, _vp]),
    "pmx_axpy": (_i, [_vp, _vp, _vp, _f, _sz, _vp]),
    "pmx_adjoint_drift": (_i, [_vp, _vp, _vp, _vp, _sz, _f, _vp, _vp]),
    "pmx_pgd_vjp": (_i, [_vp, _vp, _i, _i, _i, _i, _f, _f, _f, _vp, _vp, _i, _vp, _vp]),
    "pmx_pgd": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _f, _i, _f, _vp]),
    "pmx_paint_healpix_lightcone": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _vp, _f, _dp, _dp, _d, _d, _i, _i, _i, _vp,
                                         _vp, _vp, _i, _d, _d, _d, _d, _d, _d, _vp, _i, _vp]),
    "pmx_paint_healpix_lightcone_vjp": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _vp, _f, _dp, _dp, _d, _d, _i, _i, _i,
                                             _vp, _vp, _vp, _i, _d, _d, _d, _d, _d, _d, _vp, _vp, _vp, _i, _i, _vp,
                                             _vp, _vp]),
    "pmx_born_flat": (_i, [_vp, _vp, _i, _i, _i, _vp, _i, _vp, _vp, _i, _vp]),
    "pmx_born_flat_vjp": (_i, [_vp, _vp, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _vp, _vp]),
    "pmx_born_vjp": (_i, [_vp, _vp, _i, _c.c_int64, _vp, _i, _vp, _vp, _i, _vp, _vp]),
    "pmx_server_create": (_i, [_c.POINTER(_vp), _vp, _vp, _i, _sz, _i]),
    "pmx_server_class": (_i, [_vp, _i, _vp, _vp]),
    "pmx_server_submit": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "pmx_server_poll": (_i, [_vp, _c.c_uint64, _vp]),
    "pmx_server_wait": (_i, [_vp, _c.c_uint64, _d]),
    "pmx_server_request_times": (_i, [_vp, _c.c_uint64, _vp]),
    "pmx_server_stats": (_i, [_vp, _c.c_uint64, _vp]),
    "pmx_server_destroy": (_i, [_vp]),
}

DECONV = 1
DRIFT_PRE = 2
DRIFT_POST = 4
PAINT_ATOMIC = 8
PLAN_NO_RESERVE = 16


class PmxError(RuntimeError):
    """Error reported by libpmx."""


_lib = None
_lock = threading.Lock()


def load() -> ctypes.CDLL:
    """Load libpmx.so (once). Raises if it is missing: there is no non-native path."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise PmxError(f"{LIB_PATH} is missing; build it with __graft_entry__.build() (make -C csrc)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().pmx_last_error().decode(errors="replace")
        raise PmxError(f"{what} failed (code {rc}): {msg}")


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise PmxError("pmx kernels need device tensors")
    if not t.is_contiguous():
        raise PmxError("pmx kernels need contiguous tensors")
    return t.data_ptr()


def stream_ptr(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def dbl3(v) -> ctypes.Array:
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class Plan:
    """Owns one pmx plan (FFT tables and work buffers) for a mesh shape on one device."""

    def __init__(self, mesh_shape, device: torch.device | None = None, flags: int = 0):
        lib = load()
        self.mesh_shape = tuple(int(n) for n in mesh_shape)
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(lib.pmx_plan_create(ctypes.byref(handle), *self.mesh_shape, int(flags)), "pmx_plan_create")
        self.handle = handle
        self._box = None
        self._grids = set()
        self._acc = 0

    def reserve(self, npx: int = 0, npy: int = 0, npz: int = 0, acc: int = 0) -> None:
        """Size the paint-sort buffers for a particle grid and the map accumulator for `acc` values
        before the first call that needs them (synchronous, once per new grid or larger map): the
        entry points themselves never allocate."""
        grid = (int(npx), int(npy), int(npz)) if npx else None
        acc = int(acc) if acc > self._acc else 0
        if (grid is None or grid in self._grids) and not acc:
            return
        with torch.cuda.device(self.device):
            check(load().pmx_plan_reserve(self.handle, *(grid or (0, 0, 0)), acc), "pmx_plan_reserve")
        if grid is not None:
            self._grids.add(grid)
        if acc:
            self._acc = acc

    def escapes(self) -> int:
        """Particles the last sorted paint placed one by one (outside their sub-block's tile window;
        synchronises)."""
        c = ctypes.c_uint64()
        with torch.cuda.device(self.device):
            check(load().pmx_plan_escapes(self.handle, ctypes.byref(c)), "pmx_plan_escapes")
        return int(c.value)

    PHASES = ("paint", "r2c_z", "c2c_y", "kspace_x", "c2c_y_inv", "c2r_z", "grad4", "readout_kick_drift")

    def timing(self, enable: bool, steps: int = 64) -> None:
        """Start (events for `steps` steps created now) or stop per-phase timing of pmx_step."""
        with torch.cuda.device(self.device):
            check(load().pmx_plan_timing(self.handle, int(steps) if enable else 0), "pmx_plan_timing")

    def timing_read(self):
        """(phase -> summed ms, number of steps) of the pmx_step calls since timing(True)."""
        ms = (ctypes.c_double * len(self.PHASES))()
        n = ctypes.c_int()
        with torch.cuda.device(self.device):
            check(load().pmx_plan_timing_read(self.handle, ms, ctypes.byref(n)), "pmx_plan_timing_read")
        return dict(zip(self.PHASES, list(ms))), int(n.value)

    @property
    def workspace_bytes(self) -> int:
        return int(load().pmx_plan_workspace_bytes(self.handle))

    def set_box(self, box) -> None:
        box = tuple(float(b) for b in box)
        if box != self._box:
            with torch.cuda.device(self.device):
                check(load().pmx_plan_set_box(self.handle, dbl3(box)), "pmx_plan_set_box")
            self._box = box

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            try:
                with torch.cuda.device(self.device):
                    torch.cuda.synchronize(self.device)
                    _lib.pmx_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


_plans: dict = {}


def get_plan(mesh_shape, device: torch.device | None = None) -> Plan:
    """Cached plan per (device, mesh shape)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    key = (dev.index, tuple(int(n) for n in mesh_shape))
    p = _plans.get(key)
    if p is None:
        p = Plan(mesh_shape, dev, PLAN_NO_RESERVE)
        _plans[key] = p
    return p


def clear_plans() -> None:
    torch.cuda.synchronize()
    _plans.clear()


def release_plan(mesh_shape, device: torch.device | None = None) -> None:
    """Drop the cached plan of a mesh shape (its device buffers are freed once no caller holds it)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    _plans.pop((dev.index, tuple(int(n) for n in mesh_shape)), None)
