"""Continuous-batching forward-simulation server on one GPU (Python side of pmx_server_* in libpmx).

A request is one forward simulation of the PM path: seeded Gaussian ICs (white noise coloured by the
EH98 P(k) of its own cosmology), first-order LPT, then the EfficientDriftDoubleKick time loop
(src/jax_fli/pm/solvers.py:118-225, integrate.py:55-119) to the final scale factor, returning the final
displacements and the painted density. The reference runs such simulations one at a time
(run_simulations, src/jax_fli/scripts/fli_simulate.py:512-585); the native server keeps many in flight:

* one size class per mesh size, each with a fixed number of slots sized from the HBM budget; a slot
  holds a request's particles and meshes in HBM from its ICs to its final paint, and a request is
  admitted (in arrival order per size) as soon as a slot of its size frees up;
* every step phase of all live slots of a class is one batched launch, replayed from a captured HIP
  graph; classes run on their own streams, so different sizes overlap on the device;
* the host preparation of a request (P(k) table and sigma8 integral, growth table, time grid, KDK
  coefficients) is native code on a pool of host threads, off the scheduling thread.

The kernels are deterministic and slots do not interact, so a request's result is the same bit for
bit alone and in any mix.
"""

from __future__ import annotations

import ctypes
import dataclasses
import math
import time

import numpy as np
import torch

from . import _lib
from . import cosmology as _c
from .pm.integrate import time_grid
from .pm.interp import NoInterp
from .pm.solvers import EfficientDriftDoubleKick, StepOp
from .painting import PaintingOptions

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, -1


@dataclasses.dataclass
class SimRequest:
    rid: int
    mesh: int                 # n: n^3 particles on an n^3 mesh
    seed: int
    Omega_c: float
    sigma8: float
    arrival: float = 0.0      # seconds after the start of serve()
    box: float | None = None  # Mpc/h, default 3.9 n
    a0: float = 0.1
    a1: float = 1.0
    dt0: float = 0.09
    Omega_b: float = 0.0486
    h: float = 0.6774
    n_s: float = 0.9667

    def box_size(self) -> float:
        return float(self.box if self.box is not None else 3.9 * self.mesh)

    def cosmology(self):
        return _c.Cosmology(Omega_c=self.Omega_c, Omega_b=self.Omega_b, h=self.h, n_s=self.n_s, sigma8=self.sigma8)


@dataclasses.dataclass
class SimResult:
    rid: int
    mesh: int
    displacement: torch.Tensor  # (n, n, n, 3) final GRID_RELATIVE displacements
    density: torch.Tensor       # (n, n, n) painted particle counts at a1
    arrival: float
    admitted: float
    finished: float
    request_id: int = 0         # the server's id (pmx_server_request_times)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def forward_ops(req: SimRequest):
    """(D, vel_factor, ktab, ptab, ops, force_scale): the host side of one request in the Python layer --
    LPT factors, the colouring's P(k) table (times N / V), and the KDK time loop as StepOps with each
    step's leading drift fused into the previous step's readout (as integrate() does). The native
    server prepares the same quantities (pmx_host_request_plan); tests compare the two."""
    cosmo = req.cosmology()
    n, box = req.mesh, req.box_size()
    k = np.logspace(-4, 1, 128)
    ptab = _c.linear_matter_power(cosmo, k) * (n**3 / box**3)
    a = float(np.float32(req.a0))
    D = float(_c.growth_factor(cosmo, a))
    vf = float(a**2 * _c.growth_rate(cosmo, a) * _c.E(cosmo, a))
    solver = EfficientDriftDoubleKick(interp_kernel=NoInterp(painting=PaintingOptions(target="density")))
    raw = [solver.init_op(req.a0, req.a0 + req.dt0, req.dt0, cosmo)]
    grid = [(ev[1], ev[2]) for ev in time_grid(req.a0, [req.a1], req.dt0, req.a1) if ev[0] == "step"]
    raw += solver.step_ops(grid, req.dt0, req.a0, cosmo)
    ops, pending = [], None
    for op in raw:
        op = StepOp(op.pre, op.kick, op.post)
        if pending is not None and pending.post is None and op.pre is not None:
            pending.post, op.pre = op.pre, None
        if pending is not None:
            ops.append(pending)
        pending = op
    ops.append(pending)
    return D, vf, k, ptab, ops, 1.5 * cosmo.Omega_m


class _Request(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_int), ("seed", ctypes.c_uint64), ("Omega_c", ctypes.c_double),
                ("Omega_b", ctypes.c_double), ("h", ctypes.c_double), ("n_s", ctypes.c_double),
                ("sigma8", ctypes.c_double), ("box", ctypes.c_double), ("a0", ctypes.c_double),
                ("a1", ctypes.c_double), ("dt0", ctypes.c_double)]


class _Stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("running", ctypes.c_uint64), ("queued", ctypes.c_uint64), ("peak_running", ctypes.c_int),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("wall_s", ctypes.c_double), ("throughput_req_s", ctypes.c_double),
                ("device_bytes", ctypes.c_size_t)]


class Server:
    """The native continuous-batching server (pmx_server_*). meshes: the size classes; slots: per-size
    slot counts (dict or one int; None: from hbm_budget, default 60% of the free device memory)."""

    def __init__(self, device=None, meshes=(64, 96, 128, 192), slots=None, hbm_budget: int | None = None,
                 prep_threads: int = 4):
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.meshes = tuple(int(m) for m in meshes)
        lib = _lib.load()
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * 0.6)
        if slots is None:
            sl = [0] * len(self.meshes)
        elif isinstance(slots, dict):
            sl = [int(slots.get(m, 0)) for m in self.meshes]
        else:
            sl = [int(slots)] * len(self.meshes)
        ms = (ctypes.c_int * len(self.meshes))(*self.meshes)
        ss = (ctypes.c_int * len(sl))(*sl)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.pmx_server_create(ctypes.byref(handle), ms, ss, len(self.meshes), int(hbm_budget),
                                             int(prep_threads)), "pmx_server_create")
        self._h = handle
        self._keep = {}

    # -- classes -----------------------------------------------------------------
    def slots(self, mesh: int) -> int:
        n = ctypes.c_int()
        _lib.check(_lib.load().pmx_server_class(self._h, int(mesh), ctypes.byref(n), None), "pmx_server_class")
        return int(n.value)

    def slot_bytes(self, mesh: int) -> int:
        b = ctypes.c_size_t()
        _lib.check(_lib.load().pmx_server_class(self._h, int(mesh), None, ctypes.byref(b)), "pmx_server_class")
        return int(b.value)

    # -- requests ----------------------------------------------------------------
    def submit(self, req: SimRequest) -> int:
        """Queue one request; the result buffers are allocated here and handed to the server."""
        n = int(req.mesh)
        disp = torch.empty((n, n, n, 3), dtype=torch.float32, device=self.device)
        dens = torch.empty((n, n, n), dtype=torch.float32, device=self.device)
        r = _Request(n, int(req.seed) & (2**64 - 1), float(req.Omega_c), float(req.Omega_b), float(req.h),
                     float(req.n_s), float(req.sigma8), req.box_size(), float(req.a0), float(req.a1), float(req.dt0))
        rid = ctypes.c_uint64()
        _lib.check(_lib.load().pmx_server_submit(self._h, ctypes.byref(r), disp.data_ptr(), dens.data_ptr(),
                                                 ctypes.byref(rid)), "pmx_server_submit")
        self._keep[int(rid.value)] = (req, disp, dens)
        return int(rid.value)

    def state(self, rid: int) -> int:
        st = ctypes.c_int()
        _lib.check(_lib.load().pmx_server_poll(self._h, ctypes.c_uint64(rid), ctypes.byref(st)), "pmx_server_poll")
        return int(st.value)

    def wait(self, rid: int, timeout_s: float = 0.0) -> None:
        _lib.check(_lib.load().pmx_server_wait(self._h, ctypes.c_uint64(rid), float(timeout_s)), "pmx_server_wait")

    def times(self, rid: int):
        t = (ctypes.c_double * 3)()
        _lib.check(_lib.load().pmx_server_request_times(self._h, ctypes.c_uint64(rid), t), "pmx_server_request_times")
        return tuple(t)

    def stats(self, since_id: int = 0) -> dict:
        s = _Stats()
        _lib.check(_lib.load().pmx_server_stats(self._h, ctypes.c_uint64(since_id), ctypes.byref(s)),
                   "pmx_server_stats")
        return {"requests": int(s.completed), "failed": int(s.failed), "wall_s": s.wall_s,
                "throughput_req_s": s.throughput_req_s, "p50_ms": s.p50_ms, "p99_ms": s.p99_ms,
                "mean_ms": s.mean_ms, "max_ms": s.max_ms, "peak_active": int(s.peak_running),
                "device_bytes": int(s.device_bytes)}

    def _result(self, rid: int) -> SimResult:
        req, disp, dens = self._keep.pop(rid)
        t_sub, t_adm, t_done = self.times(rid)
        return SimResult(req.rid, req.mesh, disp, dens, req.arrival, req.arrival + (t_adm - t_sub),
                         req.arrival + (t_done - t_sub), rid)

    def serve(self, requests, sequential: bool = False, clock=time.perf_counter):
        """Submit `requests` at their arrival times (seconds after the call) and collect the results:
        ({rid: SimResult}, stats over these requests). sequential: one request at a time (each submitted
        when the previous one is done), the no-batching baseline."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        ids, first = [], None
        t0 = clock()
        i = 0
        while i < len(pending):
            now = clock() - t0
            if sequential:
                rid = self.submit(pending[i])
                self.wait(rid)
                ids.append(rid)
                i += 1
            elif pending[i].arrival <= now:
                ids.append(self.submit(pending[i]))
                i += 1
            else:
                time.sleep(min(1e-3, pending[i].arrival - now))
            if first is None and ids:
                first = ids[0]
        for rid in ids:
            self.wait(rid)
        results = {}
        for rid in ids:
            res = self._result(rid)
            results[res.rid] = res
        stats = self.stats(first if first is not None else 0)
        return results, stats

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            with torch.cuda.device(self.device):
                _lib._lib.pmx_server_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_requests(count: int, rate_per_s: float, seed: int = 0, meshes=(64, 96, 128, 192)):
    """The serving workload: Poisson arrivals, mesh uniform over `meshes` (box 3.9 N Mpc/h), per-request
    seed, Omega_c ~ U(0.2, 0.35), sigma8 ~ U(0.7, 0.9), 1LPT + KDK with dt0 = 0.09 from a = 0.1 to 1."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, size=count)) if math.isfinite(rate_per_s) else np.zeros(count)
    return [SimRequest(rid=i, mesh=int(rng.choice(meshes)), seed=int(rng.integers(1, 2**31)),
                       Omega_c=float(rng.uniform(0.2, 0.35)), sigma8=float(rng.uniform(0.7, 0.9)),
                       arrival=float(t[i])) for i in range(count)]
