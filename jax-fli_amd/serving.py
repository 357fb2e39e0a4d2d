"""Continuous-batching server for forward simulations on one GPU.

A request is one forward simulation of the PM path: seeded Gaussian ICs (white noise coloured by the
EH98 P(k) of its own cosmology), first-order LPT, then the EfficientDriftDoubleKick time loop
(src/jax_fli/pm/solvers.py:118-225, integrate.py:55-119) to the final scale factor, returning the final
displacements and the painted density. The reference runs such simulations one at a time
(fli_simulate.py); this server keeps many in flight:

* admission by HBM budget: a request is admitted when a slot is free and its device memory (plan
  workspace + particle and mesh buffers, measured per mesh size) fits the budget; requests that
  arrive while others run wait in arrival order and are admitted as memory frees up;
* every admitted request owns a stream and a plan (pooled per mesh size), and the scheduler issues
  the requests' steps round-robin, at most `lookahead` steps ahead of the device per request, so the
  steps of different requests overlap on the GPU (small meshes fill the GPU only together) and the
  queued steps cover the host's preparation of newly admitted requests;
* all host work of a request (cosmology tables, time grid, kick/drift coefficients) is done once at
  admission; the device work of a step is one fused pmx_step call.

The kernels are deterministic (integer-accumulated paint, fixed-order FFTs and gathers), so a
request's result is the same bit for bit alone and in any mix.
"""

from __future__ import annotations

import dataclasses
import math
import time
from collections import deque

import numpy as np
import torch

from . import _lib
from . import cosmology as _c
from . import engine
from .pm.integrate import time_grid
from .pm.interp import NoInterp
from .pm.solvers import EfficientDriftDoubleKick, StepOp
from .painting import PaintingOptions


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

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def forward_ops(req: SimRequest):
    """(D, vel_factor, ktab, ptab, ops): the host side of one request -- LPT factors, the colouring's
    P(k) table (times N / V), and the KDK time loop as StepOps with each step's leading drift fused
    into the previous step's readout (as integrate() does)."""
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


class _Slot:
    """Device state of one request in flight."""

    def __init__(self, n, device, plan):
        self.n = n
        self.plan = plan
        self.stream = torch.cuda.Stream(device=device)
        shape = (n, n, n)
        self.noise = torch.empty(shape, dtype=torch.float32, device=device)
        self.ic = torch.empty(shape, dtype=torch.float32, device=device)
        self.x = torch.empty(shape + (3,), dtype=torch.float32, device=device)
        self.v = torch.empty(shape + (3,), dtype=torch.float32, device=device)
        self.rho = torch.empty(shape, dtype=torch.float32, device=device)
        self.ktab = torch.empty(128, dtype=torch.float64, device=device)
        self.ptab = torch.empty(128, dtype=torch.float64, device=device)
        self.host_tab = torch.empty((2, 128), dtype=torch.float64, pin_memory=True)
        self.req = None
        self.stages = None
        self.next = 0
        self.events = deque()
        self.admitted = 0.0

    def bytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in (self.noise, self.ic, self.x, self.v, self.rho)) + \
            int(self.plan.workspace_bytes)  # the plan's workspace includes its paint-sort buffers


class Server:
    """Continuous-batching forward-simulation server (one GPU, one host thread)."""

    def __init__(self, device=None, hbm_budget: int | None = None, max_active: int = 64, lookahead: int = 4,
                 admit_per_pass: int = 4):
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * 0.8)
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.lookahead = int(lookahead)
        self.admit_per_pass = int(admit_per_pass)
        self._pool: dict[int, list[_Slot]] = {}
        self._slot_bytes: dict[int, int] = {}

    # -- slots -----------------------------------------------------------------
    def prewarm(self, counts: dict) -> None:
        """Create `counts[n]` pooled slots for mesh size n now (plan creation allocates synchronously)."""
        for n, c in counts.items():
            have = len(self._pool.get(n, []))
            for _ in range(max(0, int(c) - have)):
                self._new_slot(int(n), keep=True)
            self._slot_bytes.setdefault(int(n), self._pool[int(n)][0].bytes())
        torch.cuda.synchronize(self.device)

    def _slot_size(self, n: int) -> int:
        if n not in self._slot_bytes:
            self._slot_bytes[n] = self._new_slot(n, keep=True).bytes()
        return self._slot_bytes[n]

    def _new_slot(self, n: int, keep: bool = False) -> _Slot:
        with torch.cuda.device(self.device):
            plan = _lib.Plan((n, n, n), self.device)
            s = _Slot(n, self.device, plan)
        if keep:
            self._pool.setdefault(n, []).append(s)
        return s

    def _take_slot(self, n: int) -> _Slot:
        pool = self._pool.get(n)
        return pool.pop() if pool else self._new_slot(n)

    def _give_back(self, s: _Slot) -> None:
        s.req, s.stages, s.next = None, None, 0
        s.events.clear()
        self._pool.setdefault(s.n, []).append(s)

    # -- request stages ----------------------------------------------------------
    def _stages(self, s: _Slot, req: SimRequest):
        D, vf, k, ptab, ops, fs = forward_ops(req)
        n, box, plan, mesh = req.mesh, req.box_size(), s.plan, (req.mesh,) * 3
        s.host_tab[0].copy_(torch.from_numpy(k))
        s.host_tab[1].copy_(torch.from_numpy(ptab))

        def ics():
            s.ktab.copy_(s.host_tab[0], non_blocking=True)
            s.ptab.copy_(s.host_tab[1], non_blocking=True)
            engine.white_noise(mesh, req.seed, out=s.noise, plan=plan)
            engine.color_field(s.noise, (box,) * 3, s.ktab, s.ptab, out=s.ic, plan=plan)
            engine.lpt1(s.ic, D, vf, out=(s.x, s.v), plan=plan)

        def step(op):
            return lambda: engine.step(s.x, s.v, mesh, force_scale=fs, c_kick=op.kick, c_drift_pre=op.pre,
                                       c_drift_post=op.post, plan=plan)

        def final():
            engine.paint(s.x, mesh, out=s.rho, plan=plan)

        return [ics] + [step(op) for op in ops] + [final]

    # -- main loop ---------------------------------------------------------------
    def serve(self, requests, clock=time.perf_counter):
        """Run `requests` (arrival times relative to the start); returns ({rid: SimResult}, stats)."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        for n in {r.mesh for r in pending}:  # measure slot sizes up front (plan creation is synchronous)
            if self._slot_size(n) > self.hbm_budget:
                raise ValueError(f"a {n}^3 request needs {self._slot_size(n)} bytes, above the HBM budget")
        active: list[_Slot] = []
        results: dict[int, SimResult] = {}
        used = 0
        peak_active = 0
        t0 = clock()
        while pending or active:
            now = clock() - t0
            # admission: arrival order, by free slot and HBM budget; a few per pass so that the host
            # preparation of new requests interleaves with issuing the queued ones' steps
            admitted = 0
            while (pending and pending[0].arrival <= now and len(active) < self.max_active
                   and admitted < self.admit_per_pass):
                req = pending[0]
                need = self._slot_size(req.mesh)
                if used + need > self.hbm_budget:
                    break
                pending.popleft()
                s = self._take_slot(req.mesh)
                s.req, s.admitted = req, now
                s.stages = self._stages(s, req)
                used += need
                active.append(s)
                admitted += 1
            peak_active = max(peak_active, len(active))
            issued = False
            for s in active:  # round-robin: one stage per request per pass
                while s.events and s.events[0].query():
                    s.events.popleft()
                if s.next < len(s.stages) and len(s.events) < self.lookahead:
                    with torch.cuda.stream(s.stream):
                        s.stages[s.next]()
                        ev = torch.cuda.Event()
                        ev.record(s.stream)
                    s.events.append(ev)
                    s.next += 1
                    issued = True
            done = [s for s in active if s.next == len(s.stages) and all(e.query() for e in s.events)]
            for s in done:
                fin = clock() - t0
                r = s.req
                results[r.rid] = SimResult(r.rid, r.mesh, s.x.clone(), s.rho.clone(), r.arrival, s.admitted, fin)
                active.remove(s)
                used -= self._slot_size(s.n)
                self._give_back(s)
            if not issued and not done:
                if active:
                    active[0].events[0].synchronize() if active[0].events else None
                elif pending:
                    time.sleep(max(0.0, min(1e-3, pending[0].arrival - (clock() - t0))))
        torch.cuda.synchronize(self.device)
        wall = clock() - t0
        lat = np.asarray([r.latency for r in results.values()])
        stats = {"requests": len(results), "wall_s": wall, "throughput_req_s": len(results) / wall if wall > 0 else 0.0,
                 "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
                 "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
                 "mean_ms": float(lat.mean() * 1e3) if lat.size else None, "peak_active": peak_active}
        return results, stats


def poisson_requests(count: int, rate_per_s: float, seed: int = 0, meshes=(64, 96, 128, 192)):
    """The serving workload: Poisson arrivals, mesh uniform over `meshes` (box 3.9 N Mpc/h), per-request
    seed, Omega_c ~ U(0.2, 0.35), sigma8 ~ U(0.7, 0.9), 1LPT + KDK with dt0 = 0.09 from a = 0.1 to 1."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, size=count)) if math.isfinite(rate_per_s) else np.zeros(count)
    return [SimRequest(rid=i, mesh=int(rng.choice(meshes)), seed=int(rng.integers(1, 2**31)),
                       Omega_c=float(rng.uniform(0.2, 0.35)), sigma8=float(rng.uniform(0.7, 0.9)),
                       arrival=float(t[i])) for i in range(count)]
