#!/usr/bin/env python3
"""PM step throughput of the fused MI355X engine on the single-GPU forward workload.

Workload (the reference's headline single-GPU forward configuration): 256^3 particles on a 256^3 mesh, box
1000 Mpc/h, float32, seeded Gaussian ICs (Philox white noise x EH98 P(k), Omega_c=0.25,
sigma8=0.8, Omega_b=0.0486, h=0.6774, n_s=0.9667), 1LPT to a=0.1, EfficientDriftDoubleKick with
dt0=0.05 to a=1, snapshots at 4 HEALPix shells (nside 256, bilinear, equal comoving width out to
z=0.5) plus a final a=1 snapshot whose 3D P(k) is measured.

A "step" is one PM step of that schedule (paint, force mesh, readout + kick + drift), with the
shell paints that fall between steps inside the timed region. --steps K times K consecutive steps
of the schedule (wrapping around to the initial state if K exceeds its length); --warmup W runs W
untimed steps first. Prints one JSON line.
"""

from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PM steps/sec @256^3 fused vs plain GPU vs CPU; fwd+VJP cost; served p99 latency"
N, BOX, A0, DT0, NSIDE, NSHELL, ZMAX = 256, 1000.0, 0.1, 0.05, 256, 4, 0.5
PEAK_HBM = 8.0e12
CAL_FLOATS = 128 << 20  # calibration copy for the HBM-traffic counters: 512 MiB read + 512 MiB written
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pm_step_256cubed_pmc_traffic.txt")
# kernels of each pmx_step phase (substring of the demangled name), for the rocprofv3 counter summary
PHASE_KERNELS = {
    "paint": ("k_sort_count", "k_sort_scan", "k_sort_scatter", "k_paint_deposit", "k_paint_faces"),
    "r2c_z": ("k_r2c_z",),
    "c2c_y": ("k_c2c_lines<256, false>",),
    "kspace_x": ("k_xpass<256, 5>",),
    "c2c_y_inv": ("k_c2c_lines<256, true>",),
    "c2r_z": ("k_c2r_z<256, 1>",),
    "grad4": ("k_grad4",),
    "readout_kick_drift": ("k_readout_lds<3>",),
}


def build_schedule(pkg, cosmo):
    """(ops, shells): ops is a list of ('op', StepOp) / ('shell', i) / ('final',) in run order, with
    drifts fused across steps exactly as pm.integrate does."""
    from importlib import import_module

    cosmology = pkg.cosmology
    solvers = import_module("jax_fli_amd.pm.solvers")
    rmax = float(cosmology.radial_comoving_distance(cosmo, cosmology.z2a(ZMAX)))
    edges = np.linspace(0.0, rmax, NSHELL + 1)
    centers = 0.5 * (edges[1:] + edges[:-1])[::-1]  # far -> near
    widths = np.diff(edges)[::-1]
    a_c = cosmology.a_of_chi(cosmo, centers).astype(np.float32)
    ts = np.concatenate([a_c, np.array([1.0], np.float32)])
    solver = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="particles")))
    raw = [("op", solver.init_op(A0, A0 + DT0, DT0, cosmo))]
    for ev in pkg.pm.time_grid(A0, ts, DT0, 1.0):
        if ev[0] == "step":
            raw.append(("op", solver.step_op(ev[1], ev[2], DT0, A0, cosmo)))
        else:
            raw.append(("shell", ev[1]) if ev[1] < NSHELL else ("final",))
    ops, pending = [], None
    for item in raw:
        if item[0] == "op":
            op = solvers.StepOp(item[1].pre, item[1].kick, item[1].post)
            if pending is not None and pending.post is None and op.pre is not None:
                pending.post, op.pre = op.pre, None
            if pending is not None:
                ops.append(("op", pending))
            pending = op
        else:
            if pending is not None:
                ops.append(("op", pending))
                pending = None
            ops.append(item)
    return ops, list(zip(centers, widths))


def phase_bytes(n_p, n_c, nspec):
    """Algorithmic HBM bytes per launch of each pmx_step phase (float32 / complex64)."""
    return {
        "paint": 12 * n_p + 4 * n_c,
        "r2c_z": 4 * n_c + 8 * nspec,
        "c2c_y": 16 * nspec,
        "kspace_x": 16 * nspec,
        "c2c_y_inv": 16 * nspec,
        "c2r_z": 8 * nspec + 4 * n_c,
        "grad4": 4 * n_c + 12 * n_c,
        "readout_kick_drift": 48 * n_p + 12 * n_c,
    }


def _pmc_per_kernel(csv_path):
    """{kernel name: (dispatches, mean counter value)} from a rocprofv3 counter_collection.csv."""
    import csv

    acc = {}
    with open(csv_path) as fh:
        for r in csv.DictReader(fh):
            n, tot = acc.get(r["Kernel_Name"], (0, 0.0))
            acc[r["Kernel_Name"]] = (n + 1, tot + float(r["Counter_Value"]))
    return {k: (n, tot / n) for k, (n, tot) in acc.items()}


def pmc_summary(fetch_csv, write_csv):
    """HBM traffic per launch of each step phase from one FETCH_SIZE run and one WRITE_SIZE run (KiB per
    dispatch), each scaled by what the same counter reads for the calibration copy (torch.neg over
    CAL_FLOATS floats: exactly 4*CAL_FLOATS bytes read and written)."""
    fetch, write = _pmc_per_kernel(fetch_csv), _pmc_per_kernel(write_csv)
    cal = [k for k in fetch if "neg_kernel" in k]
    if not cal:
        raise SystemExit("calibration kernel (neg_kernel) not found in the counter file")
    cal_bytes = 4.0 * CAL_FLOATS
    f_scale = cal_bytes / (fetch[cal[0]][1] * 1024.0)
    w_scale = cal_bytes / (write[cal[0]][1] * 1024.0)
    phases = {}
    for ph, names in PHASE_KERNELS.items():
        fb = wb = 0.0
        found = False
        for nm in names:
            for kf, (n, val) in fetch.items():
                if nm in kf:
                    fb += val * 1024.0 * f_scale
                    wb += write.get(kf, (0, 0.0))[1] * 1024.0 * w_scale
                    found = True
        if found:
            phases[ph] = {"fetch_bytes": round(fb), "write_bytes": round(wb), "traffic_bytes": round(fb + wb)}
    return {"fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4), "phases": phases}


def read_traffic(phase):
    """HBM bytes per launch of `phase` from the committed counter summary (its last line is JSON)."""
    try:
        with open(TRAFFIC_FILE) as fh:
            last = [ln for ln in fh if ln.strip()][-1]
        return json.loads(last)["phases"][phase]["traffic_bytes"]
    except (OSError, KeyError, ValueError, IndexError):
        return None


def bench_vjp(pkg, dev):
    """Differentiable configuration: 128^3 particles on a 128^3 mesh, box 500 Mpc/h, ReversibleDoubleKickDrift
    (dt0 = 0.09, a = 0.1 -> 1, reverse adjoint), loss sum((delta - delta_target)^2) with the target from a
    second seed; gradient w.r.t. the white-noise field and (Omega_c, sigma8). Returns forward and
    forward+VJP wall times and the agreement with central finite differences (eps = 1e-3, 16 random
    directions of the noise field, float32 forward)."""
    import torch

    n, box, a0, dt0 = 128, 500.0, 0.1, 0.09
    P = dict(Omega_b=0.0486, h=0.6774, n_s=0.9667)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")))

    def loss(noise, Oc, s8, target):
        c = pkg.Cosmology(Omega_c=Oc, sigma8=s8, **P)
        ic = pkg.interpolate_initial_conditions(noise, (n,) * 3, (box,) * 3, cosmo=c)
        dx, p = pkg.lpt(c, ic, ts=a0)
        out = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=dt0, ts=[1.0], solver=solver, adjoint="reverse")
        rho = out.array.double()
        delta = rho / rho.mean() - 1.0
        return delta if target is None else ((delta - target) ** 2).sum()

    with torch.no_grad():
        target = loss(pkg.engine.white_noise((n,) * 3, 2), 0.25, 0.8, None)
    noise0 = pkg.engine.white_noise((n,) * 3, 1)

    def timed(fn, reps=5):
        ts_ = []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts_.append(time.perf_counter() - t)
        return float(np.median(ts_[1:])) * 1e3

    def fwd():
        with torch.no_grad():
            return float(loss(noise0, 0.25, 0.8, target))

    grads = {}

    def fwd_bwd():
        nz = noise0.clone().requires_grad_(True)
        Oc = torch.tensor(0.25, dtype=torch.float64, device=dev, requires_grad=True)
        s8 = torch.tensor(0.8, dtype=torch.float64, device=dev, requires_grad=True)
        L = loss(nz, Oc, s8, target)
        L.backward()
        grads.update(noise=nz.grad, Oc=Oc.grad, s8=s8.grad)

    def fwd_bwd_noise():
        nz = noise0.clone().requires_grad_(True)
        loss(nz, 0.25, 0.8, target).backward()

    fwd_ms = timed(fwd)
    fwd_bwd_noise_ms = timed(fwd_bwd_noise)
    fwd_bwd_ms = timed(fwd_bwd)
    g = grads["noise"].double()
    errs = []
    eps = 1e-3
    for j in range(16):
        d = torch.randn(noise0.shape, generator=torch.Generator(device=dev).manual_seed(100 + j), device=dev)
        with torch.no_grad():
            fd = (float(loss(noise0 + eps * d, 0.25, 0.8, target)) - float(loss(noise0 - eps * d, 0.25, 0.8, target))) / (
                2 * eps)
        errs.append(abs(float((g * d.double()).sum()) - fd) / abs(fd))
    with torch.no_grad():
        h = 1e-3
        fd_oc = (float(loss(noise0, 0.25 + h, 0.8, target)) - float(loss(noise0, 0.25 - h, 0.8, target))) / (2 * h)
        fd_s8 = (float(loss(noise0, 0.25, 0.8 + h, target)) - float(loss(noise0, 0.25, 0.8 - h, target))) / (2 * h)
    # the reference's default differentiable path: EfficientDriftDoubleKick with adjoint="checkpointed"
    kdk = pkg.EfficientDriftDoubleKick(interp_kernel=pkg.NoInterp(painting=pkg.PaintingOptions(target="density")))

    def loss_kdk(noise, Oc, s8):
        c = pkg.Cosmology(Omega_c=Oc, sigma8=s8, **P)
        ic = pkg.interpolate_initial_conditions(noise, (n,) * 3, (box,) * 3, cosmo=c)
        dx, p = pkg.lpt(c, ic, ts=a0)
        rho = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=dt0, ts=[1.0], solver=kdk).array.double()
        return ((rho / rho.mean() - 1.0 - target) ** 2).sum()

    def kdk_fwd():
        with torch.no_grad():
            return float(loss_kdk(noise0, 0.25, 0.8))

    def kdk_fwd_bwd():
        nz = noise0.clone().requires_grad_(True)
        Oc = torch.tensor(0.25, dtype=torch.float64, device=dev, requires_grad=True)
        s8 = torch.tensor(0.8, dtype=torch.float64, device=dev, requires_grad=True)
        loss_kdk(nz, Oc, s8).backward()

    kdk_fwd_ms, kdk_fwd_bwd_ms = timed(kdk_fwd), timed(kdk_fwd_bwd)
    checkpointed = {"workload": "the same model with EfficientDriftDoubleKick and the default adjoint='checkpointed' "
                                "(every op's state kept in HBM), gradient w.r.t. the noise and (Omega_c, sigma8)",
                    "forward_ms": round(kdk_fwd_ms, 2), "forward_plus_vjp_ms": round(kdk_fwd_bwd_ms, 2),
                    "vjp_over_forward": round(kdk_fwd_bwd_ms / kdk_fwd_ms, 2)}
    return {"workload": "128^3 particles, 128^3 mesh, box 500 Mpc/h, ReversibleDoubleKickDrift dt0=0.09 a=0.1->1 "
                        "(10 steps), reverse adjoint, L2 loss on the final density vs a second seed",
            "checkpointed_kdk": checkpointed,
            "forward_ms": round(fwd_ms, 2), "forward_plus_vjp_ms": round(fwd_bwd_ms, 2),
            "vjp_over_forward": round(fwd_bwd_ms / fwd_ms, 2),
            "noise_only": {"forward_plus_vjp_ms": round(fwd_bwd_noise_ms, 2),
                           "vjp_over_forward": round(fwd_bwd_noise_ms / fwd_ms, 2)},
            "note": "wall clock; the (Omega_c, sigma8) gradient adds float64 finite differences of the host "
                    "growth/coefficient/P(k) functions (see vjp_over_forward vs noise_only)",
            "fd_check": {"eps": eps, "directions": 16, "median_rel_err": float(np.median(errs)),
                         "max_rel_err": float(np.max(errs)),
                         "Omega_c": {"adjoint": float(grads["Oc"]), "fd": fd_oc},
                         "sigma8": {"adjoint": float(grads["s8"]), "fd": fd_s8}}}


def bench_large(pkg, dev, fd_dirs=2, sweep=(1e-2, 3e-3, 1e-3, 3e-4)):
    """Large single-GPU configuration: 512^3 particles (the regular sub-lattice, 2 cells apart) on a
    1024^3 force mesh, box 1000 Mpc/h, float32; seeded ICs, 1LPT at a = 0.1, ReversibleDoubleKickDrift
    with dt0 = 0.03 (reverse adjoint); 8 HEALPix shells (nside 512, bilinear) of equal comoving width out
    to z = 1, the shells beyond the box painted from its periodic images (OnionTiler with layers=2,
    rotate=False, no drift to the lightcone); Born convergence for one top-hat n(z) on 0 < z < 1. Times
    the forward (noise -> kappa) and the forward + VJP of sum(kappa^2) w.r.t. the white-noise field.
    The gradient is also set beside central differences of the float32 forward along the noise field
    itself (an amplitude rescaling, where they agree) and along the gradient's own direction at eps from
    1e-2 to 3e-4 (float32 positions cannot resolve smaller steps). Along the gradient the differences do
    not reach the adjoint: the loss is strongly curved along the direction that raises it fastest, and
    particle-images crossing shell edges move whole particle weights between shells (shell membership is
    a step, which the adjoint takes as zero, as autodiff does). The float64 test of the same structure at
    16^3 particles (tests/test_gpu_sky.py) shows both effects: there the differences reach the adjoint only
    at eps <= 1e-5 with the membership frozen, and a single crossing shifts them by 9%."""
    import torch

    n, npart, box, a0, dt0, nside, nsh, zmax = 1024, 512, 1000.0, 0.1, 0.03, 512, 8, 1.0
    c = pkg.Cosmology(Omega_c=0.25, sigma8=0.8, Omega_b=0.0486, h=0.6774, n_s=0.9667)
    chi_max = float(pkg.cosmology.radial_comoving_distance(c, 1.0 / (1.0 + zmax)))
    edges = np.linspace(0.0, chi_max, nsh + 1)
    r_c = (0.5 * (edges[1:] + edges[:-1]))[::-1]  # far -> near: ascending time
    w = np.diff(edges)[::-1]
    ts = np.stack([pkg.cosmology.a_of_chi(c, r_c - w / 2), pkg.cosmology.a_of_chi(c, r_c + w / 2)])
    layers = int(math.ceil((chi_max - box / 2) / box))
    painting = pkg.PaintingOptions(target="spherical", scheme="bilinear")
    tiler = pkg.OnionTiler(painting=painting, drift_on_lightcone=False, layers=layers, rotate=False)
    solver = pkg.ReversibleDoubleKickDrift(interp_kernel=tiler)
    lat = torch.stack(torch.meshgrid(*[torch.arange(npart, dtype=torch.float32, device=dev) * (n // npart)] * 3,
                                     indexing="ij"), dim=-1)
    nz = [pkg.tophat_z(0.0, zmax)]
    info = {}

    def loss(noise):
        ic = pkg.interpolate_initial_conditions(noise, (n,) * 3, (box,) * 3, cosmo=c, nside=nside)
        dx, p = pkg.lpt(c, ic, ts=a0, initial_particles=lat)
        del ic
        lc = pkg.nbody(c, dx, p, t0=a0, t1=1.0, dt0=dt0, ts=ts, solver=solver, adjoint="reverse")
        kap = pkg.born(c, lc, nz)
        info["kappa"] = kap.array
        return (kap.array.double() ** 2).sum()

    noise0 = pkg.engine.white_noise((n,) * 3, 5)

    def timed(fn, reps):
        out = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            out.append((time.perf_counter() - t) * 1e3)
        return float(np.median(out))

    def fwd():
        with torch.no_grad():
            return loss(noise0)

    grad = {}

    def fwd_bwd():
        nzt = noise0.clone().requires_grad_(True)
        loss(nzt).backward()
        grad["g"] = nzt.grad

    fwd()  # warm-up: plans, pools
    fwd_ms = timed(fwd, 2)
    fwd_bwd()
    fwd_bwd_ms = timed(fwd_bwd, 1)
    kap = info.pop("kappa")
    krms = float(kap.double().pow(2).mean().sqrt())
    g = grad.pop("g").double()
    errs = {}
    eps = 1e-3
    gn = float(g.norm())
    for name in ("noise", "gradient")[:fd_dirs]:
        d = noise0 if name == "noise" else (g * (math.sqrt(g.numel()) / gn)).float()
        with torch.no_grad():
            fd = (float(loss(noise0 + eps * d)) - float(loss(noise0 - eps * d))) / (2 * eps)
        adj = float((g * d.double()).sum())
        errs[name] = {"adjoint": adj, "fd": fd, "rel_err": round(abs(adj - fd) / abs(fd), 6)}
        if sweep and name == "gradient":  # central differences at several steps along the gradient
            errs[name]["sweep"] = {}
            for e in sweep:
                with torch.no_grad():
                    fde = (float(loss(noise0 + e * d)) - float(loss(noise0 - e * d))) / (2 * e)
                errs[name]["sweep"][f"{e:g}"] = fde
        del d
    nsteps = pkg.pm.count_steps(a0, np.sort(np.asarray(pkg.cosmology.a_of_chi(c, r_c), np.float32)), dt0, 1.0)
    del g, noise0, lat
    pkg._lib.release_plan((n,) * 3, dev)
    torch.cuda.empty_cache()
    return {"workload": "512^3 particles (sub-lattice) on a 1024^3 mesh, box 1000 Mpc/h, 1LPT a=0.1, "
                        f"ReversibleDoubleKickDrift dt0=0.03 ({nsteps} steps to the nearest shell), 8 HEALPix shells "
                        f"nside 512 to z=1 from {(2 * layers + 1) ** 3} periodic images, Born for tophat n(z) 0<z<1",
            "forward_ms": round(fwd_ms, 1), "forward_plus_vjp_ms": round(fwd_bwd_ms, 1),
            "vjp_over_forward": round(fwd_bwd_ms / fwd_ms, 2), "steps": nsteps,
            "kappa_rms": krms, "fd_check": {"eps": eps, **errs},
            "note": "wall clock including white noise, IC colouring, LPT, the N-body run with its 8 shell paints and "
                    "Born; the VJP runs the reverse adjoint with the shell-paint and Born adjoints"}


def cpu_threads():
    """Host threads of the CPU baseline: the process's CPU share (OMP_NUM_THREADS when set, as on a box
    that shares its host between GPUs), else every core in the affinity mask."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(aff, int(env))) if env and env.isdigit() else aff


def bench_config1(pkg, dev, cpu=True):
    """The 64^3 workload (64^3 particles on a 64^3 mesh, box 250 Mpc/h, 1LPT to a = 0.1, then
    EfficientDriftDoubleKick with dt0 = 0.09 to a = 1): PM steps/s of the fused step, the unfused per-op
    GPU step and the CPU restatement on all threads and on one."""
    import torch
    from importlib import import_module

    from oracle import pm_oracle as orc

    n, box = 64, 250.0
    req = pkg.serving.SimRequest(rid=0, mesh=n, seed=0, Omega_c=0.25, sigma8=0.8, box=box, dt0=0.09)
    cosmo = req.cosmology()
    _, _, _, _, ops, force_scale = pkg.serving.forward_ops(req)
    ic = pkg.gaussian_initial_conditions(0, (n,) * 3, (box,) * 3, cosmo=cosmo)
    dx, p = pkg.lpt(cosmo, ic, ts=req.a0)
    x0, v0 = dx.array.contiguous(), p.array.contiguous()
    x, v = x0.clone(), v0.clone()
    unf = import_module("jax_fli_amd.pm.unfused").UnfusedPM((n,) * 3, dev)

    def fused(o):
        pkg.engine.step(x, v, (n,) * 3, force_scale=force_scale, c_kick=o.kick, c_drift_pre=o.pre, c_drift_post=o.post)

    def unfused(o):
        unf.step(x, v, force_scale, o.kick, o.post or 0.0, c_drift_pre=o.pre)

    def per_step(fn, reps=20):
        for _ in range(2):
            x.copy_(x0)
            v.copy_(v0)
            for o in ops:
                fn(o)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            x.copy_(x0)
            v.copy_(v0)
            for o in ops:
                fn(o)
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / (reps * len(ops))

    out = {"workload": "64^3 particles on 64^3 mesh, box 250 Mpc/h, KDK dt0 0.09 from a=0.1 to 1",
           "steps": len(ops), "fused_gpu_steps_per_s": round(1e3 / per_step(fused), 1),
           "unfused_gpu_steps_per_s": round(1e3 / per_step(unfused), 1)}
    if cpu:
        xc0, vc0 = x0.cpu().numpy(), v0.cpu().numpy()
        for label, w in (("cpu_steps_per_s", cpu_threads()), ("cpu_1core_steps_per_s", 1)):
            xc, vc = xc0, vc0
            orc.cpu_pm_step(xc, vc, force_scale, ops[1].kick, ops[1].post or 0.0, workers=w)  # plans, pages
            t = time.perf_counter()
            for o in ops:
                if o.pre is not None:
                    xc = xc + vc * np.float32(o.pre)
                xc, vc = orc.cpu_pm_step(xc, vc, force_scale, o.kick, o.post or 0.0, workers=w)
            out[label] = round(len(ops) / (time.perf_counter() - t), 2)
        out["cpu_threads"] = cpu_threads()
    return out


def bench_serving(pkg, dev, count=512, slots=None):
    """Serving workload: `count` forward-simulation requests, mesh uniform over {64, 96, 128, 192}^3
    (box 3.9 N Mpc/h), own seed, Omega_c ~ U(0.2, 0.35), sigma8 ~ U(0.7, 0.9), 1LPT + KDK dt0 = 0.09 to
    a = 1 (10 steps), float32, through the native server (batched graph-replayed steps per size class).
    Reports the saturated throughput (all requests at t = 0), the same requests one at a time (each
    submitted when the previous one is done), and p50/p99 latency of a Poisson stream offered at 70% of
    the saturated throughput."""
    meshes = (64, 96, 128, 192)
    slots = slots or {64: 32, 96: 24, 128: 16, 192: 8}
    srv = pkg.serving.Server(device=dev, meshes=meshes, slots=slots)
    try:
        burst = pkg.serving.poisson_requests(count, float("inf"), seed=11)
        srv.serve(burst[:32])  # warm-up
        _, sat = srv.serve(burst)
        _, serial = srv.serve(burst, sequential=True)
        rate = 0.7 * sat["throughput_req_s"]
        _, pois = srv.serve(pkg.serving.poisson_requests(count, rate, seed=12))
        device_gb = srv.stats()["device_bytes"] / 1e9
    finally:
        srv.close()
    r3 = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items() if k != "device_bytes"}
    return {"workload": f"{count} requests, mesh uniform over 64/96/128/192^3, box 3.9 N Mpc/h, 1LPT + KDK "
                        "dt0=0.09 a=0.1->1, per-request seed and (Omega_c, sigma8)",
            "server": {"slots": {str(m): int(slots[m]) for m in meshes}, "device_GB": round(device_gb, 2),
                       "form": "native server: one batched launch per step phase over a size class's live slots, "
                               "replayed from a captured HIP graph; host preparation on 4 native threads"},
            "saturated": r3(sat), "one_at_a_time": r3(serial),
            "poisson": {"offered_req_s": round(rate, 1), **r3(pois)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed PM steps (0: one full schedule)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=1, help="CPU-baseline steps at full size (0: skip)")
    ap.add_argument("--no-vjp", action="store_true", help="skip the forward+VJP measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-large", action="store_true", help="skip the 512^3-on-1024^3 forward + VJP measurement")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the 64^3 workload (keeps profiler runs to 256^3 dispatches)")
    ap.add_argument("--vjp-only", action="store_true",
                    help="print only the differentiable (128^3 forward + VJP) measurement and exit")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the 512^3-on-1024^3 forward + VJP timing (no differences) and print its JSON")
    ap.add_argument("--large-sweep", action="store_true",
                    help="run only the 512^3-on-1024^3 configuration with its central differences (along the noise, "
                         "and along the gradient at eps 1e-2, 3e-3, 1e-3, 3e-4) and print its JSON")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the per-phase HBM traffic from two rocprofv3 --pmc runs of this script and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary)))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (this engine's multi-GPU "
              "decomposition is a design, not a runnable path)", file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__

    pkg = __graft_entry__.load_package()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.vjp_only:
        print(json.dumps({"vjp": bench_vjp(pkg, dev)}))
        return
    if args.large_only:
        print(json.dumps({"large": bench_large(pkg, dev, fd_dirs=0)}))
        return
    if args.large_sweep:
        print(json.dumps({"large": bench_large(pkg, dev, fd_dirs=2)}))
        return
    cosmo = pkg.Cosmology(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
    mesh = (N, N, N)

    t_fwd0 = time.perf_counter()
    ic = pkg.gaussian_initial_conditions(0, mesh, (BOX,) * 3, cosmo=cosmo, nside=NSIDE)
    dx, p = pkg.lpt(cosmo, ic, ts=A0)
    torch.cuda.synchronize()
    x0, v0 = dx.array.contiguous(), p.array.contiguous()
    x, v = x0.clone(), v0.clone()
    ops, shells = build_schedule(pkg, cosmo)
    nsteps_sched = sum(1 for o in ops if o[0] == "op")
    obs = tuple(b * 0.5 for b in (BOX,) * 3)
    maps = torch.empty((NSHELL, 12 * NSIDE**2), device=dev)
    kedges = np.geomspace(2 * np.pi / BOX, np.pi * N / BOX, 33)
    force_scale = 1.5 * cosmo.Omega_m
    plan = pkg._lib.get_plan(mesh)

    def run_item(item):
        kind = item[0]
        if kind == "op":
            o = item[1]
            pkg.engine.step(x, v, mesh, force_scale=force_scale, c_kick=o.kick, c_drift_pre=o.pre, c_drift_post=o.post)
        elif kind == "shell":
            c, w = shells[item[1]]
            pkg.engine.paint_healpix(x, mesh, (BOX,) * 3, obs, c - w / 2, c + w / 2, NSIDE, scheme="bilinear",
                                     out=maps[item[1]])

    # full forward (LPT already done above) + final P(k): one untimed pass that also warms up
    for item in ops:
        run_item(item)
    rho = pkg.engine.paint(x, mesh)
    k_fin, pk_fin = pkg.power(rho, box_shape=(BOX,) * 3, kedges=kedges)
    torch.cuda.synchronize()
    forward_ms = (time.perf_counter() - t_fwd0) * 1e3

    def steps_iter(k, start=0):
        """Yield the schedule items covering k PM steps from position `start` (wrapping)."""
        done, i = 0, start
        while done < k:
            if i == len(ops):
                i = 0
                yield ("reset",)
            item = ops[i]
            i += 1
            yield item
            if item[0] == "op":
                done += 1

    shell_events = []

    def run(items, time_shells=False):
        for item in items:
            if item[0] == "reset":
                x.copy_(x0)
                v.copy_(v0)
            elif item[0] == "shell" and time_shells:
                a_, b_ = time_shells[len(shell_events)]
                a_.record()
                run_item(item)
                b_.record()
                shell_events.append((a_, b_))
            elif item[0] != "final":
                run_item(item)

    x.copy_(x0)
    v.copy_(v0)
    run(steps_iter(args.warmup))
    x.copy_(x0)
    v.copy_(v0)
    K = args.steps if args.steps > 0 else nsteps_sched
    items = list(steps_iter(K))
    torch.cuda.synchronize()
    plan.timing(True, steps=K)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    shell_pool = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(sum(1 for it in items if it[0] == "shell"))]
    gc.collect()
    gc.disable()  # no collector pauses while the host issues the timed work
    e0.record()
    run(items, time_shells=shell_pool)
    e1.record()
    torch.cuda.synchronize()
    gc.enable()
    elapsed_ms = e0.elapsed_time(e1)
    shell_ms = [round(a_.elapsed_time(b_), 3) for a_, b_ in shell_events]
    phases, nrec = plan.timing_read()
    plan.timing(False)
    ms_per_step = elapsed_ms / K
    value = 1000.0 / ms_per_step

    # roofline of the dominant kernel (largest summed time over the timed region)
    nspec = (N // 2 + 1) * N * N
    pb = phase_bytes(N**3, N**3, nspec)
    dom = max(phases, key=lambda k: phases[k])
    avg_ms = phases[dom] / max(nrec, 1)
    achieved = pb[dom] / (avg_ms * 1e-3)
    step_bytes = 60 * N**3 + 64 * N**3
    per_phase = {k: {"avg_us": round(phases[k] / max(nrec, 1) * 1e3, 2),
                     "GB_per_s": round(pb[k] / (phases[k] / max(nrec, 1) * 1e-3) / 1e9, 1)} for k in phases}

    # calibration copy for the traffic counters (one dispatch of a known 512 MiB read + 512 MiB write)
    cal = torch.ones(CAL_FLOATS, device=dev)
    torch.neg(cal, out=torch.empty_like(cal))
    del cal

    # unfused per-op GPU form of the same step (rocFFT c2c x4, 3 readouts, separate kick/drift)
    from importlib import import_module

    unf = import_module("jax_fli_amd.pm.unfused").UnfusedPM(mesh, dev)
    f_unf = unf.forces(x0, force_scale)
    f_fused = pkg.engine.forces(x0, mesh, scale=force_scale)
    force_err = float((f_unf - f_fused).norm() / f_unf.norm())
    xu, vu = x0.clone(), v0.clone()
    sched = [it[1] for it in ops if it[0] == "op"]

    def time_schedule(step_fn):
        """ms per step of the schedule's PM steps (shell paints left out) from the initial state."""
        xu.copy_(x0)
        vu.copy_(v0)
        torch.cuda.synchronize()
        a_, b_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a_.record()
        for o in sched:
            step_fn(o)
        b_.record()
        torch.cuda.synchronize()
        return a_.elapsed_time(b_) / len(sched)

    def unfused_step(o):
        unf.step(xu, vu, force_scale, o.kick, o.post or 0.0, c_drift_pre=o.pre)

    def fused_step(o):
        pkg.engine.step(xu, vu, mesh, force_scale=force_scale, c_kick=o.kick, c_drift_pre=o.pre, c_drift_post=o.post)

    time_schedule(unfused_step)
    unf_ms = time_schedule(unfused_step)
    fused_ms = time_schedule(fused_step)
    del xu, vu, unf, f_unf, f_fused

    # rocFFT beside the hand-written FFT (r2c + c2r round trip of one 256^3 mesh)
    def t_gpu(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    spec = pkg.engine.rfftn(rho)
    fft_ours = t_gpu(lambda: pkg.engine.irfftn(pkg.engine.rfftn(rho), mesh))
    fft_roc = t_gpu(lambda: torch.fft.irfftn(torch.fft.rfftn(rho), s=mesh))
    del spec

    vjp = None if args.no_vjp else bench_vjp(pkg, dev)
    large = None if args.no_large else bench_large(pkg, dev)
    serving = None if args.no_serving else bench_serving(pkg, dev)

    # CPU baseline: the NumPy/SciPy/C restatement of the reference step on the host's CPU share, and on
    # one core; the 64^3 workload beside it (GPU fused, GPU unfused, CPU)
    cpu = None
    config1 = None if args.no_config1 else bench_config1(pkg, dev, cpu=args.cpu_steps > 0)
    if args.cpu_steps > 0:
        from oracle import pm_oracle as orc

        cores = cpu_threads()
        o = ops[1][1]

        def cpu_time(workers, steps):
            xc, vc = x0.cpu().numpy(), v0.cpu().numpy()
            t = time.perf_counter()
            for _ in range(steps):
                xc, vc = orc.cpu_pm_step(xc, vc, force_scale, o.kick, o.post or 0.0, workers=workers)
            return (time.perf_counter() - t) / steps

        cpu_s = cpu_time(cores, args.cpu_steps)
        one_s = cpu_time(1, 1)
        cpu = {"value": round(1.0 / cpu_s, 5), "unit": "PM steps/s", "cores": cores, "kind": "port",
               "sample": f"{args.cpu_steps} full 256^3 PM step(s) of the NumPy/SciPy/C restatement of the reference "
                         f"(c2c scipy.fft, OpenMP CIC scatter/gather, {cores} threads): {cpu_s:.2f} s/step; "
                         f"one 256^3 step on 1 thread: {one_s:.1f} s",
               "one_core": {"value": round(1.0 / one_s, 5), "s_per_step": round(one_s, 2)},
               "config1_64cubed": config1}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "PM steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox white noise x EH98 P(k), 1LPT at a=0.1)",
        "config": {"workload": "single-GPU forward: 256^3 particles, 256^3 mesh, box 1000 Mpc/h, KDK dt0=0.05 "
                               "a=0.1->1, 4 HEALPix shells nside 256 to z=0.5",
                   "mesh": N, "particles": N**3, "box_mpc_h": BOX, "dt0": DT0, "schedule_steps": nsteps_sched,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1), "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM, 4), "traffic": read_traffic(dom),
                     "algorithmic_bytes_per_launch": pb[dom], "avg_launch_us": round(avg_ms * 1e3, 2)},
        "step_roofline": {"algorithmic_bytes_per_step": step_bytes,
                          "achieved_GB_per_s": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                          "frac": round(step_bytes / (ms_per_step * 1e-3) / PEAK_HBM, 4)},
        "phases": per_phase,
        "shell_paint_ms": shell_ms,
        "cpu_baseline": cpu,
        "config1_64cubed": config1,
        "unfused_gpu": {"ms_per_step": round(unf_ms, 4), "fused_ms_per_step": round(fused_ms, 4),
                        "speedup": round(unf_ms / fused_ms, 2), "force_rel_l2_vs_fused": force_err,
                        "form": "atomic scatter paint + rocFFT c2c fwd + 3 x (k-space multiply, rocFFT c2c inv, "
                                "1-component readout) + separate kick and drift (torch ops); both forms timed over "
                                "the schedule's PM steps without the shell paints"},
        "forward_total_ms": round(forward_ms, 1),
        "vjp": vjp,
        "large": large,
        "serving": serving,
        "fft_roundtrip_ms": {"pmx": round(fft_ours, 4), "rocfft": round(fft_roc, 4)},
        "final_pk": {"k": [round(float(a), 5) for a in k_fin[:8]], "pk": [round(float(a), 3) for a in pk_fin[:8]]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
