"""Writes config3_gradient_fd.json: float64 central differences of the oracle's differentiable forward
model at the differentiable workload's full size, for tests/test_gpu_configs.py.

Model (the differentiable configuration): 128^3 white noise (numpy default_rng(301), float32) ->
EH98 colouring (box 500 Mpc/h) -> LPT1 at a = 0.1 -> ReversibleDoubleKickDrift, dt0 = 0.09, to a = 1 ->
CIC density -> L = sum((delta - delta_target)^2), delta = rho / mean(rho) - 1, with delta_target the
same model's density from the noise of default_rng(302). Directional derivatives along the noise
directions default_rng(310 + j).standard_normal (j = 0..3) and in Omega_c and sigma8, all in float64
by the oracle (oracle/nbody_oracle.py, oracle/pm_oracle.py).

Run from the repository root: python tests/golden/make_config3_fd.py (about 5 minutes on 8 cores).
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import nbody_oracle as nbo  # noqa: E402
from oracle import pm_oracle as orc  # noqa: E402

P = dict(Omega_c=0.25, Omega_b=0.0486, h=0.6774, n_s=0.9667, sigma8=0.8)
N, BOX, A0, T1, DT0 = 128, 500.0, 0.1, 1.0, 0.09
KT = np.logspace(-4, 1, 128)
SEED_NOISE, SEED_TARGET, SEED_DIR = 301, 302, 310
NDIR, EPS, EPS_CHECK, H = 4, 1e-5, 1e-6, 1e-6
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config3_gradient_fd.json")


def noise_of(seed):
    return np.random.default_rng(seed).standard_normal((N,) * 3).astype(np.float32)


def delta_of(noise, Oc=P["Omega_c"], s8=P["sigma8"]):
    o = nbo.Cosmo(**{**P, "Omega_c": Oc, "sigma8": s8})
    ic = nbo.color_field(noise, (BOX,) * 3, KT, nbo.linear_matter_power(o, KT))
    dx, p = nbo.lpt1(o, ic, A0)
    x, _ = nbo.run_nbody(o, dx, p, A0, T1, DT0, [T1], solver="kkd")
    rho = orc.cic_paint(x, (N,) * 3, "relative", dtype=np.float64)
    return rho / rho.mean() - 1.0


def main():
    t0 = time.time()
    target = delta_of(noise_of(SEED_TARGET).astype(np.float64))
    noise = noise_of(SEED_NOISE).astype(np.float64)

    def loss(nz, **kw):
        return float(((delta_of(nz, **kw) - target) ** 2).sum())

    out = {"model": __doc__.split("\n\n")[1], "N": N, "box": BOX, "a0": A0, "t1": T1, "dt0": DT0,
           "params": P, "seed_noise": SEED_NOISE, "seed_target": SEED_TARGET, "seed_dir": SEED_DIR,
           "loss": loss(noise), "eps": EPS, "h": H, "noise_fd": [], "noise_fd_check": None}
    print(f"loss {out['loss']:.10e} ({time.time() - t0:.0f} s)", flush=True)
    for j in range(NDIR):
        d = np.random.default_rng(SEED_DIR + j).standard_normal(noise.shape)
        fd = (loss(noise + EPS * d) - loss(noise - EPS * d)) / (2 * EPS)
        out["noise_fd"].append(fd)
        print(f"direction {j}: fd {fd:.10e} ({time.time() - t0:.0f} s)", flush=True)
        if j == 0:
            fd2 = (loss(noise + EPS_CHECK * d) - loss(noise - EPS_CHECK * d)) / (2 * EPS_CHECK)
            out["noise_fd_check"] = {"eps": EPS_CHECK, "fd": fd2}
            print(f"direction 0 at eps {EPS_CHECK}: fd {fd2:.10e}", flush=True)
    for name, key in (("Omega_c", "Oc"), ("sigma8", "s8")):
        v = P[name]
        fd = (loss(noise, **{key: v + H}) - loss(noise, **{key: v - H})) / (2 * H)
        out[name + "_fd"] = fd
        print(f"{name}: fd {fd:.10e} ({time.time() - t0:.0f} s)", flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
