"""Adjoint (VJP) entry points against the float64 oracle.

Cotangent vectors are compared with the oracle's float64 VJPs (oracle/pm_oracle.py, themselves
pinned by float64 finite differences in tests/test_oracle.py) by relative L2 error, tolerance 1e-5
(1e-4 for the force Jacobian, where the readout- and paint-weight derivative terms are differences
of neighbouring cells and cancel to a few float32 ulps). The coefficient cotangents and the state
of the reverse step are checked by float64 central differences of the oracle step. Linear parts
are also checked by exact transpose identities <A u, w> = <u, A^T w>.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import pm_oracle as orc

pytestmark = pytest.mark.gpu

EPS = 1e-5


def _disp(shape, scale, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(tuple(shape) + (3,)) * scale).astype(np.float32)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


def _rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dd(fn, x, d):
    """Central difference of scalar fn at float64 x along d."""
    x = x.astype(np.float64)
    return (fn(x + EPS * d) - fn(x - EPS * d)) / (2 * EPS)


@pytest.mark.parametrize("mode", ["relative", "absolute", "relative-sublattice"])
def test_paint_vjp_directional(pkg, cuda, mode):
    n = 16
    mesh = (n, n, n)
    if mode == "relative-sublattice":
        pos = _disp((8, 8, 8), 1.3, 31)
        mode = "relative"
    elif mode == "relative":
        pos = _disp(mesh, 1.3, 31)
    else:
        pos = (np.random.default_rng(31).random((n, n, n, 3)) * n).astype(np.float32)
    rng = np.random.default_rng(32)
    mbar = rng.standard_normal(mesh).astype(np.float32)
    pb, wb = pkg.engine.paint_vjp(torch.from_numpy(pos).to(cuda), torch.from_numpy(mbar).to(cuda), mode=mode,
                                  weights=0.7, want_weights=True)
    ref_pb, ref_wb = orc.cic_paint_vjp(pos, mbar, mode, weight=0.7)
    assert _rel_l2(pb.cpu().numpy(), ref_pb) < 1e-5
    assert _rel_l2(wb.cpu().numpy(), ref_wb) < 1e-5


@pytest.mark.parametrize("ncomp", [1, 3])
def test_readout_vjp_transpose_and_directional(pkg, cuda, ncomp):
    n = 16
    mesh = (n, n, n)
    pos = _disp(mesh, 1.1, 33)
    rng = np.random.default_rng(34)
    m = rng.standard_normal(mesh + ((3,) if ncomp == 3 else ())).astype(np.float32)
    obar = rng.standard_normal((n, n, n) + ((3,) if ncomp == 3 else ())).astype(np.float32)
    md, pd, od = (torch.from_numpy(a).to(cuda) for a in (m, pos, obar))
    mbar, pbar = pkg.engine.readout_vjp(md, pd, od)
    # transpose: <readout(m), obar> = <m, mbar>
    lhs = float((pkg.engine.readout(md, pd).double() * od.double()).sum().cpu())
    rhs = float((md.double() * mbar.double()).sum().cpu())
    assert _rel(rhs, lhs) < 1e-5
    ref_mb, ref_pb = orc.cic_read_vjp(m, pos, obar, "relative")
    assert _rel_l2(mbar.cpu().numpy(), ref_mb) < 1e-5
    assert _rel_l2(pbar.cpu().numpy(), ref_pb) < 1e-5


def test_force_mesh_vjp_transpose(pkg, cuda):
    n = 32
    rng = np.random.default_rng(35)
    delta = torch.from_numpy(rng.standard_normal((n, n, n)).astype(np.float32)).to(cuda)
    fbar = torch.from_numpy(rng.standard_normal((n, n, n, 3)).astype(np.float32)).to(cuda)
    for deconv in (False, True):
        fm = pkg.engine.force_mesh(delta, scale=0.8, deconv=deconv)
        db = pkg.engine.force_mesh_vjp(fbar, scale=0.8, deconv=deconv)
        lhs = float((fm.double() * fbar.double()).sum().cpu())
        rhs = float((delta.double() * db.double()).sum().cpu())
        assert _rel(rhs, lhs) < 1e-5


@pytest.mark.parametrize("mode", ["relative", "absolute"])
def test_forces_vjp_directional(pkg, cuda, mode):
    n = 32
    mesh = (n, n, n)
    if mode == "relative":
        pos = _disp(mesh, 1.5, 36)
    else:
        pos = (np.random.default_rng(36).random((16, 16, 16, 3)) * n).astype(np.float32)
    rng = np.random.default_rng(37)
    fbar = rng.standard_normal(pos.shape).astype(np.float32)
    dot = torch.zeros((), dtype=torch.float64, device=cuda)
    pb = pkg.engine.forces_vjp(torch.from_numpy(pos).to(cuda), mesh, torch.from_numpy(fbar).to(cuda), mode=mode,
                               scale=0.45, dot=dot)
    ref = orc.pm_forces_vjp(pos, mesh, fbar, mode, scale=0.45)
    err = _rel_l2(pb.cpu().numpy(), ref)
    print(f"forces vjp {mode}: rel L2 {err:.2e}")
    assert err < 1e-4
    ref_dot = float((orc.pm_forces(pos, mesh, mode, scale=0.45, dtype=np.float64) * fbar).sum())
    assert _rel(float(dot.cpu()), ref_dot) < 1e-5


def test_forces_vjp_staged_and_global_blocks(pkg, cuda):
    """The adjoint gather on the readout's 4x8x8 particle blocks: a smooth displacement field (blocks
    whose corner box is staged in LDS) with one slab of widely scattered particles (blocks that gather
    from global memory), against the float64 oracle's forces VJP and dot."""
    n = 32
    mesh = (n, n, n)
    rng = np.random.default_rng(43)
    ii = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).astype(np.float64)
    dx = 1.5 * np.sin(2 * np.pi * ii[..., ::-1] / n + np.arange(3)) + 0.2 * rng.standard_normal(ii.shape)
    dx[8:12] = rng.standard_normal((4, n, n, 3)) * 7.0
    pos = dx.astype(np.float32)
    fbar = rng.standard_normal(pos.shape).astype(np.float32)
    dot = torch.zeros((), dtype=torch.float64, device=cuda)
    pb = pkg.engine.forces_vjp(torch.from_numpy(pos).to(cuda), mesh, torch.from_numpy(fbar).to(cuda),
                               mode="relative", scale=0.45, dot=dot)
    ref = orc.pm_forces_vjp(pos, mesh, fbar, "relative", scale=0.45)
    err = _rel_l2(pb.cpu().numpy(), ref)
    print(f"forces vjp staged + global blocks: rel L2 {err:.2e}")
    assert err < 1e-4
    ref_dot = float((orc.pm_forces(pos, mesh, "relative", scale=0.45, dtype=np.float64) * fbar).sum())
    assert _rel(float(dot.cpu()), ref_dot) < 1e-5


@pytest.mark.parametrize("saved", [False, True])
def test_forces_vjp_bar_scale_and_accumulate(pkg, cuda, saved):
    """bar_scale folds a kick coefficient into the pull-back (pos_bar += c J^T u) and leaves the dot
    <u, F> unscaled; accumulate adds to pos_bar."""
    n = 32
    mesh = (n, n, n)
    pos = torch.from_numpy(_disp(mesh, 1.5, 38)).to(cuda)
    u = torch.from_numpy(np.random.default_rng(39).standard_normal(tuple(pos.shape)).astype(np.float32)).to(cuda)
    fm = None
    if saved:
        fm = torch.empty(mesh + (3,), dtype=torch.float32, device=cuda)
        xs, vs = pos.clone(), torch.zeros_like(pos)
        pkg.engine.step(xs, vs, mesh, mode="relative", force_scale=0.45, c_kick=1.0, force_mesh=fm)
    d1 = torch.zeros(1, dtype=torch.float64, device=cuda)
    d2 = torch.zeros(1, dtype=torch.float64, device=cuda)
    g1 = pkg.engine.forces_vjp(pos, mesh, u, scale=0.45, dot=d1, force_mesh=fm)
    base = torch.from_numpy(np.random.default_rng(40).standard_normal(tuple(pos.shape)).astype(np.float32)).to(cuda)
    acc = base.clone()
    pkg.engine.forces_vjp(pos, mesh, u, scale=0.45, dot=d2, force_mesh=fm, pos_bar=acc, accumulate=True,
                          bar_scale=0.37)
    expect = base.double() + 0.37 * g1.double()
    assert _rel_l2(acc.cpu().numpy(), expect.cpu().numpy()) < 1e-5
    assert _rel(float(d2.cpu()), float(d1.cpu())) < 1e-6


def test_axpy_and_adjoint_drift(pkg, cuda):
    rng = np.random.default_rng(44)
    x, v, xb, vb = (rng.standard_normal((16, 16, 16, 3)).astype(np.float32) for _ in range(4))
    out = pkg.engine.axpy(torch.from_numpy(x).to(cuda), torch.from_numpy(v).to(cuda), 0.3).cpu().numpy()
    ref = x.astype(np.float64) + 0.3 * v.astype(np.float64)
    assert np.abs(out - ref).max() <= 2e-7 * np.abs(ref).max()
    vbt = torch.from_numpy(vb).to(cuda)
    dot = torch.full((1,), 1.5, dtype=torch.float64, device=cuda)
    pkg.engine.adjoint_drift(torch.from_numpy(xb).to(cuda), vbt, torch.from_numpy(v).to(cuda), -0.7, dot, (16, 16, 16))
    ref_vb = vb.astype(np.float64) - 0.7 * xb.astype(np.float64)
    assert np.abs(vbt.cpu().numpy() - ref_vb).max() <= 2e-7 * np.abs(ref_vb).max()
    ref_dot = 1.5 + float((xb.astype(np.float64) * v.astype(np.float64)).sum())
    assert _rel(float(dot.cpu()), ref_dot) < 1e-12


@pytest.mark.parametrize("npart", [32, 16])
def test_step_reverse_vjp(pkg, cuda, npart):
    """One KKD step (v += F ck; x += v cd): the reverse restores the state and the cotangents match
    the float64 directional derivative of L = <xb, x'> + <vb, v'>; coefficient cotangents too. npart
    16: particles on the sub-lattice two cells apart (the 512^3-on-1024^3 set-up)."""
    n = 32
    mesh = (n, n, n)
    x = _disp((npart,) * 3, 1.2, 38)
    v = _disp((npart,) * 3, 0.4, 39)
    fs, ck, cd = 0.45, 0.7, 0.3
    rng = np.random.default_rng(40)
    xb = rng.standard_normal(x.shape).astype(np.float32)
    vb = rng.standard_normal(x.shape).astype(np.float32)

    def step64(x_, v_, ck_=ck, cd_=cd):
        f = orc.pm_forces(x_, mesh, "relative", scale=fs, dtype=np.float64)
        v2 = v_ + ck_ * f
        return x_ + cd_ * v2, v2

    x1, v1 = step64(x.astype(np.float64), v.astype(np.float64))
    xd = torch.from_numpy(x1.astype(np.float32)).to(cuda)
    vd = torch.from_numpy(v1.astype(np.float32)).to(cuda)
    xbd, vbd = torch.from_numpy(xb).to(cuda), torch.from_numpy(vb).to(cuda)
    cb = pkg.engine.step_reverse_vjp(xd, vd, xbd, vbd, mesh, force_scale=fs, c_kick=ck, c_drift=cd)
    assert np.abs(xd.cpu().numpy() - x).max() < 1e-4
    assert np.abs(vd.cpu().numpy() - v).max() < 1e-4

    def L(x_, v_, ck_=ck, cd_=cd):
        a, b = step64(x_, v_, ck_, cd_)
        return (a * xb).sum() + (b * vb).sum()

    # v_bar = vb + cd xb ; x_bar = xb + J(x)^T (ck v_bar)
    vmid = vb.astype(np.float64) + cd * xb
    ref_xb = xb + orc.pm_forces_vjp(x, mesh, ck * vmid, "relative", scale=fs)
    err_x = _rel_l2(xbd.cpu().numpy(), ref_xb)
    print(f"step reverse vjp: x_bar rel L2 {err_x:.2e}")
    assert _rel_l2(vbd.cpu().numpy(), vmid) < 1e-6
    assert err_x < 1e-4
    # the composed oracle against a float64 central difference of the whole step (1e-5: at eps = 1e-5
    # about one of the 10^5 coordinates crosses a cell boundary inside the difference)
    dx = rng.standard_normal(x.shape)
    fd = _dd(lambda q: L(q, v.astype(np.float64)), x, dx)
    assert _rel(float((ref_xb * dx).sum()), float(fd)) < 1e-5
    h = 1e-6
    ref_cd = (L(x.astype(np.float64), v.astype(np.float64), cd_=cd + h) -
              L(x.astype(np.float64), v.astype(np.float64), cd_=cd - h)) / (2 * h)
    ref_ck = (L(x.astype(np.float64), v.astype(np.float64), ck_=ck + h) -
              L(x.astype(np.float64), v.astype(np.float64), ck_=ck - h)) / (2 * h)
    c = cb.cpu().numpy()
    assert _rel(float(c[0]), ref_cd) < 1e-5
    assert _rel(float(c[1]), ref_ck) < 1e-5


def test_step_keep_and_reverse_with_the_kept_force_mesh(pkg, cuda):
    """The forward step that keeps its force mesh moves the state exactly as the plain step, and the
    reverse step that reuses that mesh restores the state and the cotangents as the one that
    recomputes it (the restored positions differ from the forward's by float32 rounding only)."""
    n = 32
    mesh = (n, n, n)
    x = torch.from_numpy(_disp(mesh, 1.2, 41)).to(cuda)
    v = torch.from_numpy(_disp(mesh, 0.4, 42)).to(cuda)
    fs, ck, cd = 0.45, 0.7, 0.3
    M = torch.empty(mesh + (3,), dtype=torch.float32, device=cuda)
    xa, va = x.clone(), v.clone()
    pkg.engine.step(xa, va, mesh, force_scale=fs, c_kick=ck, c_drift_post=cd, force_mesh=M)
    xb_, vb_ = x.clone(), v.clone()
    pkg.engine.step(xb_, vb_, mesh, force_scale=fs, c_kick=ck, c_drift_post=cd)
    assert torch.equal(xa, xb_) and torch.equal(va, vb_)
    F = pkg.engine.readout(M, x)  # the kept mesh read at the step's positions is the kick's force
    assert _rel_l2(((va - v) / ck).cpu().numpy(), F.cpu().numpy()) < 1e-5
    rng = np.random.default_rng(43)
    xbar = torch.from_numpy(rng.standard_normal(mesh + (3,)).astype(np.float32)).to(cuda)
    vbar = torch.from_numpy(rng.standard_normal(mesh + (3,)).astype(np.float32)).to(cuda)
    out = []
    for keep in (True, False):
        xs, vs, xbs, vbs = xa.clone(), va.clone(), xbar.clone(), vbar.clone()
        cb = pkg.engine.step_reverse_vjp(xs, vs, xbs, vbs, mesh, force_scale=fs, c_kick=ck, c_drift=cd,
                                         force_mesh=M if keep else None)
        out.append((xs, vs, xbs, vbs, cb))
    (xk, vk, xbk, vbk, cbk), (xr, vr, xbr, vbr, cbr) = out
    assert (xk - x).abs().max() < 1e-4 and (vk - v).abs().max() < 1e-4
    assert torch.equal(xk, xr) and torch.equal(vbk, vbr)
    assert _rel_l2(vk.cpu().numpy(), vr.cpu().numpy()) < 1e-6
    assert _rel_l2(xbk.cpu().numpy(), xbr.cpu().numpy()) < 1e-5
    assert np.allclose(cbk.cpu().numpy(), cbr.cpu().numpy(), rtol=1e-5)
