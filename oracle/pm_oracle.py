"""CPU oracle for the PM force + integration path.

TEST INFRASTRUCTURE ONLY. This module is a NumPy restatement of the reference algorithm, used
by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker (and the
timed CPU baseline). The product (jax-fli_amd/) never imports it.

Each function cites the reference file:line it restates. The PM arithmetic of the reference lives
in the un-vendored ``jaxpm`` fork (pyproject.toml:32, branch ALL-MODIF, no commit hash); its
published algorithm (FastPM, Feng et al. 2016; pmwd's multilinear enmesh for ``cic_paint_dx``)
is restated here and anchored on the reference's call sites. How each function is pinned (closed
forms, invariants, brute force, known-answer vectors) is listed in tests/test_oracle.py.

Parity status per function:
  cic_paint_rel / cic_paint_abs / cic_read_*   pinned (mass, zero displacement, roll, adjointness,
                                               single-particle closed form, brute force)
  fftk / gradient / invlaplace / pm_forces     pinned (plane wave, momentum conservation,
                                               two-particle antisymmetry, brute-force DFT)
  power                                        pinned (plane wave, exact restatement of _power)
  philox4x32                                   pinned (Random123 known-answer vectors)
  healpix_ang2pix / healpix_interp_weights     pinned (pixel centres round-trip, nside=1 base
                                               pixels, equal-area counts, interpolation identities)
  paint_healpix normalisation / shell edges    parity unpinned (jaxpm.spherical is not in the
                                               reference tree and no reference test covers it)
  pgd_range / pgd_displacement                 pinned (kl = 0, ks -> inf gives the force; a single
                                               plane wave is scaled by the filter at its k)
  cic_paint_vjp / cic_read_vjp /               pinned (float64 central differences along random
  force_mesh_vjp / pm_forces_vjp                 directions; exact transpose identities; also with the
                                               PGD filter against pgd_displacement)
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

# The per-particle loops (CIC paint/read, HEALPix deposits) also exist in C (oracle/oracle_c.c, the same
# operations in the same order; built by __graft_entry__.build() into oracle/liboracle_c.so). They are
# used when that library is present and USE_C is true; tests/test_oracle.py checks both agree.
USE_C = True
_C = None


def _clib():
    global _C
    if _C is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle_c.so")
        if not os.path.exists(path):
            _C = False
        else:
            lib = ctypes.CDLL(path)
            i64, vp, d, f = ctypes.c_int64, ctypes.c_void_p, ctypes.c_double, ctypes.c_float
            for name, wt in (("cic_paint_f64", d), ("cic_paint_f32", f)):
                fn = getattr(lib, name)
                fn.argtypes = [vp, i64, i64, i64, ctypes.c_int, i64, i64, i64, vp, wt, vp]
                fn.restype = ctypes.c_int
            for name in ("cic_read_f64", "cic_read_f32"):
                fn = getattr(lib, name)
                fn.argtypes = [vp, ctypes.c_int, vp, i64, i64, i64, ctypes.c_int, i64, i64, i64, vp]
                fn.restype = ctypes.c_int
            lib.cic_paint_mt.argtypes = [vp, i64, i64, i64, i64, i64, i64, vp, ctypes.c_int]
            lib.cic_paint_mt.restype = ctypes.c_int
            lib.cic_read3_mt.argtypes = [vp, vp, i64, i64, i64, i64, i64, i64, vp, ctypes.c_int]
            lib.cic_read3_mt.restype = ctypes.c_int
            lib.healpix_deposit.argtypes = [i64, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, i64, vp]
            lib.healpix_deposit.restype = ctypes.c_int
            _C = lib
    return _C if (USE_C and _C) else None


def _cptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _workers() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def fftn(a, **kw):
    """Complex 3D FFT (pocketfft through scipy.fft, threaded over the host cores)."""
    import scipy.fft

    return scipy.fft.fftn(a, workers=_workers(), **kw)


def ifftn(a, **kw):
    import scipy.fft

    return scipy.fft.ifftn(a, workers=_workers(), **kw)

# ---------------------------------------------------------------------------- CIC
# jaxpm.painting.cic_paint_dx (called at src/jax_fli/_src/fields/_painting.py:67-73): particle
# (i,j,k) at lattice (i,j,k)*r + dx, corners floor(dx) + {0,1}, weights 1 - |frac - {0,1}|,
# periodic wrap. jaxpm.painting.cic_paint (_painting.py:76-82): weights 1 - |x - (floor(x)+{0,1})|.


def _lattice(np_shape, mesh_shape):
    r = [m // n for m, n in zip(mesh_shape, np_shape)]
    grids = np.meshgrid(*[np.arange(n) * ri for n, ri in zip(np_shape, r)], indexing="ij")
    return np.stack(grids, axis=-1).reshape(-1, 3).astype(np.int64)


def _corners(pos, mesh_shape, mode, np_shape, dtype):
    """Per-dimension base indices and the two weights, in `dtype` arithmetic."""
    x = np.asarray(pos, dtype=dtype).reshape(-1, 3)
    fl = np.floor(x)
    one = dtype(1.0)
    if mode == "relative":
        fr = x - fl
        w0 = one - np.abs(fr)
        w1 = one - np.abs(fr - one)
        base = _lattice(np_shape, mesh_shape) + fl.astype(np.int64)
    else:
        w0 = one - np.abs(x - fl)
        w1 = one - np.abs(x - (fl + one))
        base = fl.astype(np.int64)
    return base, w0, w1


def _offsets():
    # x fastest: (0,0,0),(1,0,0),(0,1,0),(1,1,0),(0,0,1),...
    return [((o >> 0) & 1, (o >> 1) & 1, (o >> 2) & 1) for o in range(8)]


def cic_paint(pos, mesh_shape, mode="relative", weight=1.0, dtype=np.float32):
    """CIC scatter; accumulation in float64, result cast to `dtype`."""
    pos = np.asarray(pos)
    np_shape = pos.shape[:3]
    lib = _clib()
    if lib is not None and dtype in (np.float32, np.float64):
        x = np.ascontiguousarray(pos, dtype=dtype).reshape(-1, 3)
        wa = np.asarray(weight, dtype=dtype).reshape(-1)
        wp = None if wa.size == 1 else np.ascontiguousarray(np.broadcast_to(wa, (x.shape[0],)))
        mesh = np.empty(int(np.prod(mesh_shape)), dtype=np.float64)
        fn = lib.cic_paint_f64 if dtype == np.float64 else lib.cic_paint_f32
        rc = fn(_cptr(x), *[int(n) for n in np_shape], 0 if mode == "relative" else 1, *[int(n) for n in mesh_shape],
                None if wp is None else _cptr(wp), float(wa[0]) if wp is None else 0.0, _cptr(mesh))
        assert rc == 0
        return mesh.reshape(mesh_shape).astype(dtype)
    base, w0, w1 = _corners(pos, mesh_shape, mode, np_shape, dtype)
    wt = np.broadcast_to(np.asarray(weight, dtype=dtype).reshape(-1), (base.shape[0],))
    mesh = np.zeros(int(np.prod(mesh_shape)), dtype=np.float64)
    nx, ny, nz = mesh_shape
    for ox, oy, oz in _offsets():
        wx = w1[:, 0] if ox else w0[:, 0]
        wy = w1[:, 1] if oy else w0[:, 1]
        wz = w1[:, 2] if oz else w0[:, 2]
        w = ((wx * wy) * wz) * wt
        ix = np.mod(base[:, 0] + ox, nx)
        iy = np.mod(base[:, 1] + oy, ny)
        iz = np.mod(base[:, 2] + oz, nz)
        flat = (ix * ny + iy) * nz + iz
        mesh += np.bincount(flat, weights=w.astype(np.float64), minlength=mesh.size)
    return mesh.reshape(mesh_shape).astype(dtype)


def cic_read(mesh, pos, mode="relative", dtype=np.float32):
    """CIC gather of a (X,Y,Z) or (X,Y,Z,C) mesh; corners summed in x-fastest order."""
    mesh = np.asarray(mesh, dtype=dtype)
    mesh_shape = mesh.shape[:3]
    ncomp = 1 if mesh.ndim == 3 else mesh.shape[3]
    m = mesh.reshape(-1, ncomp)
    pos = np.asarray(pos)
    np_shape = pos.shape[:3]
    lib = _clib()
    if lib is not None and dtype in (np.float32, np.float64) and ncomp <= 8:
        x = np.ascontiguousarray(pos, dtype=dtype).reshape(-1, 3)
        mc = np.ascontiguousarray(m)
        out = np.empty((x.shape[0], ncomp), dtype=dtype)
        fn = lib.cic_read_f64 if dtype == np.float64 else lib.cic_read_f32
        rc = fn(_cptr(mc), int(ncomp), _cptr(x), *[int(n) for n in np_shape], 0 if mode == "relative" else 1,
                *[int(n) for n in mesh_shape], _cptr(out))
        assert rc == 0
        out = out.reshape(tuple(np_shape) + (ncomp,))
        return out[..., 0] if mesh.ndim == 3 else out
    base, w0, w1 = _corners(pos, mesh_shape, mode, np_shape, dtype)
    nx, ny, nz = mesh_shape
    out = np.zeros((base.shape[0], ncomp), dtype=dtype)
    for ox, oy, oz in _offsets():
        wx = w1[:, 0] if ox else w0[:, 0]
        wy = w1[:, 1] if oy else w0[:, 1]
        wz = w1[:, 2] if oz else w0[:, 2]
        w = (wx * wy) * wz
        ix = np.mod(base[:, 0] + ox, nx)
        iy = np.mod(base[:, 1] + oy, ny)
        iz = np.mod(base[:, 2] + oz, nz)
        out = out + m[(ix * ny + iy) * nz + iz] * w[:, None]
    out = out.reshape(tuple(np_shape) + (ncomp,))
    return out[..., 0] if mesh.ndim == 3 else out


# ---------------------------------------------------------------------------- k-space kernels
# jaxpm.kernels.fftk / gradient_kernel / invlaplace_kernel as used by pm_forces (called from
# _forces, src/jax_fli/pm/solvers.py:19-32): k = 2 pi fftfreq(n) in cell units, FastPM 4-point
# gradient i (8 sin k - sin 2k)/6, Green's function -1/k^2 with 0 at k = 0, no CIC compensation.


def fftk(mesh_shape, dtype=np.float64):
    ks = []
    for d, n in enumerate(mesh_shape):
        k = (2.0 * np.pi * np.fft.fftfreq(n)).astype(dtype)
        shape = [1, 1, 1]
        shape[d] = n
        ks.append(k.reshape(shape))
    return ks


def gradient_kernel(kvec, direction):
    w = kvec[direction]
    return 1j * (8.0 * np.sin(w) - np.sin(2.0 * w)) / 6.0


def invlaplace_kernel(kvec):
    kk = sum(k**2 for k in kvec)
    kk_nz = np.where(kk == 0, 1.0, kk)
    return -np.where(kk == 0, 0.0, 1.0 / kk_nz)


def cic_window_comp(kvec):
    """1 / prod_d sinc(k_d/2)^4 (compensation of paint + readout; off in the reference)."""
    out = 1.0
    for k in kvec:
        s = np.sinc(k / (2.0 * np.pi))
        out = out / s**4
    return out


def pgd_range(kvec, kl, ks):
    """PGD filter exp(-kl^2/k^2) exp(-k^4/ks^4), 0 at k = 0 (src/jax_fli/pm/correction.py:182-190)."""
    kk = sum(k**2 for k in kvec)
    kk_safe = np.where(kk == 0, 1.0, kk)
    r = np.exp(-(kl**2) / kk_safe) * np.exp(-(kk**2) / ks**4)
    return np.where(kk == 0, 0.0, r)


def force_mesh(delta, deconv=False, dtype=np.float64, pgd=None):
    """(X,Y,Z,3) force mesh of a density mesh: ifft(-grad_i * invlaplace * fft(delta)).real.
    The Nyquist gradient is exactly 0 (sin(pi) = 0). pgd = (kl, ks) multiplies the potential by
    the PGD filter."""
    mesh_shape = delta.shape
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    dk = fftn(np.asarray(delta, dtype=dtype)).astype(cdt)
    kvec = fftk(mesh_shape, np.float64)
    pot = dk * invlaplace_kernel(kvec)
    if pgd is not None:
        pot = pot * pgd_range(kvec, *pgd)
    if deconv:
        pot = pot * cic_window_comp(kvec)
    out = np.empty(tuple(mesh_shape) + (3,), dtype=dtype)
    for i in range(3):
        g = gradient_kernel(kvec, i)
        n = mesh_shape[i]
        if n % 2 == 0:
            sl = [slice(None)] * 3
            sl[i] = n // 2
            g = np.array(np.broadcast_to(g, kvec[i].shape))
            g[tuple(sl)] = 0.0
        out[..., i] = ifftn(-g * pot).real.astype(dtype)
    return out


def hessian(delta):
    """H_ij = ifft(grad_i grad_j invlaplace fft(delta)) = ifft(g_i g_j delta_k / k^2), the shear of the
    potential with the FastPM gradient kernel (upstream JaxPM lpt, order 2). Returns {(i, j): mesh}."""
    mesh_shape = delta.shape
    dk = fftn(np.asarray(delta, dtype=np.float64))
    kvec = fftk(mesh_shape, np.float64)
    pot = dk * invlaplace_kernel(kvec)
    g = [gradient_kernel(kvec, i) for i in range(3)]
    return {(i, j): ifftn(g[i] * g[j] * pot).real for i in range(3) for j in range(i, 3)}


def lpt2_source(delta):
    """delta2 = sum_{i<j} (H_ii H_jj - H_ij^2), accumulated as the upstream loop does: per i, add
    H_ii times the running sum of the earlier diagonals, subtract the squares H_ij^2 for j > i."""
    H = hessian(delta)
    d2 = np.zeros(delta.shape)
    acc = np.zeros(delta.shape)
    for i in range(3):
        d2 = d2 + H[i, i] * acc
        acc = acc + H[i, i]
        for j in range(i + 1, 3):
            d2 = d2 - H[i, j] ** 2
    return d2


def pm_forces(pos, mesh_shape, mode="relative", scale=1.0, deconv=False, dtype=np.float64):
    """jaxpm.pm.pm_forces(pos) * scale (scale = 1.5 Omega_m in _forces, solvers.py:29-30)."""
    rho = cic_paint(pos, mesh_shape, mode, dtype=dtype)
    fm = force_mesh(rho, deconv=deconv, dtype=dtype) * dtype(scale)
    return cic_read(fm.astype(dtype), pos, mode, dtype=dtype)


def pgd_displacement(pos, mesh_shape, alpha, kl, ks, mode="relative", dtype=np.float64):
    """S(x) = alpha * read(x, force mesh of paint(x) with the PGD filter): the displacement of
    PGDKernel.apply and, divided by the drift factor, the velocity boost of SharpeningKernel
    (src/jax_fli/pm/correction.py:169-200, 301-333)."""
    rho = cic_paint(pos, mesh_shape, mode, dtype=dtype)
    fm = force_mesh(rho, dtype=dtype, pgd=(kl, ks))
    return cic_read(fm.astype(dtype), pos, mode, dtype=dtype) * dtype(alpha)


# ---------------------------------------------------------------------------- P(k)
# _power / _initialize_pk (src/jax_fli/_src/power/_compute.py:23-119), auto spectrum, monopole.


def legendre(mu, ell):
    """P_ell(mu) by the three-term recurrence (src/jax_fli/_src/power/_compute.py:8-20)."""
    mu = np.asarray(mu, dtype=np.float64)
    if ell == 0:
        return np.ones_like(mu)
    p0, p1 = np.ones_like(mu), mu
    for n in range(2, ell + 1):
        p0, p1 = p1, ((2 * n - 1) * mu * p1 - (n - 1) * p0) / n
    return p1


def power(mesh, box_shape, kedges=None, mesh2=None, multipoles=0, los=(0.0, 0.0, 1.0)):
    """_power (src/jax_fli/_src/power/_compute.py:23-119) in float64 over the full c2c grid: bins by
    digitize(|k|, kedges); auto |m|^2 or cross m1 conj(m2) (ortho FFTs), weighted by (2l+1) P_l(mu),
    mu = k.los/|k| (0 at k = 0; no weighting when multipoles == 0); the cross sum per bin is
    sqrt(re^2 + im^2); P = sum / count * prod(box / mesh), under- and overflow bins dropped. One pole:
    (kavg, pk); several: (kavg, pk of shape (npoles, nbins))."""
    mesh = np.asarray(mesh, dtype=np.float64)
    mesh_shape = mesh.shape
    box = np.asarray(box_shape, dtype=np.float64)
    ms = np.asarray(mesh_shape)
    poles = tuple(multipoles) if isinstance(multipoles, (list, tuple)) else (multipoles,)
    use_los = not (isinstance(multipoles, int) and multipoles == 0)
    if kedges is None:
        kmax = np.pi * np.min(ms / box)
        dk = 2 * np.pi / np.min(box) * 2
        kedges = np.arange(dk, kmax, dk) + dk / 2
        if kedges.size < 2:
            kedges = np.linspace(kmax / 4, kmax * 0.9, 2)
    kedges = np.asarray(kedges, dtype=np.float64)
    kvec = []
    for i, (m, b) in enumerate(zip(mesh_shape, box)):
        shape = [1, 1, 1]
        shape[i] = m
        kvec.append(((2 * np.pi * m / b) * np.fft.fftfreq(m)).reshape(shape))
    kmesh = np.broadcast_to(np.sqrt(sum(k**2 for k in kvec)), mesh_shape)
    dig = np.digitize(kmesh.reshape(-1), kedges)
    nb = kedges.size + 1
    kcount = np.bincount(dig, minlength=nb)
    kavg = np.bincount(dig, weights=kmesh.reshape(-1), minlength=nb)
    kavg = (kavg / np.where(kcount == 0, 1, kcount))[1:-1]
    if use_los:
        u = np.asarray(los, dtype=np.float64)
        u = u / np.linalg.norm(u)
        mu = np.broadcast_to(sum(k * c for k, c in zip(kvec, u)), mesh_shape)
        mu = np.where(kmesh == 0, 0.0, mu / np.where(kmesh == 0, 1.0, kmesh))
    else:
        mu = np.ones(mesh_shape)
    meshk = fftn(mesh, norm="ortho")
    if mesh2 is None:
        mmk = meshk.real**2 + meshk.imag**2
    else:
        mmk = meshk * np.conj(fftn(np.asarray(mesh2, dtype=np.float64), norm="ortho"))
    out = []
    for ell in poles:
        w = (mmk * (2 * int(ell) + 1) * legendre(mu, int(ell))).reshape(-1)
        if mesh2 is None:
            psum = np.bincount(dig, weights=w, minlength=nb)
        else:
            re = np.bincount(dig, weights=w.real, minlength=nb)
            im = np.bincount(dig, weights=w.imag, minlength=nb)
            psum = np.sqrt(re**2 + im**2)
        out.append(psum)
    pk = (np.stack(out) / np.where(kcount > 0, kcount, 1))[:, 1:-1] * np.prod(box / ms)
    return (kavg, pk[0]) if len(poles) == 1 else (kavg, pk)


def transfer(mesh0, mesh1, box_shape, kedges=None):
    """sqrt(P1 / P0), monopoles (src/jax_fli/_src/power/_compute.py:240-244)."""
    k, p0 = power(mesh0, box_shape, kedges)
    _, p1 = power(mesh1, box_shape, kedges)
    return k, (p1 / p0) ** 0.5


def coherence(mesh0, mesh1, box_shape, kedges=None):
    """P01 / sqrt(P0 P1), monopoles (src/jax_fli/_src/power/_compute.py:247-252)."""
    k, p01 = power(mesh0, box_shape, kedges, mesh2=mesh1)
    _, p0 = power(mesh0, box_shape, kedges)
    _, p1 = power(mesh1, box_shape, kedges)
    return k, p01 / (p0 * p1) ** 0.5


# ---------------------------------------------------------------------------- Philox4x32-10
# Salmon, Moraes, Dror, Shaw (SC 2011), Random123. Used for the white-noise input.

_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def philox4x32(counters, key):
    """counters: (n, 4) uint32; key: (k0, k1). Returns (n, 4) uint32."""
    c = np.asarray(counters, dtype=np.uint64).copy()
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(_M0) * c[:, 0]
        p1 = np.uint64(_M1) * c[:, 2]
        n0 = (p1 >> np.uint64(32)) ^ c[:, 1] ^ k0
        n1 = p1 & mask
        n2 = (p0 >> np.uint64(32)) ^ c[:, 3] ^ k1
        n3 = p0 & mask
        c = np.stack([n0, n1, n2, n3], axis=1) & mask
        k0 = (k0 + np.uint64(_W0)) & mask
        k1 = (k1 + np.uint64(_W1)) & mask
    return c.astype(np.uint32)


def white_noise(shape, seed):
    n = int(np.prod(shape))
    nb = (n + 3) // 4
    b = np.arange(nb, dtype=np.uint64)
    ctr = np.stack([b & np.uint64(0xFFFFFFFF), b >> np.uint64(32), np.zeros_like(b), np.zeros_like(b)], axis=1)
    r = philox4x32(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)).astype(np.float64)
    u = (np.floor(r / 256.0) + 1.0) / 16777216.0
    r0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    r1 = np.sqrt(-2.0 * np.log(u[:, 2]))
    g = np.stack(
        [r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
         r1 * np.cos(2 * np.pi * u[:, 3]), r1 * np.sin(2 * np.pi * u[:, 3])], axis=1
    ).reshape(-1)[:n]
    return g.astype(np.float32).reshape(shape)


# ---------------------------------------------------------------------------- HEALPix (RING)
# Textbook HEALPix pixelisation (Gorski et al. 2005), scalar restatement of loc2pix, ring_above,
# get_ring_info2 and get_interpol of the HEALPix C++ library (the algorithm jax_healpy follows).


def healpix_ang2pix(nside, z, phi):
    za = abs(z)
    tt = np.fmod(phi * (2.0 / np.pi), 4.0)
    if tt < 0:
        tt += 4.0
    ncap = 2 * nside * (nside - 1)
    npix = 12 * nside * nside
    if za <= 2.0 / 3.0:
        nl4 = 4 * nside
        t1 = nside * (0.5 + tt)
        t2 = nside * z * 0.75
        jp = int(t1 - t2)
        jm = int(t1 + t2)
        ir = nside + 1 + jp - jm
        kshift = 1 - (ir & 1)
        t = jp + jm - nside + kshift + 1 + nl4 + nl4
        ip = (t >> 1) % nl4
        return ncap + (ir - 1) * nl4 + ip
    tp = tt - int(tt)
    tmp = nside * np.sqrt(3.0 * (1.0 - za))
    jp = int(tp * tmp)
    jm = int((1.0 - tp) * tmp)
    ir = jp + jm + 1
    ip = int(tt * ir) % (4 * ir)
    return 2 * ir * (ir - 1) + ip if z > 0 else npix - 2 * ir * (ir + 1) + ip


def healpix_pix2ang(nside, pix):
    """Pixel centre (theta, phi), RING scheme."""
    ncap = 2 * nside * (nside - 1)
    npix = 12 * nside * nside
    if pix < ncap:
        iring = int((1 + np.sqrt(1 + 2 * pix)) // 2)
        iphi = pix - 2 * iring * (iring - 1)
        z = 1.0 - iring * iring * 4.0 / npix
        phi = (iphi + 0.5) * np.pi / (2.0 * iring)
    elif pix < npix - ncap:
        ip = pix - ncap
        iring = ip // (4 * nside) + nside
        iphi = ip % (4 * nside) + 1
        fodd = 1.0 if ((iring + nside) & 1) else 0.5
        z = (2 * nside - iring) * 2.0 / (3.0 * nside)
        phi = (iphi - fodd) * np.pi / (2.0 * nside)
    else:
        ip = npix - pix
        iring = int((1 + np.sqrt(2 * ip - 1)) // 2)
        iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1))
        z = -1.0 + iring * iring * 4.0 / npix
        phi = (iphi - 0.5) * np.pi / (2.0 * iring)
    return np.arccos(z), phi


def _ring_above(nside, z):
    az = abs(z)
    if az <= 2.0 / 3.0:
        return int(nside * (2.0 - 1.5 * z))
    iring = int(nside * np.sqrt(3.0 * (1.0 - az)))
    return iring if z > 0 else 4 * nside - iring - 1


def _ring_info(nside, ring):
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    fact2 = 4.0 / npix
    fact1 = (2 * nside) * fact2
    northring = 4 * nside - ring if ring > 2 * nside else ring
    if northring < nside:
        tmp = northring * northring * fact2
        theta = np.arctan2(np.sqrt(tmp * (2.0 - tmp)), 1.0 - tmp)
        ringpix = 4 * northring
        shifted = True
        startpix = 2 * northring * (northring - 1)
    else:
        theta = np.arccos((2 * nside - northring) * fact1)
        ringpix = 4 * nside
        shifted = ((northring - nside) & 1) == 0
        startpix = ncap + (northring - nside) * ringpix
    if northring != ring:
        theta = np.pi - theta
        startpix = npix - startpix - ringpix
    return startpix, ringpix, theta, shifted


def healpix_interp_weights(nside, theta, phi):
    """Four RING pixels and bilinear weights (HEALPix get_interpol)."""
    npix = 12 * nside * nside
    z = np.cos(theta)
    ir1 = _ring_above(nside, z)
    ir2 = ir1 + 1
    pix = [0, 0, 0, 0]
    wgt = [0.0, 0.0, 0.0, 0.0]
    theta1 = theta2 = 0.0
    for slot, ir in ((0, ir1), (2, ir2)):
        if (slot == 0 and ir > 0) or (slot == 2 and ir < 4 * nside):
            sp, nr, th, shift = _ring_info(nside, ir)
            if slot == 0:
                theta1 = th
            else:
                theta2 = th
            dphi = 2.0 * np.pi / nr
            tmp = phi / dphi - 0.5 * shift
            i1 = int(tmp) - 1 if tmp < 0 else int(tmp)
            w1 = (phi - (i1 + 0.5 * shift) * dphi) / dphi
            i2 = i1 + 1
            if i1 < 0:
                i1 += nr
            if i2 >= nr:
                i2 -= nr
            pix[slot], pix[slot + 1] = sp + i1, sp + i2
            wgt[slot], wgt[slot + 1] = 1 - w1, w1
    if ir1 == 0:
        wtheta = theta / theta2
        wgt[2] *= wtheta
        wgt[3] *= wtheta
        fac = (1 - wtheta) * 0.25
        wgt[0] = fac
        wgt[1] = fac
        wgt[2] += fac
        wgt[3] += fac
        pix[0] = (pix[2] + 2) & 3
        pix[1] = (pix[3] + 2) & 3
    elif ir2 == 4 * nside:
        wtheta = (theta - theta1) / (np.pi - theta1)
        wgt[0] *= 1 - wtheta
        wgt[1] *= 1 - wtheta
        fac = wtheta * 0.25
        wgt[0] += fac
        wgt[1] += fac
        wgt[2] = fac
        wgt[3] = fac
        pix[2] = ((pix[0] + 2) & 3) + npix - 4
        pix[3] = ((pix[1] + 2) & 3) + npix - 4
    else:
        wtheta = (theta - theta1) / (theta2 - theta1)
        wgt[0] *= 1 - wtheta
        wgt[1] *= 1 - wtheta
        wgt[2] *= wtheta
        wgt[3] *= wtheta
    return pix, wgt


def healpix_deposit(nside, scheme, wrap_guard, X, r, wt, sel, acc) -> bool:
    """The per-particle loop of the shell paints in C (True), or False when the C library is absent and
    the caller runs its own loop."""
    lib = _clib()
    if lib is None or scheme not in ("ngp", "bilinear"):
        return False
    X = np.ascontiguousarray(X, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    wt = np.ascontiguousarray(wt, dtype=np.float64)
    sel = np.ascontiguousarray(sel, dtype=np.int64)
    assert acc.dtype == np.float64 and acc.flags.c_contiguous
    rc = lib.healpix_deposit(int(nside), 0 if scheme == "ngp" else 1, int(wrap_guard), _cptr(X), _cptr(r), _cptr(wt),
                             _cptr(sel), int(sel.size), _cptr(acc))
    assert rc == 0
    return True


def healpix_pix2vec(nside, pix):
    th, ph = healpix_pix2ang(nside, int(pix))
    return np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])


def healpix_rbf(nside, n, sigma):
    """rbf_neighbor: the pixel of unit direction n and those of the eight directions one pixel resolution
    sqrt(4 pi / npix) away (tangent-plane azimuths k pi/4, from the theta and phi unit vectors), each
    weighted by exp(-d^2 / (2 sigma^2)) of the angular distance d = 2 asin(|n - c| / 2) to its centre c,
    normalised to sum 1 (evaluated as exp(-(d^2 - dmin^2) / (2 sigma^2)), which cannot underflow). Returns (pixels, weights), distinct pixels in first-seen order."""
    n = np.asarray(n, dtype=np.float64)
    st = np.hypot(n[0], n[1])
    cp, sp = (n[0] / st, n[1] / st) if st > 0 else (1.0, 0.0)
    eth = np.array([n[2] * cp, n[2] * sp, -st])
    eph = np.array([-sp, cp, 0.0])
    step = np.sqrt(4 * np.pi / (12 * nside * nside))
    pix = []
    for k in range(-1, 8):
        if k < 0:
            u = n
        else:
            a = 0.25 * np.pi * k
            u = np.cos(step) * n + np.sin(step) * (np.cos(a) * eth + np.sin(a) * eph)
        ph = np.arctan2(u[1], u[0])
        ph = ph + 2 * np.pi if ph < 0 else ph
        ph = 0.0 if ph >= 2 * np.pi else ph
        q = healpix_ang2pix(nside, min(1.0, max(-1.0, u[2])), ph)
        if q not in pix:
            pix.append(q)
    d = np.array([2 * np.arcsin(min(1.0, 0.5 * np.linalg.norm(n - healpix_pix2vec(nside, q)))) for q in pix])
    w = np.exp(-0.5 * ((d / sigma) ** 2 - (d.min() / sigma) ** 2))  # relative to the nearest: no underflow
    return np.array(pix), w / w.sum()


def paint_healpix(pos, mesh_shape, box, observer, rmin, rmax, nside, scheme="bilinear", mode="relative", weight=1.0,
                  sigma=None):
    """Shell paint (build-defined normalisation: weighted count / (pixel solid angle x shell
    volume per steradian)); positions wrapped into the periodic box; float64 geometry. rbf_neighbor:
    healpix_rbf with width sigma (radians; default the pixel resolution read as a FWHM)."""
    pos = np.asarray(pos, dtype=np.float32)
    np_shape = pos.shape[:3]
    g = pos.reshape(-1, 3).astype(np.float64)
    if mode == "relative":
        g = g + _lattice(np_shape, mesh_shape).astype(np.float64)
    ms = np.asarray(mesh_shape, dtype=np.float64)
    g = g - np.floor(g / ms) * ms
    rel = g * (np.asarray(box, dtype=np.float64) / ms) - np.asarray(observer, dtype=np.float64)
    r = np.sqrt((rel**2).sum(axis=1))
    sel = np.nonzero((r > rmin) & (r <= rmax))[0]
    wt = np.broadcast_to(np.asarray(weight, dtype=np.float64).reshape(-1), (g.shape[0],))
    npix = 12 * nside * nside
    acc = np.zeros(npix)
    if healpix_deposit(nside, scheme, 0, rel, r, wt, sel, acc):
        sel = ()
    for p in sel:
        X, Y, Z = rel[p]
        phi = np.arctan2(Y, X)
        if phi < 0:
            phi += 2 * np.pi
        if scheme == "ngp":
            acc[healpix_ang2pix(nside, Z / r[p], phi)] += wt[p]
        elif scheme == "rbf_neighbor":
            sg = sigma if sigma else np.sqrt(4 * np.pi / npix) / (2 * np.sqrt(2 * np.log(2)))
            pix, w = healpix_rbf(nside, rel[p] / r[p], sg)
            for q, wq in zip(pix, w):
                acc[q] += wt[p] * wq
        else:
            theta = np.arctan2(np.sqrt(X * X + Y * Y), Z)
            pix, w = healpix_interp_weights(nside, theta, phi)
            for q in range(4):
                acc[pix[q]] += wt[p] * w[q]
    vol = (4 * np.pi / npix) * (rmax**3 - rmin**3) / 3.0
    return (acc / vol).astype(np.float32)


# ---------------------------------------------------------------------------- adjoints
# Reverse-mode derivatives of the functions above, as jax.vjp produces them for the reference's
# cic_paint_dx / cic_read_dx / pm_forces (differentiated inside _integrate_bwd,
# src/jax_fli/pm/integrate.py:304-358): d(1 - |f|)/df = -sign(f), d(1 - |f - 1|)/df = -sign(f - 1).


def _corner_grads(pos, mesh_shape, mode, np_shape):
    """Base indices, weights w0/w1 and their derivatives dw0/dw1 per axis (float64)."""
    x = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    fl = np.floor(x)
    f = x - fl
    w0, w1 = 1.0 - np.abs(f), 1.0 - np.abs(f - 1.0)
    d0, d1 = -np.sign(f), -np.sign(f - 1.0)
    if mode == "relative":
        base = _lattice(np_shape, mesh_shape) + fl.astype(np.int64)
    else:
        base = fl.astype(np.int64)
    return base, (w0, w1), (d0, d1)


def _corner_terms(pos, mesh_shape, mode):
    """Yields (flat cell index, weight, gradient of the weight (N,3)) for the 8 corners."""
    pos = np.asarray(pos)
    base, (w0, w1), (d0, d1) = _corner_grads(pos, mesh_shape, mode, pos.shape[:3])
    nx, ny, nz = mesh_shape
    for ox, oy, oz in _offsets():
        wx, dx = (w1[:, 0], d1[:, 0]) if ox else (w0[:, 0], d0[:, 0])
        wy, dy = (w1[:, 1], d1[:, 1]) if oy else (w0[:, 1], d0[:, 1])
        wz, dz = (w1[:, 2], d1[:, 2]) if oz else (w0[:, 2], d0[:, 2])
        flat = (np.mod(base[:, 0] + ox, nx) * ny + np.mod(base[:, 1] + oy, ny)) * nz + np.mod(base[:, 2] + oz, nz)
        grad = np.stack([dx * wy * wz, wx * dy * wz, wx * wy * dz], axis=-1)
        yield flat, (wx * wy) * wz, grad


def cic_paint_vjp(pos, mesh_bar, mode="relative", weight=1.0):
    """(pos_bar, weight_bar) of <mesh_bar, cic_paint(pos, weight)> (float64)."""
    pos = np.asarray(pos)
    mb = np.asarray(mesh_bar, dtype=np.float64).reshape(-1)
    n = int(np.prod(pos.shape[:3]))
    wt = np.broadcast_to(np.asarray(weight, dtype=np.float64).reshape(-1), (n,))
    pb = np.zeros((n, 3))
    wb = np.zeros(n)
    for flat, w, g in _corner_terms(pos, np.asarray(mesh_bar).shape, mode):
        pb += g * (wt * mb[flat])[:, None]
        wb += w * mb[flat]
    return pb.reshape(pos.shape), wb.reshape(pos.shape[:3])


def cic_read_vjp(mesh, pos, out_bar, mode="relative"):
    """(mesh_bar, pos_bar) of <out_bar, cic_read(mesh, pos)> (float64); mesh (X,Y,Z[,C])."""
    mesh = np.asarray(mesh, dtype=np.float64)
    shape = mesh.shape[:3]
    nc = 1 if mesh.ndim == 3 else mesh.shape[3]
    m = mesh.reshape(-1, nc)
    pos = np.asarray(pos)
    ob = np.asarray(out_bar, dtype=np.float64).reshape(-1, nc)
    mb = np.zeros_like(m)
    pb = np.zeros((ob.shape[0], 3))
    for flat, w, g in _corner_terms(pos, shape, mode):
        for c in range(nc):
            mb[:, c] += np.bincount(flat, weights=w * ob[:, c], minlength=m.shape[0])
        pb += g * (ob * m[flat]).sum(axis=1)[:, None]
    return mb.reshape(mesh.shape), pb.reshape(pos.shape)


def force_mesh_vjp(fmesh_bar, deconv=False, pgd=None):
    """Density cotangent of force_mesh: each component is ifft(K_i fft(delta)) with K_i imaginary
    and odd (times the real PGD filter when pgd is given), whose transpose is its negative, so
    delta_bar = -sum_i force_mesh(fbar_i)[..., i]."""
    fb = np.asarray(fmesh_bar, dtype=np.float64)
    return -sum(force_mesh(fb[..., i], deconv=deconv, dtype=np.float64, pgd=pgd)[..., i] for i in range(3))


def pm_forces_vjp(pos, mesh_shape, forces_bar, mode="relative", scale=1.0, deconv=False, pgd=None):
    """pos_bar of <forces_bar, pm_forces(pos) * scale> (float64): readout-weight term plus the
    paint-weight term through the transposed force operator. pgd = (kl, ks): the PGD-filtered force
    of pgd_displacement (scale = alpha)."""
    rho = cic_paint(pos, mesh_shape, mode, dtype=np.float64)
    fm = force_mesh(rho, deconv=deconv, dtype=np.float64, pgd=pgd) * scale
    mesh_bar, pb_read = cic_read_vjp(fm, pos, forces_bar, mode)
    rho_bar = force_mesh_vjp(mesh_bar, deconv=deconv, pgd=pgd) * scale
    pb_paint, _ = cic_paint_vjp(pos, rho_bar, mode)
    return pb_read + pb_paint


# ---------------------------------------------------------------------------- CPU baseline step
def cpu_pm_step(x, v, force_scale, c_kick, c_drift, workers=1):
    """One PM step of the reference's op structure on the host, float32, for the CPU baseline:
    CIC paint, complex-to-complex 3D FFT, Green x gradient per component, 3 inverse c2c FFTs, a
    3-component CIC readout, kick, drift (pm_forces as called by _forces, solvers.py:19-32, then the
    kick/drift of solvers.py:296-318). FFTs use scipy.fft and the paint/readout the C loops
    (cic_paint_mt, cic_read3_mt), all on `workers` threads; NumPy loops when the C library is absent."""
    import scipy.fft as sfft

    mesh_shape = x.shape[:3]
    lib = _clib()
    x = np.ascontiguousarray(x, dtype=np.float32)
    if lib is not None:
        rho = np.empty(mesh_shape, dtype=np.float32)
        lib.cic_paint_mt(_cptr(x), *x.shape[:3], *mesh_shape, _cptr(rho), int(workers))
    else:
        rho = cic_paint(x, mesh_shape, "relative", dtype=np.float32)
    dk = sfft.fftn(rho.astype(np.complex64), workers=workers)
    kvec = fftk(mesh_shape, np.float32)
    pot = dk * invlaplace_kernel(kvec).astype(np.float32)
    fm = np.empty(tuple(mesh_shape) + (3,), dtype=np.float32)
    for i in range(3):
        g = gradient_kernel(kvec, i).astype(np.complex64)
        fm[..., i] = sfft.ifftn(-g * pot, workers=workers).real
    if lib is not None:
        f = np.empty(x.shape, dtype=np.float32)
        lib.cic_read3_mt(_cptr(fm), _cptr(x), *x.shape[:3], *mesh_shape, _cptr(f), int(workers))
    else:
        f = np.stack([cic_read(fm[..., i], x, "relative", dtype=np.float32) for i in range(3)], axis=-1)
    f *= np.float32(force_scale)
    v = v + f * np.float32(c_kick)
    x = x + v * np.float32(c_drift)
    return x, v
