"""CPU oracle for the integration path: background cosmology, LPT, IC colouring, the two solvers and
the time loop.

TEST INFRASTRUCTURE ONLY (imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg). An independent restatement of the reference: the growth ODE is solved with SciPy's DOP853 at
each requested scale factor, dGf/da is built from D'' exactly as jaxpm.growth.dGfa does (not from the
closed form 1.5 Om D / (a^2 E)), distances use scipy.integrate.quad, and the solver algebra follows
src/jax_fli/pm/solvers.py line by line.

Parity status:
  growth / Gf / gp / dGfa / E          pinned (EdS closed forms; growth-equation identity)
  comoving distance                     pinned (EdS closed form 2 c/H0 (1 - sqrt(a)))
  EH98 transfer, sigma8 normalisation   pinned by the paper's own limits and fits: T -> 1 as k -> 0;
                                        f_b -> 0 reduces it to T0(k, 1, 1) exactly and to the
                                        independent zero-baryon fit (eqs. 28-31) within 1%; the sound
                                        horizon at the drag epoch (eqs. 4, 6) within 2% of the fit of
                                        eq. 26; sigma(8) = sigma8 by adaptive quadrature. The jax_cosmo
                                        discretisation (its k table, its sigma8 integral) is parity
                                        unpinned (not in the reference tree)
  time_grid                             pinned (hand enumeration of integrate.py semantics)
  kdk / kkd coefficient algebra         pinned (linear-force runs reproduce D(a)/D(a0) exactly in
                                        EdS and LCDM; KKD step/reverse round trip)
"""

from __future__ import annotations

import math

import numpy as np
from scipy import integrate as _si

from . import pm_oracle as orc

C_OVER_H0 = 2997.92458
T_CMB = 2.726


class Cosmo:
    def __init__(self, Omega_c=0.2607, Omega_b=0.0490, h=0.6766, n_s=0.9665, sigma8=0.8102, Omega_k=0.0, w0=-1.0,
                 wa=0.0):
        self.Omega_c, self.Omega_b, self.h, self.n_s, self.sigma8 = Omega_c, Omega_b, h, n_s, sigma8
        self.Omega_k, self.w0, self.wa = Omega_k, w0, wa
        self.Omega_m = Omega_c + Omega_b
        self.Omega_de = 1.0 - Omega_k - self.Omega_m
        self._sol = None

    # E^2(a) = Om a^-3 + Ok a^-2 + Ode a^(-3(1+w0+wa)) exp(-3 wa (1-a))
    def Esqr(self, a):
        a = np.asarray(a, dtype=np.float64)
        de = a ** (-3.0 * (1 + self.w0 + self.wa)) * np.exp(-3.0 * self.wa * (1 - a))
        return self.Omega_m / a**3 + self.Omega_k / a**2 + self.Omega_de * de

    def E(self, a):
        return np.sqrt(self.Esqr(np.atleast_1d(a)))

    def dEda(self, a, eps=1e-6):
        a = np.asarray(a, dtype=np.float64)
        return (self.E(a * (1 + eps)) - self.E(a * (1 - eps))) / (2 * eps * a)

    # growth: D'' + (3/a + E'/E) D' - 1.5 Om D / (a^5 E^2) = 0, D = a in matter domination
    def _rhs(self, a, y):
        D, Dp = y
        dlnE = float(self.dEda(a)[0] / self.E(a)[0])
        return [Dp, -(3.0 / a + dlnE) * Dp + 1.5 * self.Omega_m * D / (a**5 * float(self.Esqr(a)))]

    def _solve(self):
        if self._sol is None:
            a0 = 1e-5
            # past a = 1: the KDK's nominal next half-kick time sqrt(t1 (t1 + dt)) exceeds 1
            self._sol = _si.solve_ivp(self._rhs, (a0, 2.1), [a0, 1.0], method="DOP853", rtol=1e-12, atol=1e-16,
                                      dense_output=True)
            self._norm = self._sol.sol(1.0)[0]
        return self._sol

    def D_and_derivs(self, a):
        a = np.atleast_1d(np.asarray(a, dtype=np.float64))
        sol = self._solve()
        y = sol.sol(a) / self._norm
        D, Dp = y[0], y[1]
        Dpp = np.array([self._rhs(ai, [di, dpi])[1] for ai, di, dpi in zip(a, D, Dp)])
        return D, Dp, Dpp

    def growth_factor(self, a):
        return self.D_and_derivs(a)[0]

    def growth_rate(self, a):
        D, Dp, _ = self.D_and_derivs(a)
        return Dp * np.atleast_1d(a) / D

    def Gf(self, a):
        a = np.atleast_1d(np.asarray(a, dtype=np.float64))
        D, Dp, _ = self.D_and_derivs(a)
        return Dp * a**3 * self.E(a)

    def gp(self, a):
        return self.D_and_derivs(a)[1]

    def dGfa(self, a):
        # d(a^3 E D')/da with D'' from the growth equation (the form jaxpm.growth.dGfa evaluates)
        a = np.atleast_1d(np.asarray(a, dtype=np.float64))
        D, Dp, Dpp = self.D_and_derivs(a)
        return Dpp * a**3 * self.E(a) + Dp * a**3 * self.dEda(a) + 3 * a**2 * self.E(a) * Dp

    # second order: D2'' + (3/a + E'/E) D2' = 1.5 Om (D2 - D^2) / (a^5 E^2), D2 = -3/7 a^2 in matter
    # domination, integrated with D (unnormalised) and normalised by D2(1)
    def _rhs2(self, a, y):
        D, Dp, D2, D2p = y
        dlnE = float(self.dEda(a)[0] / self.E(a)[0])
        r = 1.5 * self.Omega_m / (a**5 * float(self.Esqr(a)))
        q = 3.0 / a + dlnE
        return [Dp, -q * Dp + r * D, D2p, -q * D2p + r * (D2 - D * D)]

    def _solve2(self):
        if getattr(self, "_sol2", None) is None:
            a0 = 1e-5
            self._sol2 = _si.solve_ivp(self._rhs2, (a0, 2.1), [a0, 1.0, -3.0 / 7.0 * a0**2, -6.0 / 7.0 * a0],
                                       method="DOP853", rtol=1e-12, atol=1e-20, dense_output=True)
            self._norm2 = self._sol2.sol(1.0)[2]
        return self._sol2

    def growth_factor_second(self, a):
        a = np.atleast_1d(np.asarray(a, dtype=np.float64))
        return self._solve2().sol(a)[2] / self._norm2

    def growth_rate_second(self, a):
        a = np.atleast_1d(np.asarray(a, dtype=np.float64))
        y = self._solve2().sol(a)
        return a * y[3] / y[2]

    def chi(self, a):
        a = np.atleast_1d(np.asarray(a, dtype=np.float64))
        f = lambda x: C_OVER_H0 / (x * x * float(self.E(x)[0]))
        return np.array([_si.quad(f, ai, 1.0, epsabs=1e-12, epsrel=1e-12)[0] for ai in a])


# ------------------------------------------------------------------------------ EH98 + sigma8
def eh98_scales(c: Cosmo):
    """EH98 (Eisenstein & Hu 1998, ApJ 496, 605) eqs. 2-7: k_eq and the Silk scale in h/Mpc, z_eq, the
    drag redshift z_d, R at z_eq and z_d, and the sound horizon s at the drag epoch in Mpc/h."""
    T2 = (T_CMB / 2.7) ** 2
    wm, wb = c.Omega_m * c.h**2, c.Omega_b * c.h**2
    keq = 7.46e-2 * wm / T2 / c.h
    zeq = 2.50e4 * wm / T2**2
    b1 = 0.313 * wm**-0.419 * (1 + 0.607 * wm**0.674)
    b2 = 0.238 * wm**0.223
    zd = 1291 * wm**0.251 / (1 + 0.659 * wm**0.828) * (1 + b1 * wb**b2)  # eq. 4: the last power is of Omega_b h^2
    Req = 31.5 * wb / T2**2 * (1e3 / zeq)
    Rd = 31.5 * wb / T2**2 * (1e3 / zd)
    s = 2 / (3 * keq) * math.sqrt(6 / Req) * math.log((math.sqrt(1 + Rd) + math.sqrt(Rd + Req)) / (1 + math.sqrt(Req)))
    ksilk = 1.6 * wb**0.52 * wm**0.73 * (1 + (10.4 * wm) ** -0.95) / c.h
    return dict(keq=keq, zeq=zeq, zd=zd, Req=Req, Rd=Rd, s=s, ksilk=ksilk)


def eisenstein_hu(c: Cosmo, k):
    k = np.atleast_1d(np.asarray(k, dtype=np.float64))
    wm = c.Omega_m * c.h**2
    fb, fc = c.Omega_b / c.Omega_m, c.Omega_c / c.Omega_m
    sc = eh98_scales(c)
    keq, zeq, zd, Rd, s, ksilk = sc["keq"], sc["zeq"], sc["zd"], sc["Rd"], sc["s"], sc["ksilk"]
    q = k / (13.41 * keq)

    def T0(alpha, beta):
        L = np.log(math.e + 1.8 * beta * q)
        return L / (L + (14.2 / alpha + 386.0 / (1 + 69.9 * q**1.08)) * q * q)

    a1 = (46.9 * wm) ** 0.670 * (1 + (32.1 * wm) ** -0.532)
    a2 = (12.0 * wm) ** 0.424 * (1 + (45.0 * wm) ** -0.582)
    ac = a1**-fb * a2 ** -(fb**3)
    bc = 1.0 / (1 + 0.944 / (1 + (458 * wm) ** -0.708) * (fc ** ((0.395 * wm) ** -0.0266) - 1))
    f = 1 / (1 + (k * s / 5.4) ** 4)
    Tc = f * T0(1.0, bc) + (1 - f) * T0(ac, bc)
    y = (1 + zeq) / (1 + zd)
    x = math.sqrt(1 + y)
    G = y * (-6 * x + (2 + 3 * y) * math.log((x + 1) / (x - 1)))
    ab = 2.07 * keq * s * (1 + Rd) ** -0.75 * G
    bnode = 8.41 * wm**0.435
    st = s / (1 + (bnode / (k * s)) ** 3) ** (1 / 3)
    bb = 0.5 + fb + (3 - 2 * fb) * math.sqrt((17.2 * wm) ** 2 + 1)
    Tb = (T0(1.0, 1.0) / (1 + (k * s / 5.2) ** 2) + ab / (1 + (bb / (k * s)) ** 3) * np.exp(-((k / ksilk) ** 1.4))) * np.sinc(k * st / np.pi)
    return fb * Tb + fc * Tc


def sigma_R(c: Cosmo, pk_fn, R=8.0):
    """sigma(R) by 64-point Gauss-Legendre on 400 segments of ln k in [1e-5, 1e3]."""
    xg, wg = np.polynomial.legendre.leggauss(64)
    edges = np.linspace(math.log(1e-5), math.log(1e3), 401)
    lo, hi = edges[:-1, None], edges[1:, None]
    lnk = (0.5 * (hi - lo) * xg[None, :] + 0.5 * (hi + lo)).reshape(-1)
    wts = (0.5 * (hi - lo) * wg[None, :]).reshape(-1)
    k = np.exp(lnk)
    x = k * R
    w = 3 * (np.sin(x) - x * np.cos(x)) / x**3
    return math.sqrt(float(np.sum(wts * k**3 * pk_fn(k) * w * w / (2 * math.pi**2))))


def linear_matter_power(c: Cosmo, k, a=1.0):
    k = np.asarray(k, dtype=np.float64)
    raw = lambda kk: kk**c.n_s * eisenstein_hu(c, kk) ** 2
    norm = (c.sigma8 / sigma_R(c, raw)) ** 2
    return norm * raw(k) * c.growth_factor(a) ** 2


# ------------------------------------------------------------------------------ IC colouring, LPT
def jc_interp(x, xp, fp):
    """Linear interpolation with linear extrapolation (jax_cosmo.scipy.interpolate.interp)."""
    ind = np.clip(np.searchsorted(xp, x) - 1, 0, len(xp) - 2)
    xi = xp[ind]
    return fp[ind] + (x - xi) * (fp[ind + 1] - fp[ind]) / (xp[ind + 1] - xi)


def color_field(noise, box, ktab, ptab):
    """interpolate_initial_conditions (initial.py:175-182) with the P(k) table given."""
    shape = noise.shape
    field = orc.fftn(np.asarray(noise, dtype=np.float64))
    kv = orc.fftk(shape)
    kmesh = np.sqrt(sum((kk / box[i] * shape[i]) ** 2 for i, kk in enumerate(kv)))
    pkm = jc_interp(kmesh.reshape(-1), np.asarray(ktab), np.asarray(ptab)).reshape(kmesh.shape)
    pkm = pkm * np.prod(shape) / np.prod(box)
    return orc.ifftn(field * np.sqrt(np.maximum(pkm, 0.0))).real


def lpt1(c: Cosmo, ic, a):
    """jaxpm lpt order 1 at the lattice: dx = D F(ic), p = a^2 f E dx."""
    F = orc.force_mesh(np.asarray(ic, dtype=np.float64))
    D = float(c.growth_factor(a)[0])
    dx = D * F
    p = a**2 * float(c.growth_rate(a)[0]) * float(c.E(a)[0]) * dx
    return dx, p


def lpt(c: Cosmo, ic, a, order=1):
    """jaxpm.pm.lpt at the lattice (src/jax_fli/pm/lpt.py:119-127; upstream JaxPM, Jenkins 2010 eqs.
    2-3): dx = D F(ic) [+ 3/7 D2(a)/D2(1) F(delta2)], p = a^2 f E D F [+ a^2 f2 E 3/7 D2 F(delta2)].
    `a` is a scalar or an array broadcasting against (X, Y, Z, 1) (per-particle scale factors)."""
    ic = np.asarray(ic, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    F = orc.force_mesh(ic)
    shp = a.shape
    af = a.reshape(-1)
    D = c.growth_factor(af).reshape(shp)
    E = c.E(af).reshape(shp)
    f = c.growth_rate(af).reshape(shp)
    dx = D * F
    p = a**2 * f * E * dx
    if order == 2:
        F2 = orc.force_mesh(orc.lpt2_source(ic))
        dx2 = 3.0 / 7.0 * c.growth_factor_second(af).reshape(shp) * F2
        p2 = a**2 * c.growth_rate_second(af).reshape(shp) * E * dx2
        dx = dx + dx2
        p = p + p2
    return dx, p


def a_of_chi(c: Cosmo, chi, a_min=1e-3, n=4001):
    """Inverse of the comoving distance: chi(a) on a geometric grid of a (per-segment quadrature),
    ln a as a cubic spline of chi."""
    key = (a_min, n)
    if getattr(c, "_achi", None) is None or c._achi[0] != key:
        a = np.geomspace(a_min, 1.0, n)
        f = lambda x: C_OVER_H0 / (x * x * float(c.E(x)[0]))
        seg = np.array([_si.quad(f, a[i], a[i + 1], epsabs=1e-13, epsrel=1e-13)[0] for i in range(n - 1)])
        chi_tab = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
        from scipy.interpolate import CubicSpline
        c._achi = (key, CubicSpline(chi_tab[::-1], np.log(a)[::-1]))
    return np.exp(c._achi[1](np.asarray(chi, dtype=np.float64)))


def lightcone_scale_factors(c: Cosmo, mesh_shape, box, observer):
    """compute_particle_scale_factors (src/jax_fli/utils.py:234-244) for the lattice particles:
    a(|cell * box / mesh - observer|), shape (X, Y, Z, 1)."""
    ax = [np.arange(n) * (b / n) - o for n, b, o in zip(mesh_shape, box, observer)]
    r = np.sqrt(ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2)
    return a_of_chi(c, r)[..., None]


# ------------------------------------------------------------------------------ time loop + solvers
def time_grid(t0, ts, dt0, t1):
    """integrate.py:_fwd_loop in float32: per snapshot, steps of dt0 clipped at the snapshot time."""
    f = np.float32
    events = []
    t = f(t0)
    cap = int(math.ceil(t1 / dt0)) + 5
    for i, target in enumerate(np.atleast_1d(np.asarray(ts, dtype=np.float32))):
        count = 0
        while t < target and count < cap:
            nxt = f(t + f(dt0))
            nxt = target if nxt > f(target - f(1e-6)) else nxt
            events.append(("step", t, nxt))
            t = nxt
            count += 1
        events.append(("save", i, target))
    return events


def _g(fn, a):
    return float(np.atleast_1d(fn(float(a)))[0])


def kdk_coeffs(c: Cosmo, t0, t1, dt):
    f = np.float32
    t0, t1 = f(t0), f(t1)
    t0t1 = f(np.sqrt(f(t0 * t1)))
    drift = (_g(c.growth_factor, t1) - _g(c.growth_factor, t0)) / _g(c.gp, t0t1)
    cd = drift / (float(t0t1) ** 3 * _g(c.E, t0t1))
    t2 = f(t1 + f(dt))
    t1t2 = f(np.sqrt(f(t1 * t2)))
    k1 = (_g(c.Gf, t1) - _g(c.Gf, t0t1)) / _g(c.dGfa, t1)
    k2 = (_g(c.Gf, t1t2) - _g(c.Gf, t1)) / _g(c.dGfa, t1)
    ck = (k1 + k2) / (float(t1) ** 2 * _g(c.E, t1))
    return cd, ck


def kdk_init_coeff(c: Cosmo, t0, t1):
    f = np.float32
    ai = f(t0)
    af = f((float(t0) * float(t1)) ** 0.5)
    return (_g(c.Gf, af) - _g(c.Gf, ai)) / _g(c.dGfa, ai) / (float(ai) ** 2 * _g(c.E, ai))


def kkd_coeffs(c: Cosmo, t0, t1, dt, t_initial):
    f = np.float32
    t0, t1 = f(t0), f(t1)
    tp = max(f(t0 - f(dt)), f(t_initial))
    ai1 = f(np.sqrt(f(tp * t0)))
    af2 = f(np.sqrt(f(t0 * t1)))
    k1 = (_g(c.Gf, t0) - _g(c.Gf, ai1)) / _g(c.dGfa, t0)
    k2 = (_g(c.Gf, af2) - _g(c.Gf, t0)) / _g(c.dGfa, t0)
    ck = (k1 + k2) / (float(t0) ** 2 * _g(c.E, t0))
    drift = (_g(c.growth_factor, t1) - _g(c.growth_factor, t0)) / _g(c.gp, af2)
    cd = drift / (float(af2) ** 3 * _g(c.E, af2))
    return ck, cd


def run_nbody(c: Cosmo, dx, p, t0, t1, dt0, ts, solver="kdk", force_fn=None, dtype=np.float64, on_save=None,
              correction=None):
    """Full integration; force_fn(x) defaults to 1.5 Om pm_forces on the particle lattice of dx.
    Arithmetic on x, v in `dtype` with the reference's operation order (x + v*c, v + F*c).
    correction = ("pgd" | "sharpen", alpha, kl, ks): KDK applies it after the drift with the drift
    factor without its prefactor (solvers.py:175), KKD as a boost before the drift with the full
    drift factor (solvers.py:311-316)."""
    x = np.asarray(dx, dtype=dtype).copy()
    v = np.asarray(p, dtype=dtype).copy()
    mesh = x.shape[:3]
    if force_fn is None:
        force_fn = lambda xx: orc.pm_forces(xx, mesh, "relative", scale=1.5 * c.Omega_m, dtype=dtype)
    cast = lambda s: dtype(np.float32(s)) if dtype == np.float32 else dtype(s)
    if solver == "kdk":
        ck = kdk_init_coeff(c, t0, t0 + dt0)
        v = v + force_fn(x).astype(dtype) * cast(ck)
    corr = None
    if correction is not None:
        kind, alpha, kl, ks = correction
        corr = lambda xx: orc.pgd_displacement(xx, mesh, alpha, kl, ks, "relative", dtype=dtype)
    for ev in time_grid(t0, ts, dt0, t1):
        if ev[0] == "save":
            if on_save is not None:
                on_save(ev[1], ev[2], x.copy(), v.copy())
            continue
        _, ta, tb = ev
        if solver == "kdk":
            cd, ck = kdk_coeffs(c, ta, tb, dt0)
            x = x + v * cast(cd)
            if corr is not None:
                s = corr(x).astype(dtype)
                if kind == "pgd":
                    x = x + s
                else:
                    ta32, tb32 = np.float32(ta), np.float32(tb)
                    tm = np.float32(np.sqrt(np.float32(ta32 * tb32)))
                    df = float((_g(c.growth_factor, tb32) - _g(c.growth_factor, ta32)) / _g(c.gp, tm))
                    v = v + s / cast(df)
            v = v + force_fn(x).astype(dtype) * cast(ck)
        else:
            ck, cd = kkd_coeffs(c, ta, tb, dt0, t0)
            f = force_fn(x).astype(dtype)
            v = v + f * cast(ck)
            if corr is not None:
                if kind != "sharpen":
                    raise ValueError("the reversible solver takes the sharpening correction")
                v = v + corr(x).astype(dtype) / cast(cd)
            x = x + v * cast(cd)
    return x, v
