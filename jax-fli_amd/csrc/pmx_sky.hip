// Sky projections: HEALPix shell painting over a table of box images (periodic replicas, optionally
// rotated by cube symmetries), its position adjoint, and Born convergence with its adjoint.
//
// The HEALPix kernels take a 16^3 Lagrangian sub-block of particles per workgroup. With more than
// one image, the workgroup first bounds its particles (observer-centred, Mpc/h) and keeps only the
// images whose copy of that box meets the shell, in table order, so a particle is tested against a
// handful of images instead of all of them. Pixel geometry is float64 throughout.
#include "pmx_internal.h"
#include "pmx_device.h"
#include <cfloat>
#include <cmath>

#define PMX_CHECK_LAUNCH(what)             \
  do {                                     \
    const int rc_ = pmx_check_launch(what); \
    if (rc_) return rc_;                   \
  } while (0)

namespace {

// ---------------------------------------------------------------- HEALPix (RING scheme)
// Restates the published HEALPix pixelisation (Gorski et al. 2005) and its bilinear ring
// interpolation: loc2pix, ring_above, get_ring_info2 and get_interpol of the HEALPix C++ library.
struct HpGeom {
  int nside;
  long npix, ncap;
  double fact1, fact2;
  const double* rth;  // colatitude of each ring 1 .. 4 nside - 1 (index = ring), built by k_ring_theta
};

__device__ __forceinline__ long hp_ang2pix(const HpGeom& g, double z, double phi) {
  const double za = fabs(z);
  const double inv_halfpi = 0.63661977236758134308;
  double tt = fmod(phi * inv_halfpi, 4.0);
  if (tt < 0) tt += 4.0;
  const long ns = g.nside;
  if (za <= 2.0 / 3.0) {
    const long nl4 = 4 * ns;
    const double t1 = ns * (0.5 + tt), t2 = ns * z * 0.75;
    const long jp = (long)(t1 - t2), jm = (long)(t1 + t2);
    const long ir = ns + 1 + jp - jm;
    const long kshift = 1 - (ir & 1);
    const long t = jp + jm - ns + kshift + 1 + nl4 + nl4;
    const long ip = (t >> 1) % nl4;
    return g.ncap + (ir - 1) * nl4 + ip;
  }
  const double tp = tt - (long)tt;
  const double tmp = ns * sqrt(3.0 * (1.0 - za));
  const long jp = (long)(tp * tmp), jm = (long)((1.0 - tp) * tmp);
  const long ir = jp + jm + 1;
  long ip = (long)(tt * ir);
  ip = ip % (4 * ir);
  if (ip < 0) ip += 4 * ir;
  return (z > 0) ? 2 * ir * (ir - 1) + ip : g.npix - 2 * ir * (ir + 1) + ip;
}

// unit vector of the centre of RING pixel `pix`
__device__ void hp_pix2vec(const HpGeom& g, long pix, double v[3]) {
  const long ns = g.nside;
  double z, st, phi;
  if (pix < g.ncap) {
    long iring = (long)((1.0 + sqrt(1.0 + 2.0 * (double)pix)) * 0.5);
    while (2 * iring * (iring - 1) > pix) --iring;
    while (2 * (iring + 1) * iring <= pix) ++iring;
    const long iphi = pix - 2 * iring * (iring - 1);
    const double t = (double)iring * iring * 4.0 / (double)g.npix;
    z = 1.0 - t;
    st = sqrt(t * (2.0 - t));
    phi = ((double)iphi + 0.5) * M_PI / (2.0 * iring);
  } else if (pix < g.npix - g.ncap) {
    const long ip = pix - g.ncap, iring = ip / (4 * ns) + ns, iphi = ip % (4 * ns) + 1;
    const double fodd = ((iring + ns) & 1) ? 1.0 : 0.5;
    z = (double)(2 * ns - iring) * 2.0 / (3.0 * ns);
    st = sqrt((1.0 - z) * (1.0 + z));
    phi = ((double)iphi - fodd) * M_PI / (2.0 * ns);
  } else {
    const long ip = g.npix - pix;
    long iring = (long)((1.0 + sqrt(2.0 * (double)ip - 1.0)) * 0.5);
    while (2 * iring * (iring - 1) >= ip) --iring;
    while (2 * (iring + 1) * iring < ip) ++iring;
    const long iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    const double t = (double)iring * iring * 4.0 / (double)g.npix;
    z = -1.0 + t;
    st = sqrt(t * (2.0 - t));
    phi = ((double)iphi - 0.5) * M_PI / (2.0 * iring);
  }
  v[0] = st * cos(phi);
  v[1] = st * sin(phi);
  v[2] = z;
}

// rbf_neighbor candidates and weights of a direction n (unit): its pixel and the pixels of the eight
// directions one pixel resolution away (azimuths k pi/4 in the tangent plane), each weighted by
// exp(-d^2 / (2 sigma^2)), d the angular distance from n to the pixel centre, normalised to sum 1.
// Returns the number of distinct pixels (<= 9).
__device__ int hp_rbf(const HpGeom& g, const double n[3], double sigma, long pix[9], double wgt[9]) {
  const double st = sqrt(n[0] * n[0] + n[1] * n[1]);
  double ct = n[2], cp = 1.0, sp = 0.0;
  if (st > 0.0) {
    cp = n[0] / st;
    sp = n[1] / st;
  }
  const double eth[3] = {ct * cp, ct * sp, -st}, eph[3] = {-sp, cp, 0.0};
  const double step = sqrt(4.0 * M_PI / (double)g.npix), cs = cos(step), ss = sin(step);
  int m = 0;
  for (int k = -1; k < 8; ++k) {
    double u[3];
    if (k < 0) {
      u[0] = n[0];
      u[1] = n[1];
      u[2] = n[2];
    } else {
      const double ca = cos(0.25 * M_PI * k), sa = sin(0.25 * M_PI * k);
      for (int c = 0; c < 3; ++c) u[c] = cs * n[c] + ss * (ca * eth[c] + sa * eph[c]);
    }
    const double z = fmin(1.0, fmax(-1.0, u[2]));
    double ph = atan2(u[1], u[0]);
    if (ph < 0.0) ph += 2.0 * M_PI;
    if (ph >= 2.0 * M_PI) ph = 0.0;
    const long q = hp_ang2pix(g, z, ph);
    bool seen = false;
    for (int j = 0; j < m; ++j) seen |= (pix[j] == q);
    if (!seen) pix[m++] = q;
  }
  // exp(-(d^2 - dmin^2) / (2 sigma^2)): the nearest candidate weighs 1, so no width underflows to 0 / 0
  double q2[9], qmin = 1e300, sum = 0.0;
  for (int j = 0; j < m; ++j) {
    double c[3];
    hp_pix2vec(g, pix[j], c);
    const double dx = n[0] - c[0], dy = n[1] - c[1], dz = n[2] - c[2];
    const double d = 2.0 * asin(fmin(1.0, 0.5 * sqrt(dx * dx + dy * dy + dz * dz)));
    q2[j] = (d / sigma) * (d / sigma);
    qmin = fmin(qmin, q2[j]);
  }
  for (int j = 0; j < m; ++j) {
    wgt[j] = exp(-0.5 * (q2[j] - qmin));
    sum += wgt[j];
  }
  for (int j = 0; j < m; ++j) wgt[j] /= sum;
  return m;
}

__device__ __forceinline__ long hp_ring_above(const HpGeom& g, double z) {
  const double az = fabs(z);
  if (az <= 2.0 / 3.0) return (long)(g.nside * (2.0 - 1.5 * z));
  const long iring = (long)(g.nside * sqrt(3.0 * (1.0 - az)));
  return (z > 0) ? iring : 4 * (long)g.nside - iring - 1;
}

// colatitude of ring `ring` (1 .. 4 nside - 1)
__device__ __forceinline__ double hp_ring_theta(const HpGeom& g, long ring) {
  const long ns = g.nside;
  const long northring = (ring > 2 * ns) ? 4 * ns - ring : ring;
  double theta;
  if (northring < ns) {
    const double tmp = (double)northring * northring * g.fact2;
    const double costheta = 1.0 - tmp;
    const double sintheta = sqrt(tmp * (2.0 - tmp));
    theta = atan2(sintheta, costheta);
  } else {
    theta = acos((2 * ns - northring) * g.fact1);
  }
  return northring != ring ? M_PI - theta : theta;
}

// first pixel, pixel count and shift of ring `ring`; its colatitude from the table g.rth
__device__ __forceinline__ void hp_ring_info(const HpGeom& g, long ring, long& startpix, long& ringpix, double& theta,
                                             bool& shifted) {
  const long ns = g.nside;
  const long northring = (ring > 2 * ns) ? 4 * ns - ring : ring;
  if (northring < ns) {
    ringpix = 4 * northring;
    shifted = true;
    startpix = 2 * northring * (northring - 1);
  } else {
    ringpix = 4 * ns;
    shifted = ((northring - ns) & 1) == 0;
    startpix = g.ncap + (northring - ns) * ringpix;
  }
  if (northring != ring) startpix = g.npix - startpix - ringpix;
  theta = g.rth[ring];
}

// pixels on one ring around phi and the phi weight of the second one; idphi = 1 / ring pixel width.
// In units of the pixel width, t = phi / width - shift / 2 puts pixel i1 = floor(t) at t = i1 and the
// weight of the second pixel is t - i1; both come from one product instead of three divisions
// (the last-bit differences only move t across a pixel edge, where the two weights are 0 and 1).
__device__ __forceinline__ void hp_ring_pair(const HpGeom& g, long ring, double phi, long& p1, long& p2, double& w1,
                                             double& theta, double& idphi) {
  long sp, nr;
  bool shift;
  hp_ring_info(g, ring, sp, nr, theta, shift);
  idphi = (double)nr * (0.5 / M_PI);
  const double tmp = phi * idphi - 0.5 * shift;
  long i1 = (tmp < 0) ? (long)tmp - 1 : (long)tmp;
  w1 = tmp - (double)i1;
  long i2 = i1 + 1;
  if (i1 < 0) i1 += nr;
  if (i2 >= nr) i2 -= nr;
  p1 = sp + i1;
  p2 = sp + i2;
}

// Bilinear weights of the four RING pixels around (theta, phi); DER also gives d wgt / d theta and
// d wgt / d phi (the integer ring/pixel choices are locally constant).
template <bool DER>
__device__ __forceinline__ void hp_interpol(const HpGeom& g, double theta, double z, double phi, long pix[4],
                                            double wgt[4], double dth[4], double dph[4]) {
  // z = cos(theta), passed in as the caller's z / r
  const long ir1 = hp_ring_above(g, z);
  const long ir2 = ir1 + 1;
  const long nring = 4 * (long)g.nside;
  double theta1 = 0, theta2 = 0, idphi;
  double w1;
  if (ir1 > 0) {
    hp_ring_pair(g, ir1, phi, pix[0], pix[1], w1, theta1, idphi);
    wgt[0] = 1 - w1;
    wgt[1] = w1;
    if (DER) {
      dph[0] = -idphi;
      dph[1] = idphi;
    }
  }
  if (ir2 < nring) {
    hp_ring_pair(g, ir2, phi, pix[2], pix[3], w1, theta2, idphi);
    wgt[2] = 1 - w1;
    wgt[3] = w1;
    if (DER) {
      dph[2] = -idphi;
      dph[3] = idphi;
    }
  }
  if (ir1 == 0) {  // north polar cap: the missing ring is the 4 pixels around the pole
    const double wtheta = theta / theta2, dw = 1.0 / theta2;
    if (DER) {
      dth[0] = dth[1] = -0.25 * dw;
      dth[2] = (wgt[2] - 0.25) * dw;
      dth[3] = (wgt[3] - 0.25) * dw;
      dph[0] = dph[1] = 0.0;
      dph[2] *= wtheta;
      dph[3] *= wtheta;
    }
    wgt[2] *= wtheta;
    wgt[3] *= wtheta;
    const double fac = (1 - wtheta) * 0.25;
    wgt[0] = fac;
    wgt[1] = fac;
    wgt[2] += fac;
    wgt[3] += fac;
    pix[0] = (pix[2] + 2) & 3;
    pix[1] = (pix[3] + 2) & 3;
  } else if (ir2 == nring) {  // south polar cap
    const double wtheta = (theta - theta1) / (M_PI - theta1), dw = 1.0 / (M_PI - theta1);
    if (DER) {
      dth[0] = (0.25 - wgt[0]) * dw;
      dth[1] = (0.25 - wgt[1]) * dw;
      dth[2] = dth[3] = 0.25 * dw;
      dph[0] *= (1 - wtheta);
      dph[1] *= (1 - wtheta);
      dph[2] = dph[3] = 0.0;
    }
    wgt[0] *= (1 - wtheta);
    wgt[1] *= (1 - wtheta);
    const double fac = wtheta * 0.25;
    wgt[0] += fac;
    wgt[1] += fac;
    wgt[2] = fac;
    wgt[3] = fac;
    pix[2] = ((pix[0] + 2) & 3) + g.npix - 4;
    pix[3] = ((pix[1] + 2) & 3) + g.npix - 4;
  } else {
    const double wtheta = (theta - theta1) / (theta2 - theta1), dw = 1.0 / (theta2 - theta1);
    if (DER) {
      dth[0] = -wgt[0] * dw;
      dth[1] = -wgt[1] * dw;
      dth[2] = wgt[2] * dw;
      dth[3] = wgt[3] * dw;
      dph[0] *= (1 - wtheta);
      dph[1] *= (1 - wtheta);
      dph[2] *= wtheta;
      dph[3] *= wtheta;
    }
    wgt[0] *= (1 - wtheta);
    wgt[1] *= (1 - wtheta);
    wgt[2] *= wtheta;
    wgt[3] *= wtheta;
  }
}

// ---------------------------------------------------------------- box images
// Image t maps an observer-centred position X to R_t X + shift_t. R_t is the cube symmetry with
// index rot[t] in the enumeration "for signs in product([1, -1], repeat=3): for perm in
// permutations((0, 1, 2))" (index = 6 * sign_index + perm_index; 0 is the identity):
// (R X)_i = sign_i * X_{perm_i}. This is the table of get_all_47_symmetries
// (src/jax_fli/pm/interp.py:31-43), whose entries 1..47 the reference uses.
struct Sym {
  int p[3];
  double s[3];
};
__device__ __forceinline__ Sym sym_of(int code) {
  Sym r;
  code = ((code % 48) + 48) % 48;  // the table has 48 entries
  const int pi = code % 6, si = code / 6;
  const int p0 = pi >> 1;
  const int a = (p0 == 0) ? 1 : 0, b = (p0 == 2) ? 1 : 2;  // the two remaining axes, a < b
  r.p[0] = p0;
  r.p[1] = (pi & 1) ? b : a;
  r.p[2] = (pi & 1) ? a : b;
  r.s[0] = (si & 4) ? -1.0 : 1.0;
  r.s[1] = (si & 2) ? -1.0 : 1.0;
  r.s[2] = (si & 1) ? -1.0 : 1.0;
  return r;
}

struct ShellArgs {
  double cell[3];   // Mpc/h per grid unit
  double obs[3];    // observer, Mpc/h
  double rmin, rmax;
  int scheme;       // 0 ngp, 1 bilinear, 2 rbf_neighbor
  double rbf_sigma;  // rbf_neighbor: Gaussian width (radians)
  int NX, NY, NZ;
  int rx, ry, rz;   // lattice spacing of the particle grid in cells
  int nbx, nby, nbz;
  int ntiles;       // image count; rot == nullptr and ntiles == 1: the box itself
  const int* rot;
  const double* shift;  // 3 per image, Mpc/h
  // drift to the lightcone (v != nullptr): each image position moves by f (R v) with
  // f = (Gp(a_t) - gp_now) / dGfa(sqrt(a_now a_t)), a_t = a(chi = |position|), from the tables
  // dtab = {a(chi) on [0, chi_max], Gp(a) and dGfa(a) on [a_lo, a_hi]}, ntab points each
  const float* v;
  const double* dtab;
  int ntab;
  double chi_max, a_lo, a_hi, a_now, gp_now;
  double fbound;    // bound on |f| over the tables' range (widens the image culling boxes)
  const double* ptab;  // adjoint: nparam tables of df/dtheta on the a(chi) grid
  int nparam;
  double* pdots;       // adjoint: += sum over particle images of (W . G) df/dtheta(|Y|) weight
};

constexpr int HP_HASH = 4096, HP_PROBES = 8, HP_PROBES_FULL = 1, HP_SB = 16, HP_PER_THREAD = HP_SB * HP_SB * HP_SB / 256;
constexpr int HP_MAX_TILES = 4096;
constexpr unsigned HP_EMPTY = 0xFFFFFFFFu;

// particle position in grid units, lattice added in relative mode, wrapped into the periodic box
__device__ __forceinline__ double grid_pos(int mode, int lat, float xc, int n) {
  double g = (double)xc + (mode == 0 ? (double)lat : 0.0);
  return g - floor(g / n) * n;
}

// the r-th particle of this thread in the workgroup's sub-block: flat index, or -1 past the grid
__device__ __forceinline__ long sb_particle(const PartArgs& a, const ShellArgs& sa, unsigned b, int r, int& i, int& j,
                                            int& k) {
  const int bk = (int)(b % (unsigned)sa.nbz), bj = (int)((b / (unsigned)sa.nbz) % (unsigned)sa.nby),
            bi = (int)(b / ((unsigned)sa.nbz * sa.nby));
  const int l = r * 256 + (int)threadIdx.x;
  i = bi * HP_SB + (l >> 8);
  j = bj * HP_SB + ((l >> 4) & 15);
  k = bk * HP_SB + (l & 15);
  if (i >= a.npx || j >= a.npy || k >= a.npz) return -1;
  return ((long)i * a.npy + j) * a.npz + k;
}

__device__ __forceinline__ void centred(const PartArgs& a, const ShellArgs& sa, long p, int i, int j, int k,
                                        double X[3]) {
  const float3 x = ld3(a.x, (unsigned)p);
  X[0] = grid_pos(a.mode, i * sa.rx, x.x, sa.NX) * sa.cell[0] - sa.obs[0];
  X[1] = grid_pos(a.mode, j * sa.ry, x.y, sa.NY) * sa.cell[1] - sa.obs[1];
  X[2] = grid_pos(a.mode, k * sa.rz, x.z, sa.NZ) * sa.cell[2] - sa.obs[2];
}

__device__ __forceinline__ void image_of(const ShellArgs& sa, int t, const double X[3], double Xt[3]) {
  if (!sa.rot) {
    for (int c = 0; c < 3; ++c) Xt[c] = X[c];
    return;
  }
  const Sym R = sym_of(sa.rot[t]);
  for (int c = 0; c < 3; ++c) Xt[c] = R.s[c] * X[R.p[c]] + sa.shift[3 * t + c];
}

// linear interpolation in a table of n values on the uniform grid [x0, x1], clamped at the ends
__device__ __forceinline__ double tab_lin(const double* t, int n, double x0, double x1, double x) {
  const double u = (x - x0) / (x1 - x0) * (double)(n - 1);
  if (!(u > 0.0)) return t[0];
  if (u >= (double)(n - 1)) return t[n - 1];
  const int i = (int)u;
  return t[i] + (t[i + 1] - t[i]) * (u - (double)i);
}

// velocity of image t (rotated like the position, not shifted)
__device__ __forceinline__ void image_vel(const ShellArgs& sa, int t, const double V[3], double Vt[3]) {
  if (!sa.rot) {
    for (int c = 0; c < 3; ++c) Vt[c] = V[c];
    return;
  }
  const Sym R = sym_of(sa.rot[t]);
  for (int c = 0; c < 3; ++c) Vt[c] = R.s[c] * V[R.p[c]];
}

// tab_lin and the slope of the segment it used (0 where the table is clamped)
__device__ __forceinline__ double tab_lin_d(const double* t, int n, double x0, double x1, double x, double& slope) {
  const double u = (x - x0) / (x1 - x0) * (double)(n - 1);
  if (!(u > 0.0)) {
    slope = 0.0;
    return t[0];
  }
  if (u >= (double)(n - 1)) {
    slope = 0.0;
    return t[n - 1];
  }
  const int i = (int)u;
  slope = (t[i + 1] - t[i]) * (double)(n - 1) / (x1 - x0);
  return t[i] + (t[i + 1] - t[i]) * (u - (double)i);
}

// drift factor f(r) = (Gp(a_t) - Gp(a_now)) / dGfa(sqrt(a_now a_t)), a_t = a(chi = r), of
// _drift_to_time (src/jax_fli/pm/interp.py:110-121), and its derivative df/dr (of the tables'
// linear interpolation, so it is the exact derivative of what the paint computes)
__device__ __forceinline__ double drift_factor(const ShellArgs& sa, double r, double& dfdr) {
  double da, dgp, ddg;
  const double at = tab_lin_d(sa.dtab, sa.ntab, 0.0, sa.chi_max, r, da);
  const double gpt = tab_lin_d(sa.dtab + sa.ntab, sa.ntab, sa.a_lo, sa.a_hi, at, dgp);
  const double bt = sqrt(sa.a_now * at);
  const double dg = tab_lin_d(sa.dtab + 2 * sa.ntab, sa.ntab, sa.a_lo, sa.a_hi, bt, ddg);
  const double num = gpt - sa.gp_now;
  const double dfda = dgp / dg - num * ddg * (bt > 0.0 ? 0.5 * sa.a_now / bt : 0.0) / (dg * dg);
  dfdr = dfda * da;
  return num / dg;
}

// _drift_to_time at the image's own lightcone time: Xt += f(|Xt|) Vt
__device__ __forceinline__ void drift_image(const ShellArgs& sa, double Xt[3], const double Vt[3]) {
  const double r = sqrt(Xt[0] * Xt[0] + Xt[1] * Xt[1] + Xt[2] * Xt[2]);
  double dfdr;
  const double f = drift_factor(sa, r, dfdr);
  for (int c = 0; c < 3; ++c) Xt[c] += f * Vt[c];
}

// Builds the workgroup's list of images (table order) whose copy of the particles' bounding box
// (grown by the largest drift when drifting) meets the shell. Returns the count; hits[] holds the
// image indices.
__device__ int shell_images(const PartArgs& a, const ShellArgs& sa, unsigned b, unsigned short* hits) {
  __shared__ double red[6][4];
  __shared__ int wcount[4];
  if (!sa.rot) {  // the box alone: every particle is tested once
    if (threadIdx.x == 0) hits[0] = 0;
    __syncthreads();
    return 1;
  }
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int r = 0; r < HP_PER_THREAD; ++r) {
    int i, j, k;
    const long p = sb_particle(a, sa, b, r, i, j, k);
    if (p < 0) continue;
    double X[3];
    centred(a, sa, p, i, j, k, X);
    double pad = 0.0;
    if (sa.v) {  // |f (R v)_c| <= fbound max_c |v_c| on every axis
      const float3 v = ld3(sa.v, (unsigned)p);
      pad = sa.fbound * (double)fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
    }
    for (int c = 0; c < 3; ++c) {
      lo[c] = fmin(lo[c], X[c] - pad);
      hi[c] = fmax(hi[c], X[c] + pad);
    }
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int c = 0; c < 3; ++c) {
      lo[c] = fmin(lo[c], __shfl_xor(lo[c], o));
      hi[c] = fmax(hi[c], __shfl_xor(hi[c], o));
    }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int c = 0; c < 3; ++c) {
      red[c][wv] = lo[c];
      red[3 + c][wv] = hi[c];
    }
  __syncthreads();
  for (int c = 0; c < 3; ++c) {
    lo[c] = fmin(fmin(red[c][0], red[c][1]), fmin(red[c][2], red[c][3]));
    hi[c] = fmax(fmax(red[3 + c][0], red[3 + c][1]), fmax(red[3 + c][2], red[3 + c][3]));
  }
  const bool empty = !(lo[0] <= hi[0]);
  // conservative margins; the per-particle test decides
  const double r2max = sa.rmax * sa.rmax * (1.0 + 1e-9) + 1e-9, r2min = sa.rmin * sa.rmin * (1.0 - 1e-9);
  int base = 0;
  for (int t0 = 0; t0 < sa.ntiles; t0 += 256) {
    const int t = t0 + (int)threadIdx.x;
    bool hit = false;
    if (!empty && t < sa.ntiles) {
      const Sym R = sym_of(sa.rot[t]);
      double dmin2 = 0.0, dmax2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        double l = lo[R.p[c]], h = hi[R.p[c]];
        if (R.s[c] < 0) {
          const double tl = -h;
          h = -l;
          l = tl;
        }
        l += sa.shift[3 * t + c];
        h += sa.shift[3 * t + c];
        const double dn = (l > 0) ? l : ((h < 0) ? -h : 0.0);
        dmin2 += dn * dn;
        dmax2 += fmax(l * l, h * h);
      }
      hit = dmin2 <= r2max && dmax2 >= r2min;
    }
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wcount[wv] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wv; ++w) off += wcount[w];
    if (hit) hits[off + __popcll(m & ((1ull << lane) - 1ull))] = (unsigned short)t;
    base += wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
  }
  return base;
}

// Per-workgroup aggregation: the (pixel, weight) deposits of the sub-block go through an LDS hash
// table, and each distinct pixel then reaches the map with one float64 atomic. A deposit that
// finds no free slot within HP_PROBES probes goes straight to the map and marks the table full;
// after that a deposit probes only HP_PROBES_FULL slots (a sub-block near the observer spans more
// pixels than the table holds, and long probe runs would cost more than the atomics they save).
__device__ __forceinline__ void hp_table_add(volatile unsigned* keys, double* vals, unsigned pix, double v,
                                             double* __restrict__ acc, volatile int* full) {
  unsigned h = (pix * 2654435761u) >> 20;  // 12-bit multiplicative hash
  const int probes = *full ? HP_PROBES_FULL : HP_PROBES;
#pragma unroll 1
  for (int t = 0; t < probes; ++t) {
    const unsigned k = keys[h];
    if (k == pix) {
      atomicAdd(&vals[h], v);
      return;
    }
    if (k == HP_EMPTY) {
      const unsigned old = atomicCAS(const_cast<unsigned*>(keys + h), HP_EMPTY, pix);
      if (old == HP_EMPTY || old == pix) {
        atomicAdd(&vals[h], v);
        return;
      }
    }
    h = (h + 1) & (HP_HASH - 1);
  }
  *full = 1;
  unsafeAtomicAdd(acc + pix, v);
}

__device__ __forceinline__ bool in_shell(const ShellArgs& sa, const double Xt[3], double& rr) {
  rr = sqrt(Xt[0] * Xt[0] + Xt[1] * Xt[1] + Xt[2] * Xt[2]);
  return rr > sa.rmin && rr <= sa.rmax;
}

__device__ __forceinline__ double phi_of(const double Xt[3]) {
  const double phi0 = atan2(Xt[1], Xt[0]);
  const double phi = phi0 < 0 ? phi0 + 2.0 * M_PI : phi0;
  return (phi < 2.0 * M_PI) ? phi : 0.0;  // -tiny + 2 pi can round up to 2 pi
}

// One instance per scheme (0 ngp, 1 bilinear, 2 rbf_neighbor): the rbf scheme's nine pixels per
// image would otherwise set the register allocation of every scheme (256 VGPRs, one wave per SIMD,
// against 155 and two for bilinear alone).
template <int SCHEME>
__global__ void __launch_bounds__(256) k_paint_healpix(PartArgs a, const float* __restrict__ w, float w0, ShellArgs sa,
                                                       HpGeom g, double* __restrict__ acc) {
  __shared__ unsigned keys[HP_HASH];
  __shared__ double vals[HP_HASH];
  // max(1, image count) entries, sized at launch: with the 48 KiB table, three workgroups per CU for
  // up to 128 images instead of two for a fixed HP_MAX_TILES list (256^3 shells 20% faster)
  extern __shared__ unsigned short hits[];
  __shared__ int full;
  for (int i = threadIdx.x; i < HP_HASH; i += blockDim.x) {
    keys[i] = HP_EMPTY;
    vals[i] = 0.0;
  }
  if (threadIdx.x == 0) full = 0;
  const unsigned b = xcd_logical(blockIdx.x, gridDim.x);
  const int nhit = shell_images(a, sa, b, hits);  // synchronises the workgroup
  for (int r = 0; r < HP_PER_THREAD; ++r) {
    int i, j, k;
    const long p = sb_particle(a, sa, b, r, i, j, k);
    if (p < 0) continue;
    double X[3], V[3] = {0.0, 0.0, 0.0};
    centred(a, sa, p, i, j, k, X);
    if (sa.v) {
      const float3 v = ld3(sa.v, (unsigned)p);
      V[0] = v.x;
      V[1] = v.y;
      V[2] = v.z;
    }
    const double wt = w ? (double)w[p] : (double)w0;
    for (int h = 0; h < nhit; ++h) {
      double Xt[3], rr;
      image_of(sa, hits[h], X, Xt);
      if (sa.v) {
        double Vt[3];
        image_vel(sa, hits[h], V, Vt);
        drift_image(sa, Xt, Vt);
      }
      if (!in_shell(sa, Xt, rr)) continue;
      const double phi = phi_of(Xt);
      if constexpr (SCHEME == 0) {
        hp_table_add(keys, vals, (unsigned)hp_ang2pix(g, Xt[2] / rr, phi), wt, acc, &full);
      } else if constexpr (SCHEME == 2) {
        const double n[3] = {Xt[0] / rr, Xt[1] / rr, Xt[2] / rr};
        long pix[9];
        double wg[9];
        const int m = hp_rbf(g, n, sa.rbf_sigma, pix, wg);
        for (int q = 0; q < m; ++q) hp_table_add(keys, vals, (unsigned)pix[q], wt * wg[q], acc, &full);
      } else {
        const double theta = atan2(sqrt(Xt[0] * Xt[0] + Xt[1] * Xt[1]), Xt[2]);
        long pix[4];
        double wg[4], d0[4], d1[4];
        hp_interpol<false>(g, theta, Xt[2] / rr, phi, pix, wg, d0, d1);
#pragma unroll
        for (int m = 0; m < 4; ++m) hp_table_add(keys, vals, (unsigned)pix[m], wt * wg[m], acc, &full);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < HP_HASH; i += blockDim.x) {
    const unsigned k = keys[i];
    if (k != HP_EMPTY) unsafeAtomicAdd(acc + k, vals[i]);
  }
}

// Position cotangent of the bilinear shell paint: for every image of the particle inside the shell,
// d/dX sum_m map_bar[pix_m] wgt_m(theta, phi), rotated back to the box frame, times
// scale * weight * (Mpc/h per grid unit). The ngp scheme is piecewise constant: zero.
constexpr int kDriftParams = 4;  // cosmological parameters the drift adjoint differentiates at once

// *dst += the workgroup's sum of v (float64 atomic: the order of the workgroups' additions varies)
__device__ __forceinline__ void block_sum_add(double v, double* dst) {
  __shared__ double part[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (__lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += part[i];
    unsafeAtomicAdd(dst, t);
  }
  __syncthreads();
}

// With drift (sa.v): Z = Y + f(|Y|) W for the image Y = R X + shift, W = R V, so
// dL/dY = G + f'(|Y|) (W . G) Y / |Y| and dL/dW = f G, G = dL/dZ; both are rotated back by R^T.
// vel_bar receives the velocity cotangent (drift only).
__global__ void __launch_bounds__(256) k_healpix_vjp(PartArgs a, const float* __restrict__ w, float w0, ShellArgs sa,
                                                     HpGeom g, const float* __restrict__ map_bar, double scale,
                                                     float* __restrict__ pos_bar, float* __restrict__ vel_bar,
                                                     int accumulate) {
  __shared__ unsigned short hits[HP_MAX_TILES];
  const unsigned b = xcd_logical(blockIdx.x, gridDim.x);
  const int nhit = shell_images(a, sa, b, hits);
  double pacc[kDriftParams] = {0.0, 0.0, 0.0, 0.0};
  for (int r = 0; r < HP_PER_THREAD; ++r) {
    int i, j, k;
    const long p = sb_particle(a, sa, b, r, i, j, k);
    if (p < 0) continue;
    double X[3], V[3] = {0.0, 0.0, 0.0};
    centred(a, sa, p, i, j, k, X);
    if (sa.v) {
      const float3 v = ld3(sa.v, (unsigned)p);
      V[0] = v.x;
      V[1] = v.y;
      V[2] = v.z;
    }
    double gX[3] = {0.0, 0.0, 0.0}, gV[3] = {0.0, 0.0, 0.0}, pp[kDriftParams] = {0.0, 0.0, 0.0, 0.0};
    if (sa.scheme == 1) {
      for (int h = 0; h < nhit; ++h) {
        const int t = hits[h];
        double Y[3], Xt[3], W[3] = {0.0, 0.0, 0.0}, rr, f = 0.0, dfdr = 0.0, ry = 0.0;
        image_of(sa, t, X, Y);
        for (int c = 0; c < 3; ++c) Xt[c] = Y[c];
        if (sa.v) {
          image_vel(sa, t, V, W);
          ry = sqrt(Y[0] * Y[0] + Y[1] * Y[1] + Y[2] * Y[2]);
          f = drift_factor(sa, ry, dfdr);
          for (int c = 0; c < 3; ++c) Xt[c] += f * W[c];
        }
        if (!in_shell(sa, Xt, rr)) continue;
        const double rho2 = Xt[0] * Xt[0] + Xt[1] * Xt[1];
        if (!(rho2 > 0.0)) continue;  // on the polar axis phi is undefined
        const double rho = sqrt(rho2);
        const double phi = phi_of(Xt);
        const double theta = atan2(rho, Xt[2]);
        long pix[4];
        double wg[4], dth[4], dph[4];
        hp_interpol<true>(g, theta, Xt[2] / rr, phi, pix, wg, dth, dph);
        double gt = 0.0, gp = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const double mb = (double)map_bar[pix[m]];
          gt += mb * dth[m];
          gp += mb * dph[m];
        }
        const double r2 = rr * rr;
        // theta = atan2(rho, Z), phi = atan2(Y, X)
        const double ct = gt * Xt[2] / (rho * r2);
        double G[3];
        G[0] = ct * Xt[0] - gp * Xt[1] / rho2;
        G[1] = ct * Xt[1] + gp * Xt[0] / rho2;
        G[2] = -gt * rho / r2;
        double GY[3] = {G[0], G[1], G[2]}, GW[3] = {0.0, 0.0, 0.0};
        if (sa.v) {
          const double wg = W[0] * G[0] + W[1] * G[1] + W[2] * G[2];
          const double c = ry > 0.0 ? dfdr * wg / ry : 0.0;
          for (int q = 0; q < 3; ++q) {
            GY[q] += c * Y[q];
            GW[q] = f * G[q];
          }
          for (int q = 0; q < sa.nparam; ++q) pp[q] += wg * tab_lin(sa.ptab + q * sa.ntab, sa.ntab, 0.0, sa.chi_max, ry);
        }
        if (sa.rot) {
          const Sym R = sym_of(sa.rot[t]);
          for (int c = 0; c < 3; ++c) {
            gX[R.p[c]] += R.s[c] * GY[c];
            gV[R.p[c]] += R.s[c] * GW[c];
          }
        } else {
          for (int c = 0; c < 3; ++c) {
            gX[c] += GY[c];
            gV[c] += GW[c];
          }
        }
      }
    }
    const double wt = (w ? (double)w[p] : (double)w0) * scale;
    for (int q = 0; q < sa.nparam; ++q) pacc[q] += pp[q] * wt;
    float3 o = make_float3((float)(gX[0] * wt * sa.cell[0]), (float)(gX[1] * wt * sa.cell[1]),
                           (float)(gX[2] * wt * sa.cell[2]));
    if (accumulate) {
      const float3 q = ld3(pos_bar, (unsigned)p);
      o = make_float3(q.x + o.x, q.y + o.y, q.z + o.z);
    }
    st3(pos_bar, (unsigned)p, o);
    if (vel_bar) {
      float3 ov = make_float3((float)(gV[0] * wt), (float)(gV[1] * wt), (float)(gV[2] * wt));
      if (accumulate) {
        const float3 q = ld3(vel_bar, (unsigned)p);
        ov = make_float3(q.x + ov.x, q.y + ov.y, q.z + ov.z);
      }
      st3(vel_bar, (unsigned)p, ov);
    }
  }
  for (int q = 0; q < sa.nparam; ++q) block_sum_add(pacc[q], sa.pdots + q);
}

__global__ void k_finalize_map_d(const double* __restrict__ acc, float* __restrict__ map, long n, double scale,
                                 int accumulate) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float v = (float)(acc[i] * scale);
    map[i] = accumulate ? map[i] + v : v;
  }
}

__global__ void k_zero_d(double* __restrict__ a, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0.0;
}

// ---------------------------------------------------------------- Born convergence
// kappa[k][p] = sum_s C[k][s] (maps[s][p] / mean_s - 1), mean_s = the shell's pixel mean (0 is
// replaced by FLT_EPSILON, as the reference does with the float32 eps). C (float64, host-built)
// carries 3/2 Om (H0/c)^2 d_r r_s / a_s times the n(z)-weighted lensing efficiency.
// Scratch in plan->red: [0, kBornCoef) per-workgroup partial sums, then the coefficient area.
constexpr int kBornCoef = kRedMax - 8192;
constexpr int kBornMaxShells = 256, kBornMaxSources = 16;

__device__ __forceinline__ void chunk_of(long n, int nblk, int b, long& beg, long& end) {
  const long per = (n + nblk - 1) / nblk;
  beg = (long)b * per;
  end = beg + per < n ? beg + per : n;
  if (beg > n) beg = n;
}

// part[s * nblk + b] = sum of maps[s][chunk b]
__global__ void __launch_bounds__(256) k_map_sums(const float* __restrict__ maps, long npix, int nblk,
                                                  double* __restrict__ part) {
  const int s = blockIdx.y, b = blockIdx.x;
  long beg, end;
  chunk_of(npix, nblk, b, beg, end);
  const float* m = maps + (size_t)s * npix;
  double t = 0.0;
  for (long i = beg + threadIdx.x; i < end; i += blockDim.x) t += (double)m[i];
  block_sum_to(t, part + (size_t)s * nblk + b);
}

// coefficient area: f[k*S+s] = C/mean_s (K*S), mean[s] (raw), cb[k] = sum_s C[k][s]
__global__ void __launch_bounds__(256) k_born_coef(const double* __restrict__ part, int nblk, int S, int K, long npix,
                                                   const double* __restrict__ C, double* __restrict__ coef) {
  double* f = coef;
  double* mean = coef + K * S;
  double* cb = mean + S;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    double t = 0.0;
    for (int b = 0; b < nblk; ++b) t += part[(size_t)s * nblk + b];
    const double m = t / (double)npix;
    mean[s] = m;
    const double ms = (m == 0.0) ? (double)FLT_EPSILON : m;
    for (int k = 0; k < K; ++k) f[k * S + s] = C[k * S + s] / ms;
  }
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    double t = 0.0;
    for (int s = 0; s < S; ++s) t += C[k * S + s];
    cb[k] = t;
  }
}

__global__ void __launch_bounds__(256) k_born(const float* __restrict__ maps, long npix, int S, int K,
                                              const double* __restrict__ coef, float* __restrict__ kappa,
                                              int accumulate) {
  const double* f = coef;
  const double* cb = coef + K * S + S;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < npix; p += (long)gridDim.x * blockDim.x) {
    double acc[kBornMaxSources];
#pragma unroll
    for (int k = 0; k < kBornMaxSources; ++k) acc[k] = 0.0;
    for (int s = 0; s < S; ++s) {
      const double rho = (double)maps[(size_t)s * npix + p];
#pragma unroll
      for (int k = 0; k < kBornMaxSources; ++k)
        if (k < K) acc[k] += f[k * S + s] * rho;
    }
#pragma unroll
    for (int k = 0; k < kBornMaxSources; ++k)
      if (k < K) {
        const float v = (float)(acc[k] - cb[k]);
        float* o = kappa + (size_t)k * npix + p;
        *o = accumulate ? *o + v : v;
      }
  }
}

// partials of D[k][s] = sum_p kbar[k][p] maps[s][p] (blockIdx.y = s < S) and E[k] = sum_p kbar[k][p]
// (blockIdx.y = S), at part[((y * K) + k) * nblk + b]
__global__ void __launch_bounds__(256) k_born_dots(const float* __restrict__ maps, const float* __restrict__ kbar,
                                                   long npix, int S, int K, int nblk, double* __restrict__ part) {
  const int s = blockIdx.y, b = blockIdx.x;
  long beg, end;
  chunk_of(npix, nblk, b, beg, end);
  double acc[kBornMaxSources];
#pragma unroll
  for (int k = 0; k < kBornMaxSources; ++k) acc[k] = 0.0;
  for (long p = beg + threadIdx.x; p < end; p += blockDim.x) {
    const double m = (s < S) ? (double)maps[(size_t)s * npix + p] : 1.0;
#pragma unroll
    for (int k = 0; k < kBornMaxSources; ++k)
      if (k < K) acc[k] += (double)kbar[(size_t)k * npix + p] * m;
  }
  for (int k = 0; k < K; ++k) block_sum_to(acc[k], part + ((size_t)s * K + k) * nblk + b);
}

// h[s] = sum_k C[k][s] D[k][s] / (mean_s^2 npix) (0 where the mean was 0: the eps replacement is a
// constant); dots (optional) = {D (K*S, index k*S+s), E (K), mean (S, as used: 0 -> FLT_EPSILON)}
__global__ void __launch_bounds__(256) k_born_vjp_coef(const double* __restrict__ part, int nblk, int S, int K,
                                                       long npix, const double* __restrict__ C,
                                                       double* __restrict__ coef, double* __restrict__ dots) {
  const double* mean = coef + K * S;
  double* h = coef + K * S + S + K;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const double m = mean[s];
    double t = 0.0;
    for (int k = 0; k < K; ++k) {
      double d = 0.0;
      for (int b = 0; b < nblk; ++b) d += part[((size_t)s * K + k) * nblk + b];
      if (dots) dots[k * S + s] = d;
      t += C[k * S + s] * d;
    }
    h[s] = (m == 0.0) ? 0.0 : t / (m * m * (double)npix);
    if (dots) dots[K * S + K + s] = (m == 0.0) ? (double)FLT_EPSILON : m;
  }
  if (dots)
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
      double e = 0.0;
      for (int b = 0; b < nblk; ++b) e += part[((size_t)S * K + k) * nblk + b];
      dots[K * S + k] = e;
    }
}

__global__ void __launch_bounds__(256) k_born_vjp(const float* __restrict__ kbar, long npix, int S, int K,
                                                  const double* __restrict__ coef, float* __restrict__ maps_bar,
                                                  int accumulate) {
  const double* f = coef;
  const double* h = coef + K * S + S + K;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < npix; p += (long)gridDim.x * blockDim.x) {
    double kb[kBornMaxSources];
#pragma unroll
    for (int k = 0; k < kBornMaxSources; ++k) kb[k] = (k < K) ? (double)kbar[(size_t)k * npix + p] : 0.0;
    for (int s = 0; s < S; ++s) {
      double t = -h[s];
#pragma unroll
      for (int k = 0; k < kBornMaxSources; ++k)
        if (k < K) t += f[k * S + s] * kb[k];
      float* o = maps_bar + (size_t)s * npix + p;
      *o = accumulate ? *o + (float)t : (float)t;
    }
  }
}

// ---------------------------------------------------------------- flat-sky Born
// Plane s (ny rows, nx columns) is resampled at output pixel (i, j) by periodic bilinear
// interpolation at row coordinate u = j rs_s - 1/2 and column coordinate w = i cs_s - 1/2: the
// angular grid of _born_core_impl (src/jax_fli/_src/lensing/_born.py:53-98: x angles index rows,
// y angles columns, both scaled by chi_s / pixel size) through map_coordinates(order=1, mode="wrap").
// sc[2 s] = rs_s, sc[2 s + 1] = cs_s.
struct FlatSample {
  double val, du, dw;  // value and its derivatives along u and w
};

__device__ __forceinline__ long wrap_long(long i, long n) {
  const long r = i % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ void flat_corners(int ny, int nx, double u, double w, long& r0, long& r1, long& c0, long& c1,
                                             double& tu, double& tw) {
  const double fu = floor(u), fw = floor(w);
  tu = u - fu;
  tw = w - fw;
  r0 = wrap_long((long)fu, ny);
  c0 = wrap_long((long)fw, nx);
  r1 = (r0 + 1 == ny) ? 0 : r0 + 1;
  c1 = (c0 + 1 == nx) ? 0 : c0 + 1;
}

__device__ __forceinline__ FlatSample flat_sample(const float* __restrict__ pl, int ny, int nx, double u, double w) {
  long r0, r1, c0, c1;
  double tu, tw;
  flat_corners(ny, nx, u, w, r0, r1, c0, c1, tu, tw);
  const double v00 = pl[r0 * nx + c0], v01 = pl[r0 * nx + c1], v10 = pl[r1 * nx + c0], v11 = pl[r1 * nx + c1];
  FlatSample f;
  f.val = (1.0 - tu) * ((1.0 - tw) * v00 + tw * v01) + tu * ((1.0 - tw) * v10 + tw * v11);
  f.du = (1.0 - tw) * (v10 - v00) + tw * (v11 - v01);
  f.dw = (1.0 - tu) * (v01 - v00) + tu * (v11 - v10);
  return f;
}

// kappa[k][p] = sum_s f[k][s] R_s(p) - cb[k], R_s the resampled plane (f = C / mean of the plane)
__global__ void __launch_bounds__(256) k_born_flat(const float* __restrict__ maps, int ny, int nx, int S, int K,
                                                   const double* __restrict__ coef, const double* __restrict__ sc,
                                                   float* __restrict__ kappa, int accumulate) {
  const double* f = coef;
  const double* cb = coef + K * S + S;
  const long npix = (long)ny * nx;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < npix; p += (long)gridDim.x * blockDim.x) {
    const long i = p / nx, j = p - i * nx;
    double acc[kBornMaxSources];
#pragma unroll
    for (int k = 0; k < kBornMaxSources; ++k) acc[k] = 0.0;
    for (int s = 0; s < S; ++s) {
      const double R = flat_sample(maps + (size_t)s * npix, ny, nx, (double)j * sc[2 * s] - 0.5,
                                   (double)i * sc[2 * s + 1] - 0.5).val;
#pragma unroll
      for (int k = 0; k < kBornMaxSources; ++k)
        if (k < K) acc[k] += f[k * S + s] * R;
    }
#pragma unroll
    for (int k = 0; k < kBornMaxSources; ++k)
      if (k < K) {
        const float v = (float)(acc[k] - cb[k]);
        float* o = kappa + (size_t)k * npix + p;
        *o = accumulate ? *o + v : v;
      }
  }
}

// partials of D[k][s] = <kbar_k, R_s> (y = s < S), E[k] = sum kbar_k (y = S), as k_born_dots, and of
// P[s] = sum_p g_s(p) (dR_s/du j rs_s + dR_s/dw i cs_s) = chi_s d<g_s, R_s>/d chi_s with
// g_s = sum_k f[k][s] kbar_k, at part[((S + 1) K + s) nblk + b]
__global__ void __launch_bounds__(256) k_born_flat_dots(const float* __restrict__ maps, const float* __restrict__ kbar,
                                                        int ny, int nx, int S, int K, int nblk,
                                                        const double* __restrict__ coef, const double* __restrict__ sc,
                                                        double* __restrict__ part) {
  const int s = blockIdx.y, b = blockIdx.x;
  const long npix = (long)ny * nx;
  long beg, end;
  chunk_of(npix, nblk, b, beg, end);
  const double* f = coef;
  double acc[kBornMaxSources];
#pragma unroll
  for (int k = 0; k < kBornMaxSources; ++k) acc[k] = 0.0;
  double accp = 0.0;
  for (long p = beg + threadIdx.x; p < end; p += blockDim.x) {
    if (s < S) {
      const long i = p / nx, j = p - i * nx;
      const double rs = sc[2 * s], cs = sc[2 * s + 1];
      const FlatSample fs = flat_sample(maps + (size_t)s * npix, ny, nx, (double)j * rs - 0.5, (double)i * cs - 0.5);
      double g = 0.0;
#pragma unroll
      for (int k = 0; k < kBornMaxSources; ++k)
        if (k < K) {
          const double kb = (double)kbar[(size_t)k * npix + p];
          acc[k] += kb * fs.val;
          g += f[k * S + s] * kb;
        }
      accp += g * (fs.du * (double)j * rs + fs.dw * (double)i * cs);
    } else {
#pragma unroll
      for (int k = 0; k < kBornMaxSources; ++k)
        if (k < K) acc[k] += (double)kbar[(size_t)k * npix + p];
    }
  }
  for (int k = 0; k < K; ++k) block_sum_to(acc[k], part + ((size_t)s * K + k) * nblk + b);
  if (s < S) block_sum_to(accp, part + ((size_t)(S + 1) * K + s) * nblk + b);
}

// dots[K S + K + S + s] = P[s]
__global__ void __launch_bounds__(256) k_born_flat_p(const double* __restrict__ part, int nblk, int S, int K,
                                                     double* __restrict__ dots) {
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    double t = 0.0;
    for (int b = 0; b < nblk; ++b) t += part[((size_t)(S + 1) * K + s) * nblk + b];
    dots[K * S + K + S + s] = t;
  }
}

// maps_bar[s] (+)= -h[s] everywhere (the mean's term); the scatter pass adds the rest
__global__ void __launch_bounds__(256) k_born_flat_vjp_fill(long npix, int S, int K, const double* __restrict__ coef,
                                                            float* __restrict__ maps_bar, int accumulate) {
  const double* h = coef + K * S + S + K;
  const size_t n = (size_t)S * npix;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const float v = (float)(-h[q / npix]);
    maps_bar[q] = accumulate ? maps_bar[q] + v : v;
  }
}

// maps_bar[s] += R_s^T g_s: each output pixel adds g_s(p) times its 4 bilinear weights (float atomics)
__global__ void __launch_bounds__(256) k_born_flat_scatter(const float* __restrict__ kbar, int ny, int nx, int S, int K,
                                                           const double* __restrict__ coef,
                                                           const double* __restrict__ sc, float* __restrict__ maps_bar) {
  const double* f = coef;
  const long npix = (long)ny * nx;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < npix; p += (long)gridDim.x * blockDim.x) {
    const long i = p / nx, j = p - i * nx;
    double kb[kBornMaxSources];
#pragma unroll
    for (int k = 0; k < kBornMaxSources; ++k) kb[k] = (k < K) ? (double)kbar[(size_t)k * npix + p] : 0.0;
    for (int s = 0; s < S; ++s) {
      double g = 0.0;
#pragma unroll
      for (int k = 0; k < kBornMaxSources; ++k)
        if (k < K) g += f[k * S + s] * kb[k];
      if (g == 0.0) continue;
      long r0, r1, c0, c1;
      double tu, tw;
      flat_corners(ny, nx, (double)j * sc[2 * s] - 0.5, (double)i * sc[2 * s + 1] - 0.5, r0, r1, c0, c1, tu, tw);
      float* mb = maps_bar + (size_t)s * npix;
      unsafeAtomicAdd(mb + r0 * nx + c0, (float)(g * (1.0 - tu) * (1.0 - tw)));
      unsafeAtomicAdd(mb + r0 * nx + c1, (float)(g * (1.0 - tu) * tw));
      unsafeAtomicAdd(mb + r1 * nx + c0, (float)(g * tu * (1.0 - tw)));
      unsafeAtomicAdd(mb + r1 * nx + c1, (float)(g * tu * tw));
    }
  }
}

HpGeom hp_geom(int nside) {
  HpGeom g;
  g.nside = nside;
  g.npix = 12L * nside * nside;
  g.ncap = 2L * nside * (nside - 1);
  g.fact2 = 4.0 / (double)g.npix;
  g.fact1 = (double)(2 * nside) * g.fact2;
  g.rth = nullptr;
  return g;
}

namespace {
__global__ void k_ring_theta(HpGeom g, double* __restrict__ out) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r < 4L * g.nside) out[r] = r == 0 ? 0.0 : hp_ring_theta(g, r);
}
}  // namespace

// HEALPix geometry with its ring colatitude table, built on `s` after the npix map accumulators in
// p->acc (the caller sized p->acc for npix + 4 nside doubles)
static HpGeom hp_geom_table(pmx_plan_t* p, int nside, hipStream_t s, int& rc) {
  HpGeom g = hp_geom(nside);
  double* tab = p->acc + g.npix;
  hipLaunchKernelGGL(k_ring_theta, dim3((unsigned)((4L * nside + 255) / 256)), dim3(256), 0, s, g, tab);
  rc = pmx_check_launch("ring_theta");
  g.rth = tab;
  return g;
}

ShellArgs shell_args(const pmx_plan_t* p, const PartArgs& a, const double box[3], const double obs[3], double rmin,
                     double rmax, int scheme, int ntiles, const int* rot, const double* shift) {
  ShellArgs sa;
  sa.NX = p->nx;
  sa.NY = p->ny;
  sa.NZ = p->nz;
  const int n[3] = {p->nx, p->ny, p->nz};
  for (int c = 0; c < 3; ++c) {
    sa.cell[c] = box[c] / n[c];
    sa.obs[c] = obs[c];
  }
  sa.rmin = rmin;
  sa.rmax = rmax;
  sa.scheme = scheme;
  sa.rbf_sigma = p->rbf_sigma;
  sa.rx = p->nx / a.npx;
  sa.ry = p->ny / a.npy;
  sa.rz = p->nz / a.npz;
  sa.nbx = (a.npx + HP_SB - 1) / HP_SB;
  sa.nby = (a.npy + HP_SB - 1) / HP_SB;
  sa.nbz = (a.npz + HP_SB - 1) / HP_SB;
  sa.ntiles = rot ? ntiles : 1;
  sa.rot = rot;
  sa.shift = shift;
  sa.v = nullptr;
  sa.dtab = nullptr;
  sa.ntab = 0;
  sa.chi_max = sa.a_lo = sa.a_hi = sa.a_now = sa.gp_now = sa.fbound = 0.0;
  sa.ptab = nullptr;
  sa.nparam = 0;
  sa.pdots = nullptr;
  return sa;
}

int born_blocks(long npix, long rows) {
  long nblk = (npix + 1023) / 1024;
  const long cap = kBornCoef / (rows > 0 ? rows : 1);
  if (nblk > 256) nblk = 256;
  if (nblk > cap) nblk = cap;
  return (int)(nblk < 1 ? 1 : nblk);
}

}  // namespace

// ------------------------------------------------------------------ HEALPix orderings and ud_grade
namespace {
// Gorski et al. (2005): face f has corner ring jrll[f] nside and phi index jpll[f]; a NESTED index
// holds the face above the bit-interleaved (ix, iy) of the pixel inside it.
__constant__ int kJrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
__constant__ int kJpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

__device__ __forceinline__ long long spread_bits(long long v) {
  long long o = 0;
  for (int b = 0; b < 31; ++b) o |= ((v >> b) & 1ll) << (2 * b);
  return o;
}
__device__ __forceinline__ long long compact_bits(long long v) {
  long long o = 0;
  for (int b = 0; b < 31; ++b) o |= ((v >> (2 * b)) & 1ll) << b;
  return o;
}
__device__ __forceinline__ long long isqrt_ll(long long v) {
  long long r = (long long)sqrt((double)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}

__device__ long long hp_nest2ring(long long n, long long pix) {
  const long long npface = n * n, face = pix / npface, r = pix - face * npface;
  const long long ix = compact_bits(r), iy = compact_bits(r >> 1);
  const long long jr = kJrll[face] * n - ix - iy - 1;
  long long nr, kshift, before;
  const long long ncap = 2 * n * (n - 1), npix = 12 * n * n;
  if (jr < n) {
    nr = jr;
    kshift = 0;
    before = 2 * nr * (nr - 1);
  } else if (jr > 3 * n) {
    nr = 4 * n - jr;
    kshift = 0;
    before = npix - 2 * (nr + 1) * nr;
  } else {
    nr = n;
    kshift = (jr - n) & 1;
    before = ncap + (jr - n) * 4 * n;
  }
  long long jp = (kJpll[face] * nr + ix - iy + 1 + kshift) / 2;
  if (jp > 4 * n) jp -= 4 * n;
  if (jp < 1) jp += 4 * n;
  return before + jp - 1;
}

__device__ long long hp_ring2nest(long long n, long long pix) {
  const long long ncap = 2 * n * (n - 1), npix = 12 * n * n, nl2 = 2 * n;
  long long iring, iphi, kshift, nr, face;
  if (pix < ncap) {
    iring = (1 + isqrt_ll(1 + 2 * pix)) >> 1;
    iphi = (pix + 1) - 2 * iring * (iring - 1);
    kshift = 0;
    nr = iring;
    face = (iphi - 1) / nr;
  } else if (pix < npix - ncap) {
    const long long ip = pix - ncap, tmp = ip / (4 * n);
    iring = tmp + n;
    iphi = ip - tmp * 4 * n + 1;
    kshift = (iring + n) & 1;
    nr = n;
    const long long ire = iring - n + 1, irm = nl2 + 2 - ire;
    const long long ifm = (iphi - ire / 2 + n - 1) / n, ifp = (iphi - irm / 2 + n - 1) / n;
    face = (ifp == ifm) ? (ifp | 4) : ((ifp < ifm) ? ifp : ifm + 8);
  } else {
    const long long ip = npix - pix;
    iring = (1 + isqrt_ll(2 * ip - 1)) >> 1;
    iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    kshift = 0;
    nr = iring;
    iring = 2 * nl2 - iring;
    face = 8 + (iphi - 1) / nr;
  }
  const long long irt = iring - kJrll[face] * n + 1;
  long long ipt = 2 * iphi - kJpll[face] * nr - kshift - 1;
  if (ipt >= nl2) ipt -= 8 * n;
  const long long ix = (ipt - irt) >> 1, iy = (-ipt - irt) >> 1;
  return face * n * n + spread_bits(ix) + (spread_bits(iy) << 1);
}

__device__ __forceinline__ bool hp_bad(float v) {
  const double unseen = -1.6375e30;
  return !isfinite(v) || fabs((double)v - unseen) <= 1e-8 + 1e-5 * 1.6375e30;
}

// out pixel q (in order_out) from the map at nside_in (in order_in): the NESTED children average
// (finite, not UNSEEN ones; pess: UNSEEN unless all are good, else UNSEEN when none is) or the
// parent's value, scaled by ratio = (nside_out / nside_in)^power
__global__ void __launch_bounds__(256) k_ud_grade(const float* __restrict__ in, long long nin, float* __restrict__ out,
                                                  long long nout, int ring_in, int ring_out, int pess, double ratio) {
  const long long npix = 12 * nout * nout;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < npix;
       q += (long long)gridDim.x * blockDim.x) {
    const long long nq = ring_out ? hp_ring2nest(nout, q) : q;
    float v;
    if (nout >= nin) {
      const long long r2 = (nout / nin) * (nout / nin), par = nq / r2;
      v = (float)((double)in[ring_in ? hp_nest2ring(nin, par) : par] * ratio);
    } else {
      const long long r2 = (nin / nout) * (nin / nout);
      double sum = 0.0;
      long long nhit = 0;
      for (long long c = nq * r2; c < (nq + 1) * r2; ++c) {
        const float x = in[ring_in ? hp_nest2ring(nin, c) : c];
        if (!hp_bad(x)) {
          sum += (double)x;
          ++nhit;
        }
      }
      const bool badout = pess ? (nhit != r2) : (nhit == 0);
      const double hits = (double)nhit / ratio;
      v = badout ? -1.6375e30f : (float)(sum / hits);
    }
    out[q] = v;
  }
}
}  // namespace

int pmx_launch_ud_grade(const float* in, int nside_in, float* out, int nside_out, double ratio, int ring_in,
                        int ring_out, int pess, hipStream_t s) {
  const long long npix = 12ll * nside_out * nside_out;
  long long g = (npix + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_ud_grade, dim3((unsigned)g), dim3(256), 0, s, in, (long long)nside_in, out,
                     (long long)nside_out, ring_in, ring_out, pess, ratio);
  return pmx_check_launch("ud_grade");
}

// ------------------------------------------------------------------ launchers used by pmx_api.hip
int pmx_sky_max_tiles() { return HP_MAX_TILES; }
int pmx_born_max_shells() { return kBornMaxShells; }
int pmx_born_max_sources() { return kBornMaxSources; }

namespace {
// One image of the particles as absolute grid positions: X = unwrapped position (Mpc/h) - observer,
// Y = R X + shift, with sa.v Y += f(|Y|) R v (drift_image), out = (Y + observer) / cell. Grid-stride
// over particles; float64 arithmetic, rounded once to float32.
__global__ void __launch_bounds__(256) k_image_positions(PartArgs a, ShellArgs sa, int code, double sh0, double sh1,
                                                         double sh2, float* __restrict__ out) {
  const size_t n = (size_t)a.npx * a.npy * a.npz;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(p % (size_t)a.npz);
    const size_t q = p / (size_t)a.npz;
    const int j = (int)(q % (size_t)a.npy), i = (int)(q / (size_t)a.npy);
    const float3 x = ld3(a.x, (unsigned)p);
    const bool rel = a.mode == 0;
    const double X[3] = {((double)x.x + (rel ? (double)(i * sa.rx) : 0.0)) * sa.cell[0] - sa.obs[0],
                         ((double)x.y + (rel ? (double)(j * sa.ry) : 0.0)) * sa.cell[1] - sa.obs[1],
                         ((double)x.z + (rel ? (double)(k * sa.rz) : 0.0)) * sa.cell[2] - sa.obs[2]};
    const Sym R = sym_of(code);
    const double sh[3] = {sh0, sh1, sh2};
    double Y[3];
    for (int c = 0; c < 3; ++c) Y[c] = R.s[c] * X[R.p[c]] + sh[c];
    if (sa.v) {
      const float3 v = ld3(sa.v, (unsigned)p);
      const double V[3] = {v.x, v.y, v.z};
      double Vt[3];
      for (int c = 0; c < 3; ++c) Vt[c] = R.s[c] * V[R.p[c]];
      drift_image(sa, Y, Vt);
    }
    for (int c = 0; c < 3; ++c) out[3 * p + c] = (float)((Y[c] + sa.obs[c]) / sa.cell[c]);
  }
}

// Cotangents of k_image_positions for gbar (per particle, of the absolute grid positions out):
// Zbar = gbar / cell; with the drift, Ybar = Zbar + f'(|Y|) (W . Zbar) Y / |Y| and Wbar = f Zbar
// (W = R v); xbar (+)= cell * R^T Ybar, vbar (+)= R^T Wbar, and pdots[q] += sum (W . Zbar)
// df/dtheta_q(|Y|) from the parameter tables.
__global__ void __launch_bounds__(256) k_image_positions_vjp(PartArgs a, ShellArgs sa, int code, double sh0,
                                                             double sh1, double sh2, const float* __restrict__ gbar,
                                                             float* __restrict__ xbar, float* __restrict__ vbar,
                                                             int accumulate) {
  const size_t n = (size_t)a.npx * a.npy * a.npz;
  double pp[kDriftParams] = {0.0, 0.0, 0.0, 0.0};
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(p % (size_t)a.npz);
    const size_t q = p / (size_t)a.npz;
    const int j = (int)(q % (size_t)a.npy), i = (int)(q / (size_t)a.npy);
    const Sym R = sym_of(code);
    const float3 gb = ld3(gbar, (unsigned)p);
    const double Zb[3] = {gb.x / sa.cell[0], gb.y / sa.cell[1], gb.z / sa.cell[2]};
    double Yb[3] = {Zb[0], Zb[1], Zb[2]}, Wb[3] = {0.0, 0.0, 0.0};
    if (sa.v) {
      const float3 x = ld3(a.x, (unsigned)p);
      const bool rel = a.mode == 0;
      const double X[3] = {((double)x.x + (rel ? (double)(i * sa.rx) : 0.0)) * sa.cell[0] - sa.obs[0],
                           ((double)x.y + (rel ? (double)(j * sa.ry) : 0.0)) * sa.cell[1] - sa.obs[1],
                           ((double)x.z + (rel ? (double)(k * sa.rz) : 0.0)) * sa.cell[2] - sa.obs[2]};
      const double sh[3] = {sh0, sh1, sh2};
      const float3 v = ld3(sa.v, (unsigned)p);
      const double V[3] = {v.x, v.y, v.z};
      double Y[3], W[3];
      for (int c = 0; c < 3; ++c) {
        Y[c] = R.s[c] * X[R.p[c]] + sh[c];
        W[c] = R.s[c] * V[R.p[c]];
      }
      const double ry = sqrt(Y[0] * Y[0] + Y[1] * Y[1] + Y[2] * Y[2]);
      double dfdr;
      const double f = drift_factor(sa, ry, dfdr);
      const double wz = W[0] * Zb[0] + W[1] * Zb[1] + W[2] * Zb[2];
      const double cc = ry > 0.0 ? dfdr * wz / ry : 0.0;
      for (int c = 0; c < 3; ++c) {
        Yb[c] += cc * Y[c];
        Wb[c] = f * Zb[c];
      }
      for (int t = 0; t < sa.nparam; ++t) pp[t] += wz * tab_lin(sa.ptab + t * sa.ntab, sa.ntab, 0.0, sa.chi_max, ry);
    }
    double Xb[3] = {0.0, 0.0, 0.0}, Vb[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; ++c) {
      Xb[R.p[c]] += R.s[c] * Yb[c];
      Vb[R.p[c]] += R.s[c] * Wb[c];
    }
    float3 o = make_float3((float)(Xb[0] * sa.cell[0]), (float)(Xb[1] * sa.cell[1]), (float)(Xb[2] * sa.cell[2]));
    if (accumulate) {
      const float3 t = ld3(xbar, (unsigned)p);
      o = make_float3(t.x + o.x, t.y + o.y, t.z + o.z);
    }
    st3(xbar, (unsigned)p, o);
    if (vbar) {
      float3 ov = make_float3((float)Vb[0], (float)Vb[1], (float)Vb[2]);
      if (accumulate) {
        const float3 t = ld3(vbar, (unsigned)p);
        ov = make_float3(t.x + ov.x, t.y + ov.y, t.z + ov.z);
      }
      st3(vbar, (unsigned)p, ov);
    }
  }
  for (int t = 0; t < sa.nparam; ++t) block_sum_add(pp[t], sa.pdots + t);
}
}  // namespace

int pmx_launch_image_positions_vjp(pmx_plan_t* p, const PartArgs& a, const double box[3], const double obs[3],
                                   int code, const double shift[3], const float* gbar, float* xbar, float* vbar,
                                   int accumulate, hipStream_t s, const PmxDrift* drift) {
  ShellArgs sa = shell_args(p, a, box, obs, 0.0, 0.0, 0, 1, nullptr, nullptr);
  if (drift) {
    sa.v = drift->v;
    sa.dtab = drift->tables;
    sa.ntab = drift->ntab;
    sa.chi_max = drift->chi_max;
    sa.a_lo = drift->a_lo;
    sa.a_hi = drift->a_hi;
    sa.a_now = drift->a_now;
    sa.gp_now = drift->gp_now;
    sa.ptab = drift->ptab;
    sa.nparam = drift->ptab ? drift->nparam : 0;
    sa.pdots = drift->pdots;
  }
  const size_t n = (size_t)a.npx * a.npy * a.npz;
  hipLaunchKernelGGL(k_image_positions_vjp, dim3(grid_for(n, 256)), dim3(256), 0, s, a, sa, code, shift[0], shift[1],
                     shift[2], gbar, xbar, drift ? vbar : nullptr, accumulate);
  return pmx_check_launch("image_positions_vjp");
}

int pmx_launch_image_positions(pmx_plan_t* p, const PartArgs& a, const double box[3], const double obs[3], int code,
                               const double shift[3], float* out, hipStream_t s, const PmxDrift* drift) {
  ShellArgs sa = shell_args(p, a, box, obs, 0.0, 0.0, 0, 1, nullptr, nullptr);
  if (drift) {
    sa.v = drift->v;
    sa.dtab = drift->tables;
    sa.ntab = drift->ntab;
    sa.chi_max = drift->chi_max;
    sa.a_lo = drift->a_lo;
    sa.a_hi = drift->a_hi;
    sa.a_now = drift->a_now;
    sa.gp_now = drift->gp_now;
  }
  const size_t n = (size_t)a.npx * a.npy * a.npz;
  hipLaunchKernelGGL(k_image_positions, dim3(grid_for(n, 256)), dim3(256), 0, s, a, sa, code, shift[0], shift[1],
                     shift[2], out);
  return pmx_check_launch("image_positions");
}

int pmx_launch_healpix(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, const double box[3],
                       const double obs[3], double rmin, double rmax, int nside, int scheme, int ntiles,
                       const int* rot, const double* shift, float* map, int accumulate, hipStream_t s,
                       const PmxDrift* drift) {
  ShellArgs sa = shell_args(p, a, box, obs, rmin, rmax, scheme, ntiles, rot, shift);
  if (drift) {
    sa.v = drift->v;
    sa.dtab = drift->tables;
    sa.ntab = drift->ntab;
    sa.chi_max = drift->chi_max;
    sa.a_lo = drift->a_lo;
    sa.a_hi = drift->a_hi;
    sa.a_now = drift->a_now;
    sa.gp_now = drift->gp_now;
    sa.fbound = drift->fbound;
  }
  int rc = 0;
  const HpGeom g = hp_geom_table(p, nside, s, rc);
  if (rc) return rc;
  if (!(sa.rbf_sigma > 0.0))  // default width: the pixel resolution, read as a FWHM
    sa.rbf_sigma = sqrt(4.0 * M_PI / (double)g.npix) / (2.0 * sqrt(2.0 * log(2.0)));
  double* acc = p->acc;
  hipLaunchKernelGGL(k_zero_d, dim3(grid_for((size_t)g.npix, 256)), dim3(256), 0, s, acc, g.npix);
  PMX_CHECK_LAUNCH("healpix_zero");
  const dim3 grid((unsigned)sa.nbx * sa.nby * sa.nbz);
  const size_t shm = (((size_t)(sa.rot ? sa.ntiles : 1) * sizeof(unsigned short)) + 15) & ~(size_t)15;
  if (sa.scheme == 0)
    hipLaunchKernelGGL(k_paint_healpix<0>, grid, dim3(256), shm, s, a, w, w0, sa, g, acc);
  else if (sa.scheme == 2)
    hipLaunchKernelGGL(k_paint_healpix<2>, grid, dim3(256), shm, s, a, w, w0, sa, g, acc);
  else
    hipLaunchKernelGGL(k_paint_healpix<1>, grid, dim3(256), shm, s, a, w, w0, sa, g, acc);
  PMX_CHECK_LAUNCH("paint_healpix");
  const double vol = (4.0 * M_PI / (double)g.npix) * (rmax * rmax * rmax - rmin * rmin * rmin) / 3.0;
  hipLaunchKernelGGL(k_finalize_map_d, dim3(grid_for((size_t)g.npix, 256)), dim3(256), 0, s, acc, map, g.npix,
                     1.0 / vol, accumulate);
  return pmx_check_launch("healpix_finalize");
}

int pmx_launch_healpix_vjp(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, const double box[3],
                           const double obs[3], double rmin, double rmax, int nside, int scheme, int ntiles,
                           const int* rot, const double* shift, const float* map_bar, float* pos_bar, int accumulate,
                           hipStream_t s, const PmxDrift* drift, float* vel_bar) {
  ShellArgs sa = shell_args(p, a, box, obs, rmin, rmax, scheme, ntiles, rot, shift);
  if (drift) {
    sa.v = drift->v;
    sa.dtab = drift->tables;
    sa.ntab = drift->ntab;
    sa.chi_max = drift->chi_max;
    sa.a_lo = drift->a_lo;
    sa.a_hi = drift->a_hi;
    sa.a_now = drift->a_now;
    sa.gp_now = drift->gp_now;
    sa.fbound = drift->fbound;
    sa.ptab = drift->ptab;
    sa.nparam = drift->ptab ? drift->nparam : 0;
    sa.pdots = drift->pdots;
  }
  int rc = 0;
  const HpGeom g = hp_geom_table(p, nside, s, rc);
  if (rc) return rc;
  const double vol = (4.0 * M_PI / (double)g.npix) * (rmax * rmax * rmax - rmin * rmin * rmin) / 3.0;
  hipLaunchKernelGGL(k_healpix_vjp, dim3((unsigned)sa.nbx * sa.nby * sa.nbz), dim3(256), 0, s, a, w, w0, sa, g,
                     map_bar, 1.0 / vol, pos_bar, drift ? vel_bar : nullptr, accumulate);
  return pmx_check_launch("healpix_vjp");
}

int pmx_launch_born(pmx_plan_t* p, const float* maps, int S, long npix, const double* C, int K, float* kappa,
                    int accumulate, hipStream_t s) {
  const int nblk = born_blocks(npix, S);
  double* coef = p->red + kBornCoef;
  hipLaunchKernelGGL(k_map_sums, dim3(nblk, S), dim3(256), 0, s, maps, npix, nblk, p->red);
  PMX_CHECK_LAUNCH("born_sums");
  hipLaunchKernelGGL(k_born_coef, dim3(1), dim3(256), 0, s, p->red, nblk, S, K, npix, C, coef);
  PMX_CHECK_LAUNCH("born_coef");
  hipLaunchKernelGGL(k_born, dim3(grid_for((size_t)npix, 256)), dim3(256), 0, s, maps, npix, S, K, coef, kappa,
                     accumulate);
  return pmx_check_launch("born");
}

int pmx_launch_born_vjp(pmx_plan_t* p, const float* maps, int S, long npix, const double* C, int K,
                        const float* kappa_bar, float* maps_bar, int accumulate, double* dots, hipStream_t s) {
  double* coef = p->red + kBornCoef;
  // means and f of the forward
  int nblk = born_blocks(npix, S);
  hipLaunchKernelGGL(k_map_sums, dim3(nblk, S), dim3(256), 0, s, maps, npix, nblk, p->red);
  PMX_CHECK_LAUNCH("born_sums");
  hipLaunchKernelGGL(k_born_coef, dim3(1), dim3(256), 0, s, p->red, nblk, S, K, npix, C, coef);
  PMX_CHECK_LAUNCH("born_coef");
  nblk = born_blocks(npix, (long)(S + 1) * K);
  hipLaunchKernelGGL(k_born_dots, dim3(nblk, S + 1), dim3(256), 0, s, maps, kappa_bar, npix, S, K, nblk, p->red);
  PMX_CHECK_LAUNCH("born_dots");
  hipLaunchKernelGGL(k_born_vjp_coef, dim3(1), dim3(256), 0, s, p->red, nblk, S, K, npix, C, coef, dots);
  PMX_CHECK_LAUNCH("born_vjp_coef");
  hipLaunchKernelGGL(k_born_vjp, dim3(grid_for((size_t)npix, 256)), dim3(256), 0, s, kappa_bar, npix, S, K, coef,
                     maps_bar, accumulate);
  return pmx_check_launch("born_vjp");
}

int pmx_launch_born_flat(pmx_plan_t* p, const float* maps, int S, int ny, int nx, const double* C, int K,
                         const double* sc, float* kappa, int accumulate, hipStream_t s) {
  const long npix = (long)ny * nx;
  const int nblk = born_blocks(npix, S);
  double* coef = p->red + kBornCoef;
  hipLaunchKernelGGL(k_map_sums, dim3(nblk, S), dim3(256), 0, s, maps, npix, nblk, p->red);
  PMX_CHECK_LAUNCH("born_sums");
  hipLaunchKernelGGL(k_born_coef, dim3(1), dim3(256), 0, s, p->red, nblk, S, K, npix, C, coef);
  PMX_CHECK_LAUNCH("born_coef");
  hipLaunchKernelGGL(k_born_flat, dim3(grid_for((size_t)npix, 256)), dim3(256), 0, s, maps, ny, nx, S, K, coef, sc,
                     kappa, accumulate);
  return pmx_check_launch("born_flat");
}

int pmx_launch_born_flat_vjp(pmx_plan_t* p, const float* maps, int S, int ny, int nx, const double* C, int K,
                             const double* sc, const float* kappa_bar, float* maps_bar, int accumulate, double* dots,
                             hipStream_t s) {
  const long npix = (long)ny * nx;
  double* coef = p->red + kBornCoef;
  int nblk = born_blocks(npix, S);
  hipLaunchKernelGGL(k_map_sums, dim3(nblk, S), dim3(256), 0, s, maps, npix, nblk, p->red);
  PMX_CHECK_LAUNCH("born_sums");
  hipLaunchKernelGGL(k_born_coef, dim3(1), dim3(256), 0, s, p->red, nblk, S, K, npix, C, coef);
  PMX_CHECK_LAUNCH("born_coef");
  nblk = born_blocks(npix, (long)(S + 1) * K + S);
  hipLaunchKernelGGL(k_born_flat_dots, dim3(nblk, S + 1), dim3(256), 0, s, maps, kappa_bar, ny, nx, S, K, nblk, coef,
                     sc, p->red);
  PMX_CHECK_LAUNCH("born_flat_dots");
  hipLaunchKernelGGL(k_born_vjp_coef, dim3(1), dim3(256), 0, s, p->red, nblk, S, K, npix, C, coef, dots);
  PMX_CHECK_LAUNCH("born_vjp_coef");
  if (dots) {
    hipLaunchKernelGGL(k_born_flat_p, dim3(1), dim3(256), 0, s, p->red, nblk, S, K, dots);
    PMX_CHECK_LAUNCH("born_flat_p");
  }
  hipLaunchKernelGGL(k_born_flat_vjp_fill, dim3(grid_for((size_t)S * npix, 256)), dim3(256), 0, s, npix, S, K, coef,
                     maps_bar, accumulate);
  PMX_CHECK_LAUNCH("born_flat_vjp_fill");
  hipLaunchKernelGGL(k_born_flat_scatter, dim3(grid_for((size_t)npix, 256)), dim3(256), 0, s, kappa_bar, ny, nx, S, K,
                     coef, sc, maps_bar);
  return pmx_check_launch("born_flat_scatter");
}
