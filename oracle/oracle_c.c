/* Loops of the CPU oracle in plain C (test infrastructure only, like the rest of oracle/).
 *
 * These are the same computations as the NumPy loops of oracle/pm_oracle.py and oracle/sky_oracle.py,
 * written out per particle so that the oracle reaches the sizes of the real workloads. Every function
 * performs the operations of its NumPy counterpart in the same order and precision, so the two give
 * bitwise equal results (tests/test_oracle.py checks that on random inputs):
 *
 *   cic_paint_{f64,f32}   pm_oracle.cic_paint: weights in the working precision T, deposits summed in
 *                         float64 per corner pass (as np.bincount), corner passes added in x-fastest
 *                         order (jaxpm.painting.cic_paint_dx / cic_paint, called at
 *                         src/jax_fli/_src/fields/_painting.py:66-82)
 *   cic_read_{f64,f32}    pm_oracle.cic_read: per particle, out = out + mesh[corner] * w in T, corners in
 *                         x-fastest order (cic_read_dx / cic_read, fields/particles.py:183-229)
 *   cic_paint_mt,         the CPU baseline's paint and 3-component readout (pm_oracle.cpu_pm_step): one
 *   cic_read3_mt          pass over the particles on `nthreads` OpenMP threads, float32 weights, deposits
 *                         added with atomic float adds (summation order, hence rounding, varies between
 *                         runs: not used where bitwise equality is checked)
 *   healpix_deposit       the per-particle body of pm_oracle.paint_healpix / sky_oracle.paint_healpix_tiled:
 *                         RING angles, ngp pixel or the four bilinear pixels and weights of HEALPix
 *                         get_interpol (Gorski et al. 2005), added to acc in particle order
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (no fast-math, no FMA contraction: the NumPy
 * side rounds every product and sum).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline int64_t pmod(int64_t a, int64_t n) {
  int64_t r = a % n;
  return r < 0 ? r + n : r;
}

/* floor, weights and base index of one coordinate (T = double) */
#define CIC_CORNERS(T, x, lat, mode, base, w0, w1)      \
  do {                                                  \
    const T fl_ = (T)floor((double)(x));                \
    if ((mode) == 0) {                                  \
      const T fr_ = (x) - fl_;                          \
      (w0) = (T)1 - (T)fabs((double)fr_);               \
      (w1) = (T)1 - (T)fabs((double)(fr_ - (T)1));      \
      (base) = (lat) + (int64_t)fl_;                    \
    } else {                                            \
      (w0) = (T)1 - (T)fabs((double)((x) - fl_));       \
      (w1) = (T)1 - (T)fabs((double)((x) - (fl_ + (T)1))); \
      (base) = (int64_t)fl_;                            \
    }                                                   \
  } while (0)

#define DEF_PAINT(NAME, T)                                                                                     \
  int NAME(const T* pos, int64_t npx, int64_t npy, int64_t npz, int mode, int64_t nx, int64_t ny, int64_t nz,  \
           const T* w, T w0s, double* mesh) {                                                                  \
    const int64_t np_ = npx * npy * npz, nc = nx * ny * nz;                                                    \
    const int64_t rx = nx / npx, ry = ny / npy, rz = nz / npz;                                                 \
    int64_t* base = (int64_t*)malloc(sizeof(int64_t) * 3 * np_);                                               \
    T* wt = (T*)malloc(sizeof(T) * 6 * np_);                                                                   \
    double* tmp = (double*)malloc(sizeof(double) * nc);                                                        \
    if (!base || !wt || !tmp) {                                                                                \
      free(base);                                                                                              \
      free(wt);                                                                                                \
      free(tmp);                                                                                               \
      return 1;                                                                                                \
    }                                                                                                          \
    for (int64_t p = 0; p < np_; ++p) {                                                                        \
      const int64_t k = p % npz, j = (p / npz) % npy, i = p / (npz * npy);                                     \
      const int64_t lat[3] = {i * rx, j * ry, k * rz};                                                         \
      for (int d = 0; d < 3; ++d) CIC_CORNERS(T, pos[3 * p + d], lat[d], mode, base[3 * p + d], wt[6 * p + d], \
                                              wt[6 * p + 3 + d]);                                              \
    }                                                                                                          \
    memset(mesh, 0, sizeof(double) * nc);                                                                      \
    for (int o = 0; o < 8; ++o) {                                                                              \
      const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;                                              \
      memset(tmp, 0, sizeof(double) * nc);                                                                     \
      for (int64_t p = 0; p < np_; ++p) {                                                                      \
        const T wx = wt[6 * p + 3 * ox + 0], wy = wt[6 * p + 3 * oy + 1], wz = wt[6 * p + 3 * oz + 2];         \
        const T ww = w ? w[p] : w0s;                                                                           \
        const T v = ((wx * wy) * wz) * ww;                                                                     \
        const int64_t ix = pmod(base[3 * p] + ox, nx), iy = pmod(base[3 * p + 1] + oy, ny),                    \
                      iz = pmod(base[3 * p + 2] + oz, nz);                                                     \
        tmp[(ix * ny + iy) * nz + iz] += (double)v;                                                            \
      }                                                                                                        \
      for (int64_t c = 0; c < nc; ++c) mesh[c] += tmp[c];                                                      \
    }                                                                                                          \
    free(base);                                                                                                \
    free(wt);                                                                                                  \
    free(tmp);                                                                                                 \
    return 0;                                                                                                  \
  }

DEF_PAINT(cic_paint_f64, double)
DEF_PAINT(cic_paint_f32, float)

#define DEF_READ(NAME, T)                                                                                      \
  int NAME(const T* mesh, int ncomp, const T* pos, int64_t npx, int64_t npy, int64_t npz, int mode, int64_t nx, \
           int64_t ny, int64_t nz, T* out) {                                                                   \
    const int64_t np_ = npx * npy * npz;                                                                       \
    const int64_t rx = nx / npx, ry = ny / npy, rz = nz / npz;                                                 \
    if (ncomp < 1 || ncomp > 8) return 1;                                                                      \
    for (int64_t p = 0; p < np_; ++p) {                                                                        \
      const int64_t k = p % npz, j = (p / npz) % npy, i = p / (npz * npy);                                     \
      const int64_t lat[3] = {i * rx, j * ry, k * rz};                                                         \
      int64_t b[3];                                                                                            \
      T w[2][3];                                                                                               \
      for (int d = 0; d < 3; ++d) CIC_CORNERS(T, pos[3 * p + d], lat[d], mode, b[d], w[0][d], w[1][d]);        \
      T acc[8];                                                                                                \
      for (int c = 0; c < ncomp; ++c) acc[c] = (T)0;                                                           \
      for (int o = 0; o < 8; ++o) {                                                                            \
        const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;                                            \
        const T ww = (w[ox][0] * w[oy][1]) * w[oz][2];                                                         \
        const int64_t ix = pmod(b[0] + ox, nx), iy = pmod(b[1] + oy, ny), iz = pmod(b[2] + oz, nz);            \
        const T* m = mesh + ((ix * ny + iy) * nz + iz) * ncomp;                                                \
        for (int c = 0; c < ncomp; ++c) acc[c] = acc[c] + m[c] * ww;                                           \
      }                                                                                                        \
      for (int c = 0; c < ncomp; ++c) out[p * ncomp + c] = acc[c];                                             \
    }                                                                                                          \
    return 0;                                                                                                  \
  }

DEF_READ(cic_read_f64, double)
DEF_READ(cic_read_f32, float)

/* relative-mode paint and readout for the CPU baseline, parallel over particles */
int cic_paint_mt(const float* pos, int64_t npx, int64_t npy, int64_t npz, int64_t nx, int64_t ny, int64_t nz,
                 float* mesh, int nthreads) {
  const int64_t np_ = npx * npy * npz, nc = nx * ny * nz;
  const int64_t rx = nx / npx, ry = ny / npy, rz = nz / npz;
  memset(mesh, 0, sizeof(float) * nc);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t p = 0; p < np_; ++p) {
    const int64_t k = p % npz, j = (p / npz) % npy, i = p / (npz * npy);
    const int64_t lat[3] = {i * rx, j * ry, k * rz};
    int64_t b[3];
    float w[2][3];
    for (int d = 0; d < 3; ++d) CIC_CORNERS(float, pos[3 * p + d], lat[d], 0, b[d], w[0][d], w[1][d]);
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
      const float v = (w[ox][0] * w[oy][1]) * w[oz][2];
      const int64_t c = (pmod(b[0] + ox, nx) * ny + pmod(b[1] + oy, ny)) * nz + pmod(b[2] + oz, nz);
#pragma omp atomic
      mesh[c] += v;
    }
  }
  return 0;
}

int cic_read3_mt(const float* mesh, const float* pos, int64_t npx, int64_t npy, int64_t npz, int64_t nx, int64_t ny,
                 int64_t nz, float* out, int nthreads) {
  const int64_t np_ = npx * npy * npz;
  const int64_t rx = nx / npx, ry = ny / npy, rz = nz / npz;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t p = 0; p < np_; ++p) {
    const int64_t k = p % npz, j = (p / npz) % npy, i = p / (npz * npy);
    const int64_t lat[3] = {i * rx, j * ry, k * rz};
    int64_t b[3];
    float w[2][3];
    for (int d = 0; d < 3; ++d) CIC_CORNERS(float, pos[3 * p + d], lat[d], 0, b[d], w[0][d], w[1][d]);
    float acc[3] = {0.f, 0.f, 0.f};
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
      const float ww = (w[ox][0] * w[oy][1]) * w[oz][2];
      const float* m = mesh + ((pmod(b[0] + ox, nx) * ny + pmod(b[1] + oy, ny)) * nz + pmod(b[2] + oz, nz)) * 3;
      for (int c = 0; c < 3; ++c) acc[c] = acc[c] + m[c] * ww;
    }
    for (int c = 0; c < 3; ++c) out[p * 3 + c] = acc[c];
  }
  return 0;
}

/* ------------------------------------------------------------------------------------ HEALPix RING */
static const double PI = 3.141592653589793;

static int64_t ang2pix(int64_t nside, double z, double phi) {
  const double za = fabs(z);
  double tt = fmod(phi * (2.0 / PI), 4.0);
  if (tt < 0) tt += 4.0;
  const int64_t ncap = 2 * nside * (nside - 1), npix = 12 * nside * nside;
  if (za <= 2.0 / 3.0) {
    const int64_t nl4 = 4 * nside;
    const double t1 = nside * (0.5 + tt), t2 = nside * z * 0.75;
    const int64_t jp = (int64_t)(t1 - t2), jm = (int64_t)(t1 + t2);
    const int64_t ir = nside + 1 + jp - jm;
    const int64_t kshift = 1 - (ir & 1);
    const int64_t t = jp + jm - nside + kshift + 1 + nl4 + nl4;
    const int64_t ip = (t >> 1) % nl4;
    return ncap + (ir - 1) * nl4 + ip;
  }
  const double tp = tt - (double)(int64_t)tt;
  const double tmp = nside * sqrt(3.0 * (1.0 - za));
  const int64_t jp = (int64_t)(tp * tmp), jm = (int64_t)((1.0 - tp) * tmp);
  const int64_t ir = jp + jm + 1;
  const int64_t ip = ((int64_t)(tt * ir)) % (4 * ir);
  return z > 0 ? 2 * ir * (ir - 1) + ip : npix - 2 * ir * (ir + 1) + ip;
}

static int64_t ring_above(int64_t nside, double z) {
  const double az = fabs(z);
  if (az <= 2.0 / 3.0) return (int64_t)(nside * (2.0 - 1.5 * z));
  const int64_t iring = (int64_t)(nside * sqrt(3.0 * (1.0 - az)));
  return z > 0 ? iring : 4 * nside - iring - 1;
}

static void ring_info(int64_t nside, int64_t ring, int64_t* startpix, int64_t* ringpix, double* theta, int* shifted) {
  const int64_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  const double fact2 = 4.0 / npix, fact1 = (2 * nside) * fact2;
  const int64_t northring = ring > 2 * nside ? 4 * nside - ring : ring;
  if (northring < nside) {
    const double tmp = northring * northring * fact2;
    *theta = atan2(sqrt(tmp * (2.0 - tmp)), 1.0 - tmp);
    *ringpix = 4 * northring;
    *shifted = 1;
    *startpix = 2 * northring * (northring - 1);
  } else {
    *theta = acos((2 * nside - northring) * fact1);
    *ringpix = 4 * nside;
    *shifted = ((northring - nside) & 1) == 0;
    *startpix = ncap + (northring - nside) * *ringpix;
  }
  if (northring != ring) {
    *theta = PI - *theta;
    *startpix = npix - *startpix - *ringpix;
  }
}

static void interp_weights(int64_t nside, double theta, double phi, int64_t pix[4], double wgt[4]) {
  const int64_t npix = 12 * nside * nside;
  const double z = cos(theta);
  const int64_t ir1 = ring_above(nside, z), ir2 = ir1 + 1;
  double theta1 = 0.0, theta2 = 0.0;
  for (int q = 0; q < 4; ++q) {
    pix[q] = 0;
    wgt[q] = 0.0;
  }
  for (int s = 0; s < 2; ++s) {
    const int slot = 2 * s;
    const int64_t ir = s == 0 ? ir1 : ir2;
    if ((s == 0 && ir > 0) || (s == 1 && ir < 4 * nside)) {
      int64_t sp, nr;
      double th;
      int shift;
      ring_info(nside, ir, &sp, &nr, &th, &shift);
      if (s == 0)
        theta1 = th;
      else
        theta2 = th;
      const double dphi = 2.0 * PI / nr;
      const double tmp = phi / dphi - 0.5 * shift;
      int64_t i1 = tmp < 0 ? (int64_t)tmp - 1 : (int64_t)tmp;
      const double w1 = (phi - (i1 + 0.5 * shift) * dphi) / dphi;
      int64_t i2 = i1 + 1;
      if (i1 < 0) i1 += nr;
      if (i2 >= nr) i2 -= nr;
      pix[slot] = sp + i1;
      pix[slot + 1] = sp + i2;
      wgt[slot] = 1 - w1;
      wgt[slot + 1] = w1;
    }
  }
  if (ir1 == 0) {
    const double wtheta = theta / theta2;
    wgt[2] *= wtheta;
    wgt[3] *= wtheta;
    const double fac = (1 - wtheta) * 0.25;
    wgt[0] = fac;
    wgt[1] = fac;
    wgt[2] += fac;
    wgt[3] += fac;
    pix[0] = (pix[2] + 2) & 3;
    pix[1] = (pix[3] + 2) & 3;
  } else if (ir2 == 4 * nside) {
    const double wtheta = (theta - theta1) / (PI - theta1);
    wgt[0] *= 1 - wtheta;
    wgt[1] *= 1 - wtheta;
    const double fac = wtheta * 0.25;
    wgt[0] += fac;
    wgt[1] += fac;
    wgt[2] = fac;
    wgt[3] = fac;
    pix[2] = ((pix[0] + 2) & 3) + npix - 4;
    pix[3] = ((pix[1] + 2) & 3) + npix - 4;
  } else {
    const double wtheta = (theta - theta1) / (theta2 - theta1);
    wgt[0] *= 1 - wtheta;
    wgt[1] *= 1 - wtheta;
    wgt[2] *= wtheta;
    wgt[3] *= wtheta;
  }
}

/* Deposits of the selected particles idx[0..nsel) (observer-centred positions X (n x 3), radii r,
 * weights wt) into acc (12 nside^2), in the order of idx. scheme 0 ngp, 1 bilinear. wrap_guard 1:
 * an azimuth that rounds up to 2 pi is taken as 0 (sky_oracle._angles); 0: not (pm_oracle). */
int healpix_deposit(int64_t nside, int scheme, int wrap_guard, const double* X, const double* r, const double* wt,
                    const int64_t* idx, int64_t nsel, double* acc) {
  for (int64_t s = 0; s < nsel; ++s) {
    const int64_t p = idx[s];
    const double x = X[3 * p], y = X[3 * p + 1], z = X[3 * p + 2];
    double phi = atan2(y, x);
    if (phi < 0) phi += 2 * PI;
    if (wrap_guard && !(phi < 2 * PI)) phi = 0.0;
    if (scheme == 0) {
      acc[ang2pix(nside, z / r[p], phi)] += wt[p];
    } else {
      const double theta = atan2(sqrt(x * x + y * y), z);
      int64_t pix[4];
      double w[4];
      interp_weights(nside, theta, phi, pix, w);
      for (int q = 0; q < 4; ++q) acc[pix[q]] += wt[p] * w[q];
    }
  }
  return 0;
}
