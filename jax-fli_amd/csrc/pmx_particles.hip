// Particle-side kernels: CIC paint (with an optional drift fused in front), CIC readout fused with
// the kick and the next drift, HEALPix and flat-sky shell painting, counter-based white noise.
//
// Every particle kernel walks the particle grid in its Lagrangian (i, j, k) order, so a wavefront
// touches a compact patch of the mesh and the 8-corner gathers/scatters stay in L2.
// FP contraction is disabled here so x + v*c and the CIC weights round exactly like the
// unfused float32 expressions of the reference (two roundings, no FMA).
#include "pmx_internal.h"
#include <cmath>

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ int wrap_idx(int i, int n) {
  if ((unsigned)i < (unsigned)n) return i;
  int r = i % n;
  return r < 0 ? r + n : r;
}

struct Corner1D {
  int i0, i1;
  float w0, w1;
};

// relative mode: lattice + floor(d) + {0,1}, weights 1-|frac - {0,1}| (multilinear enmesh)
__device__ __forceinline__ Corner1D corner_rel(int lattice, float d, int n) {
  const float fl = floorf(d);
  const float fr = d - fl;
  const int b = lattice + (int)fl;
  Corner1D c;
  c.i0 = wrap_idx(b, n);
  c.i1 = wrap_idx(b + 1, n);
  c.w0 = 1.0f - fabsf(fr);
  c.w1 = 1.0f - fabsf(fr - 1.0f);
  return c;
}
// absolute mode: floor(x) + {0,1}, weights 1-|x - (floor(x) + {0,1})|
__device__ __forceinline__ Corner1D corner_abs(float x, int n) {
  const float fl = floorf(x);
  const int b = (int)fl;
  Corner1D c;
  c.i0 = wrap_idx(b, n);
  c.i1 = wrap_idx(b + 1, n);
  c.w0 = 1.0f - fabsf(x - fl);
  c.w1 = 1.0f - fabsf(x - (fl + 1.0f));
  return c;
}

struct Cic {
  Corner1D c[3];
};

__device__ __forceinline__ Cic make_cic(const PartArgs& a, size_t p, float3 x, int NX, int NY, int NZ, int rx, int ry,
                                        int rz) {
  Cic q;
  if (a.mode == 0) {
    const int k = (int)(p % a.npz);
    const size_t q2 = p / a.npz;
    const int j = (int)(q2 % a.npy);
    const int i = (int)(q2 / a.npy);
    q.c[0] = corner_rel(i * rx, x.x, NX);
    q.c[1] = corner_rel(j * ry, x.y, NY);
    q.c[2] = corner_rel(k * rz, x.z, NZ);
  } else {
    q.c[0] = corner_abs(x.x, NX);
    q.c[1] = corner_abs(x.y, NY);
    q.c[2] = corner_abs(x.z, NZ);
  }
  return q;
}

__device__ __forceinline__ float3 ld3(const float* a, size_t p) {
  return make_float3(a[3 * p], a[3 * p + 1], a[3 * p + 2]);
}
__device__ __forceinline__ void st3(float* a, size_t p, float3 v) {
  a[3 * p] = v.x;
  a[3 * p + 1] = v.y;
  a[3 * p + 2] = v.z;
}
__device__ __forceinline__ float3 axpy3(float3 x, float3 v, float c) {  // x + v*c, two roundings
  return make_float3(x.x + v.x * c, x.y + v.y * c, x.z + v.z * c);
}

// Contiguous particle chunk of this workgroup. Hardware dispatches workgroup b to XCD b % 8; the
// logical id remaps so each XCD walks one contiguous slab of the Lagrangian grid and the mesh
// planes it gathers from / scatters to stay in that XCD's L2.
struct Chunk {
  size_t begin, end;
};
__device__ __forceinline__ unsigned xcd_logical(unsigned b, unsigned G) {
  return (G % 8u == 0u) ? (b % 8u) * (G / 8u) + b / 8u : b;
}
__device__ __forceinline__ Chunk my_chunk(size_t n) {
  const unsigned G = gridDim.x;
  const size_t per = (n + G - 1) / G;
  const size_t beg = min(n, (size_t)xcd_logical(blockIdx.x, G) * per);
  return Chunk{beg, min(n, beg + per)};
}

__global__ void k_fill(float* __restrict__ a, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}
__global__ void k_fill_d(double* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0;
}

// CIC paint. Optional fused drift: x <- x + v*cd (written back) before painting.
__global__ void __launch_bounds__(256) k_paint(PartArgs a, const float* __restrict__ w, float w0, float cd,
                                               float* __restrict__ mesh, int NX, int NY, int NZ, int rx, int ry,
                                               int rz) {
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  for (size_t p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x) {
    float3 x = ld3(a.x, p);
    if (a.v != nullptr) {
      x = axpy3(x, ld3(a.v, p), cd);
      st3(a.x, p, x);
    }
    const float wt = w ? w[p] : w0;
    const Cic q = make_cic(a, p, x, NX, NY, NZ, rx, ry, rz);
#pragma unroll
    for (int cx = 0; cx < 2; ++cx) {
      const float wx = cx ? q.c[0].w1 : q.c[0].w0;
      const size_t bx = (size_t)(cx ? q.c[0].i1 : q.c[0].i0) * NY;
#pragma unroll
      for (int cy = 0; cy < 2; ++cy) {
        const float wxy = wx * (cy ? q.c[1].w1 : q.c[1].w0);
        const size_t bxy = (bx + (cy ? q.c[1].i1 : q.c[1].i0)) * NZ;
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
          const float wv = (wxy * (cz ? q.c[2].w1 : q.c[2].w0)) * wt;
          unsafeAtomicAdd(mesh + bxy + (cz ? q.c[2].i1 : q.c[2].i0), wv);
        }
      }
    }
  }
}

// Owner-computes CIC paint for relative mode (particles on their Lagrangian lattice): one
// workgroup per mesh tile (BX, BY, BZ) accumulates in LDS every deposit that lands in its tile from
// the particles whose lattice lies within the tile grown by H + 1 cells, then writes the tile with
// plain stores -- every mesh cell is written exactly once, no global atomics and no zero fill.
// A particle with |floor(dx)| > H along any axis ("escape") is skipped by all tiles; its home tile
// records it in its own segment of the escape list (LDS counter, no global atomics) and
// k_paint_escapes deposits those with atomics afterwards.
constexpr int PT_MAX = 16 * 16 * 32;
constexpr int PT_UNROLL = 4;

__device__ __forceinline__ int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }
__device__ __forceinline__ int wrap_near(int i, int n) {
  if (i < 0) i += n;
  else if (i >= n) i -= n;
  if ((unsigned)i >= (unsigned)n) {
    i %= n;
    i += i < 0 ? n : 0;
  }
  return i;
}

__global__ void __launch_bounds__(256) k_paint_tiles(PartArgs a, const float* __restrict__ w, float w0,
                                                     float* __restrict__ mesh, int NX, int NY, int NZ, int rx,
                                                     int ry, int rz, int BX, int BY, int BZ, int H, int accumulate,
                                                     unsigned* __restrict__ esc_count, unsigned* __restrict__ esc_list,
                                                     int esc_cap) {
  __shared__ float tile[PT_MAX];
  __shared__ unsigned s_nesc;
  const int ntx = NX / BX, nty = NY / BY, ntz = NZ / BZ;
  const unsigned L = xcd_logical(blockIdx.x, gridDim.x);
  const int tz = L % ntz, ty = (L / ntz) % nty, tx = L / (ntz * nty);
  const int T0x = tx * BX, T0y = ty * BY, T0z = tz * BZ;
  const int ncell = BX * BY * BZ;
  for (int i = threadIdx.x; i < ncell; i += blockDim.x) tile[i] = 0.f;
  if (threadIdx.x == 0) s_nesc = 0u;
  __syncthreads();
  // lattice i*r in [T0 - 1 - H, T0 + B - 1 + H] (unwrapped)
  const int ilo = -floordiv(-(T0x - 1 - H), rx), ihi = floordiv(T0x + BX - 1 + H, rx);
  const int jlo = -floordiv(-(T0y - 1 - H), ry), jhi = floordiv(T0y + BY - 1 + H, ry);
  const int klo = -floordiv(-(T0z - 1 - H), rz), khi = floordiv(T0z + BZ - 1 + H, rz);
  const int ci = ihi - ilo + 1, cj = jhi - jlo + 1, ck = khi - klo + 1;
  const float fH = (float)H;
  // division-free walk over (ii, jj, kk), kk fastest, stride blockDim.x
  const int stride = blockDim.x;
  const int dk = stride % ck, drow = stride / ck;
  const int dj = drow % cj, di = drow / cj;
  int kk = threadIdx.x % ck;
  int jj = (threadIdx.x / ck) % cj;
  int ii = threadIdx.x / (ck * cj);
  unsigned* my_esc = esc_list + (size_t)L * esc_cap;
  while (ii < ci) {
    size_t pidx[PT_UNROLL];
    int li[PT_UNROLL], lj[PT_UNROLL], lk[PT_UNROLL];
    bool ok[PT_UNROLL];
#pragma unroll
    for (int u = 0; u < PT_UNROLL; ++u) {
      ok[u] = ii < ci;
      li[u] = ilo + ii;
      lj[u] = jlo + jj;
      lk[u] = klo + kk;
      pidx[u] = ((size_t)wrap_near(li[u], a.npx) * a.npy + wrap_near(lj[u], a.npy)) * a.npz + wrap_near(lk[u], a.npz);
      kk += dk;
      int carry = 0;
      if (kk >= ck) { kk -= ck; carry = 1; }
      jj += dj + carry;
      carry = 0;
      if (jj >= cj) { jj -= cj; carry = 1; }
      ii += di + carry;
    }
    float3 xv[PT_UNROLL];
#pragma unroll
    for (int u = 0; u < PT_UNROLL; ++u) xv[u] = ok[u] ? ld3(a.x, pidx[u]) : make_float3(0.f, 0.f, 0.f);
#pragma unroll
    for (int u = 0; u < PT_UNROLL; ++u) {
      if (!ok[u]) continue;
      const float3 x = xv[u];
      const float fx = floorf(x.x), fy = floorf(x.y), fz = floorf(x.z);
      const bool esc = !(fx >= -fH && fx <= fH && fy >= -fH && fy <= fH && fz >= -fH && fz <= fH);
      if (esc) {
        const bool home = (li[u] * rx >= T0x && li[u] * rx < T0x + BX && lj[u] * ry >= T0y && lj[u] * ry < T0y + BY &&
                           lk[u] * rz >= T0z && lk[u] * rz < T0z + BZ);
        if (home) my_esc[atomicAdd(&s_nesc, 1u)] = (unsigned)pidx[u];
        continue;
      }
      const float wt = w ? w[pidx[u]] : w0;
      const float frx = x.x - fx, fry = x.y - fy, frz = x.z - fz;
      const float wx[2] = {1.0f - fabsf(frx), 1.0f - fabsf(frx - 1.0f)};
      const float wy[2] = {1.0f - fabsf(fry), 1.0f - fabsf(fry - 1.0f)};
      const float wz[2] = {1.0f - fabsf(frz), 1.0f - fabsf(frz - 1.0f)};
      const int bx = li[u] * rx + (int)fx - T0x, by = lj[u] * ry + (int)fy - T0y, bz = lk[u] * rz + (int)fz - T0z;
#pragma unroll
      for (int cx = 0; cx < 2; ++cx) {
        const int ix = bx + cx;
        if (ix < 0 || ix >= BX) continue;
#pragma unroll
        for (int cy = 0; cy < 2; ++cy) {
          const int iy = by + cy;
          if (iy < 0 || iy >= BY) continue;
          const float wxy = wx[cx] * wy[cy];
#pragma unroll
          for (int cz = 0; cz < 2; ++cz) {
            const int iz = bz + cz;
            if (iz < 0 || iz >= BZ) continue;
            atomicAdd(&tile[(ix * BY + iy) * BZ + iz], (wxy * wz[cz]) * wt);
          }
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) esc_count[L] = s_nesc;
  if (accumulate) {
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
      const int iz = c % BZ, iy = (c / BZ) % BY, ix = c / (BZ * BY);
      float* dst = mesh + ((size_t)(T0x + ix) * NY + (T0y + iy)) * NZ + (T0z + iz);
      *dst = *dst + tile[c];
    }
  } else {
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
      const int iz = c % BZ, iy = (c / BZ) % BY, ix = c / (BZ * BY);
      mesh[((size_t)(T0x + ix) * NY + (T0y + iy)) * NZ + (T0z + iz)] = tile[c];
    }
  }
}

// escapes of tile blockIdx.x: segment [t * esc_cap, t * esc_cap + esc_count[t])
__global__ void __launch_bounds__(256) k_paint_escapes(PartArgs a, const float* __restrict__ w, float w0,
                                                       float* __restrict__ mesh, int NX, int NY, int NZ, int rx, int ry,
                                                       int rz, const unsigned* __restrict__ esc_count,
                                                       const unsigned* __restrict__ esc_list, int esc_cap) {
  const unsigned n = esc_count[blockIdx.x];
  const unsigned* seg = esc_list + (size_t)blockIdx.x * esc_cap;
  for (unsigned e = threadIdx.x; e < n; e += blockDim.x) {
    const size_t p = seg[e];
    const float3 x = ld3(a.x, p);
    const float wt = w ? w[p] : w0;
    const Cic q = make_cic(a, p, x, NX, NY, NZ, rx, ry, rz);
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int cx = o & 1, cy = (o >> 1) & 1, cz = (o >> 2) & 1;
      const float wv = (((cx ? q.c[0].w1 : q.c[0].w0) * (cy ? q.c[1].w1 : q.c[1].w0)) * (cz ? q.c[2].w1 : q.c[2].w0)) * wt;
      const size_t idx =
          (((size_t)(cx ? q.c[0].i1 : q.c[0].i0) * NY + (cy ? q.c[1].i1 : q.c[1].i0)) * NZ + (cz ? q.c[2].i1 : q.c[2].i0));
      unsafeAtomicAdd(mesh + idx, wv);
    }
  }
}

// x <- x + v*c
__global__ void __launch_bounds__(256) k_drift(float* __restrict__ x, const float* __restrict__ v, size_t n, float c) {
  const Chunk ch = my_chunk(n);
  for (size_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) x[i] = x[i] + v[i] * c;
}

__global__ void __launch_bounds__(256) k_paint_escapes(PartArgs a, const float* __restrict__ w, float w0,
                                                       float* __restrict__ mesh, int NX, int NY, int NZ, int rx, int ry,
                                                       int rz, const unsigned* __restrict__ esc_count,
                                                       const unsigned* __restrict__ esc_list) {
  const unsigned n = *esc_count;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const size_t p = esc_list[e];
    const float3 x = ld3(a.x, p);
    const float wt = w ? w[p] : w0;
    const Cic q = make_cic(a, p, x, NX, NY, NZ, rx, ry, rz);
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int cx = o & 1, cy = (o >> 1) & 1, cz = (o >> 2) & 1;
      const float wv = (((cx ? q.c[0].w1 : q.c[0].w0) * (cy ? q.c[1].w1 : q.c[1].w0)) * (cz ? q.c[2].w1 : q.c[2].w0)) * wt;
      const size_t idx =
          (((size_t)(cx ? q.c[0].i1 : q.c[0].i0) * NY + (cy ? q.c[1].i1 : q.c[1].i0)) * NZ + (cz ? q.c[2].i1 : q.c[2].i0));
      unsafeAtomicAdd(mesh + idx, wv);
    }
  }
}

// CIC readout of an NC-component interleaved mesh. what: bit0 store F in out, bit1 kick
// v <- v + F*ck (NC == 3), bit2 drift x <- x + v*cd after the kick. Corners are summed
// sequentially in the order x-fastest (0,0,0),(1,0,0),(0,1,0),(1,1,0),(0,0,1),...
template <int NC>
__global__ void __launch_bounds__(256) k_readout(PartArgs a, const float* __restrict__ mesh, float* __restrict__ out,
                                                 float ck, float cd, int what, int NX, int NY, int NZ, int rx,
                                                 int ry, int rz) {
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  for (size_t p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x) {
    float3 x = ld3(a.x, p);
    const Cic q = make_cic(a, p, x, NX, NY, NZ, rx, ry, rz);
    float f[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int cx = o & 1, cy = (o >> 1) & 1, cz = (o >> 2) & 1;
      const float wv = ((cx ? q.c[0].w1 : q.c[0].w0) * (cy ? q.c[1].w1 : q.c[1].w0)) * (cz ? q.c[2].w1 : q.c[2].w0);
      const size_t idx =
          (((size_t)(cx ? q.c[0].i1 : q.c[0].i0) * NY + (cy ? q.c[1].i1 : q.c[1].i0)) * NZ + (cz ? q.c[2].i1 : q.c[2].i0));
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = f[c] + mesh[idx * NC + c] * wv;
    }
    if (what & 1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) out[p * NC + c] = f[c];
    }
    if constexpr (NC == 3) {
      if (what & 2) {
        float3 v = ld3(a.v, p);
        v = axpy3(v, make_float3(f[0], f[1], f[2]), ck);
        st3(a.v, p, v);
        if (what & 4) {
          x = axpy3(x, v, cd);
          st3(a.x, p, x);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- HEALPix (RING scheme)
// Restates the published HEALPix pixelisation (Gorski et al. 2005) and its bilinear ring
// interpolation: loc2pix, ring_above, get_ring_info2, get_interpol of the HEALPix C++ library.
struct HpGeom {
  int nside;
  long npix, ncap;
  double fact1, fact2;
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

__device__ __forceinline__ long hp_ring_above(const HpGeom& g, double z) {
  const double az = fabs(z);
  if (az <= 2.0 / 3.0) return (long)(g.nside * (2.0 - 1.5 * z));
  const long iring = (long)(g.nside * sqrt(3.0 * (1.0 - az)));
  return (z > 0) ? iring : 4 * (long)g.nside - iring - 1;
}

__device__ __forceinline__ void hp_ring_info(const HpGeom& g, long ring, long& startpix, long& ringpix, double& theta,
                                             bool& shifted) {
  const long ns = g.nside;
  const long northring = (ring > 2 * ns) ? 4 * ns - ring : ring;
  if (northring < ns) {
    const double tmp = (double)northring * northring * g.fact2;
    const double costheta = 1.0 - tmp;
    const double sintheta = sqrt(tmp * (2.0 - tmp));
    theta = atan2(sintheta, costheta);
    ringpix = 4 * northring;
    shifted = true;
    startpix = 2 * northring * (northring - 1);
  } else {
    theta = acos((2 * ns - northring) * g.fact1);
    ringpix = 4 * ns;
    shifted = ((northring - ns) & 1) == 0;
    startpix = g.ncap + (northring - ns) * ringpix;
  }
  if (northring != ring) {
    theta = M_PI - theta;
    startpix = g.npix - startpix - ringpix;
  }
}

__device__ __forceinline__ void hp_interpol(const HpGeom& g, double theta, double phi, long pix[4], double wgt[4]) {
  const double z = cos(theta);
  const long ir1 = hp_ring_above(g, z);
  const long ir2 = ir1 + 1;
  const double twopi = 2.0 * M_PI;
  double theta1 = 0, theta2 = 0;
  long sp, nr;
  bool shift;
  if (ir1 > 0) {
    hp_ring_info(g, ir1, sp, nr, theta1, shift);
    const double dphi = twopi / nr;
    const double tmp = phi / dphi - 0.5 * shift;
    long i1 = (tmp < 0) ? (long)tmp - 1 : (long)tmp;
    const double w1 = (phi - (i1 + 0.5 * shift) * dphi) / dphi;
    long i2 = i1 + 1;
    if (i1 < 0) i1 += nr;
    if (i2 >= nr) i2 -= nr;
    pix[0] = sp + i1;
    pix[1] = sp + i2;
    wgt[0] = 1 - w1;
    wgt[1] = w1;
  }
  if (ir2 < 4 * (long)g.nside) {
    hp_ring_info(g, ir2, sp, nr, theta2, shift);
    const double dphi = twopi / nr;
    const double tmp = phi / dphi - 0.5 * shift;
    long i1 = (tmp < 0) ? (long)tmp - 1 : (long)tmp;
    const double w1 = (phi - (i1 + 0.5 * shift) * dphi) / dphi;
    long i2 = i1 + 1;
    if (i1 < 0) i1 += nr;
    if (i2 >= nr) i2 -= nr;
    pix[2] = sp + i1;
    pix[3] = sp + i2;
    wgt[2] = 1 - w1;
    wgt[3] = w1;
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
  } else if (ir2 == 4 * (long)g.nside) {
    const double wtheta = (theta - theta1) / (M_PI - theta1);
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
    const double wtheta = (theta - theta1) / (theta2 - theta1);
    wgt[0] *= (1 - wtheta);
    wgt[1] *= (1 - wtheta);
    wgt[2] *= wtheta;
    wgt[3] *= wtheta;
  }
}

struct ShellArgs {
  double box[3];    // Mpc/h
  double obs[3];    // observer, Mpc/h
  double rmin, rmax;
  int scheme;       // 0 ngp, 1 bilinear
};

__device__ __forceinline__ double grid_pos(const PartArgs& a, size_t p, int c, float xc, int n, int r) {
  double g = (double)xc;
  if (a.mode == 0) {
    const int k = (int)(p % a.npz);
    const size_t q2 = p / a.npz;
    const int j = (int)(q2 % a.npy);
    const int i = (int)(q2 / a.npy);
    g += (double)((c == 0 ? i : (c == 1 ? j : k)) * r);
  }
  g -= floor(g / n) * n;  // periodic box
  return g;
}

__global__ void __launch_bounds__(256) k_paint_healpix(PartArgs a, const float* __restrict__ w, float w0, ShellArgs sa,
                                                       HpGeom g, double* __restrict__ acc, int NX, int NY, int NZ,
                                                       int rx, int ry, int rz) {
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  for (size_t p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x) {
    const float3 x = ld3(a.x, p);
    const double gx = grid_pos(a, p, 0, x.x, NX, rx), gy = grid_pos(a, p, 1, x.y, NY, ry),
                 gz = grid_pos(a, p, 2, x.z, NZ, rz);
    const double X = gx * sa.box[0] / NX - sa.obs[0];
    const double Y = gy * sa.box[1] / NY - sa.obs[1];
    const double Z = gz * sa.box[2] / NZ - sa.obs[2];
    const double r = sqrt(X * X + Y * Y + Z * Z);
    if (!(r > sa.rmin && r <= sa.rmax)) continue;
    const double wt = w ? (double)w[p] : (double)w0;
    const double phi0 = atan2(Y, X);
    double phi = phi0 < 0 ? phi0 + 2.0 * M_PI : phi0;
    if (!(phi < 2.0 * M_PI)) phi = 0.0;  // -tiny + 2 pi can round up to 2 pi
    if (sa.scheme == 0) {
      const long pix = hp_ang2pix(g, Z / r, phi);
      atomicAdd(acc + pix, wt);
    } else {
      const double theta = atan2(sqrt(X * X + Y * Y), Z);
      long pix[4];
      double wg[4];
      hp_interpol(g, theta, phi, pix, wg);
#pragma unroll
      for (int m = 0; m < 4; ++m) atomicAdd(acc + pix[m], wt * wg[m]);
    }
  }
}

__global__ void k_finalize_map(const double* __restrict__ acc, float* __restrict__ map, long n, double scale, int accumulate) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float v = (float)(acc[i] * scale);
    map[i] = accumulate ? map[i] + v : v;
  }
}

// flat-sky: absolute grid positions, slab center-width/2 < z <= center+width/2 (grid units),
// (x, y) scaled to the flat grid and wrapped; 4-corner CIC
struct FlatArgs {
  int fnx, fny;
  float sxf, syf;          // fnx/nx, fny/ny
  float zlo, zhi;          // slab bounds in grid units
};
__global__ void __launch_bounds__(256) k_paint_flat(PartArgs a, const float* __restrict__ w, float w0, FlatArgs fa,
                                                    double* __restrict__ acc, int rx, int ry, int rz) {
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  for (size_t p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x) {
    float3 x = ld3(a.x, p);
    if (a.mode == 0) {
      const int k = (int)(p % a.npz);
      const size_t q2 = p / a.npz;
      const int j = (int)(q2 % a.npy);
      const int i = (int)(q2 / a.npy);
      x = make_float3(x.x + (float)(i * rx), x.y + (float)(j * ry), x.z + (float)(k * rz));
    }
    if (!(x.z > fa.zlo && x.z <= fa.zhi)) continue;
    const float wt = w ? w[p] : w0;
    float px = x.x * fa.sxf, py = x.y * fa.syf;
    px = px - floorf(px / fa.fnx) * fa.fnx;
    py = py - floorf(py / fa.fny) * fa.fny;
    const Corner1D cx = corner_abs(px, fa.fnx), cy = corner_abs(py, fa.fny);
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      const int ox = o & 1, oy = o >> 1;
      const float wv = ((ox ? cx.w1 : cx.w0) * (oy ? cy.w1 : cy.w0)) * wt;
      atomicAdd(acc + (size_t)(ox ? cx.i1 : cx.i0) * fa.fny + (oy ? cy.i1 : cy.i0), (double)wv);
    }
  }
}

// ---------------------------------------------------------------- Philox4x32-10 white noise
// Salmon et al. 2011 (Random123). counter = (idx/4 lo, idx/4 hi, 0, 0), key = (seed lo, seed hi).
// Box-Muller on u = ((r >> 8) + 1) 2^-24 in (0, 1], evaluated in float64 then rounded.
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__global__ void k_philox_raw(uint32_t* __restrict__ out, size_t nblocks, uint32_t k0, uint32_t k1) {
  for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < nblocks; b += (size_t)gridDim.x * blockDim.x) {
    uint32_t c[4] = {(uint32_t)b, (uint32_t)(b >> 32), 0u, 0u};
    philox10(c, k0, k1);
    for (int m = 0; m < 4; ++m) out[4 * b + m] = c[m];
  }
}

__global__ void k_white_noise(float* __restrict__ out, size_t n, uint32_t k0, uint32_t k1) {
  const size_t nb = (n + 3) / 4;
  for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < nb; b += (size_t)gridDim.x * blockDim.x) {
    uint32_t c[4] = {(uint32_t)b, (uint32_t)(b >> 32), 0u, 0u};
    philox10(c, k0, k1);
    double u[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) u[m] = (double)((c[m] >> 8) + 1u) * (1.0 / 16777216.0);
    double g[4];
    const double r0 = sqrt(-2.0 * log(u[0])), r1 = sqrt(-2.0 * log(u[2]));
    g[0] = r0 * cos(2.0 * M_PI * u[1]);
    g[1] = r0 * sin(2.0 * M_PI * u[1]);
    g[2] = r1 * cos(2.0 * M_PI * u[3]);
    g[3] = r1 * sin(2.0 * M_PI * u[3]);
#pragma unroll
    for (int m = 0; m < 4; ++m)
      if (4 * b + m < n) out[4 * b + m] = (float)g[m];
  }
}

// dx = D * F ; p = c * dx  (first-order LPT at the lattice, F the unit-scale force mesh)
__global__ void k_lpt1(const float* __restrict__ F, float* __restrict__ dx, float* __restrict__ pv, size_t n, float D,
                       float c) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float d = F[i] * D;
    dx[i] = d;
    pv[i] = d * c;
  }
}

inline int grid_for(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  const size_t gmax = 1u << 20;
  return (int)(g < 1 ? 1 : (g > gmax ? gmax : g));
}
// particle-chunk grids: a multiple of 8 workgroups (one slab per XCD), ~1-4 particles per thread
inline int grid_particles(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  g = (g + 7) / 8 * 8;
  return (int)(g < 8 ? 8 : g);
}

}  // namespace

// ------------------------------------------------------------------ launchers used by pmx_api.hip
int pmx_launch_fill(float* a, size_t n, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, a, n, v);
  return pmx_check_launch("fill");
}
int pmx_launch_fill_d(double* a, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_d, dim3(grid_for(n, 256)), dim3(256), 0, s, a, n);
  return pmx_check_launch("fill_d");
}

static void ratios(pmx_plan_t* p, const PartArgs& a, int& rx, int& ry, int& rz) {
  rx = p->nx / a.npx;
  ry = p->ny / a.npy;
  rz = p->nz / a.npz;
}

// Paint into `mesh`. Relative mode uses the owner-computes tiled kernel (+ escape pass) unless
// `atomic` is set; absolute mode always scatters with atomics. accumulate=0 overwrites the mesh.
// A pre-drift (a.v != null) runs as its own pass first (tiles read each particle several times).
int pmx_launch_paint(pmx_plan_t* p, const PartArgs& a_in, const float* w, float w0, float cd_pre, float* mesh,
                     int accumulate, int atomic, hipStream_t s) {
  int rx, ry, rz;
  PartArgs a = a_in;
  ratios(p, a, rx, ry, rz);
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  if (a.v != nullptr) {
    hipLaunchKernelGGL(k_drift, dim3(grid_particles(3 * np)), dim3(256), 0, s, a.x, a.v, 3 * np, cd_pre);
    int rc = pmx_check_launch("drift");
    if (rc) return rc;
    a.v = nullptr;
  }
  if (a.mode != 0 || atomic) {
    if (!accumulate) {
      int rc = pmx_launch_fill(mesh, p->ncell, 0.f, s);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_paint, dim3(grid_particles(np)), dim3(256), 0, s, a, w, w0, 0.f, mesh, p->nx, p->ny, p->nz, rx,
                       ry, rz);
    return pmx_check_launch("paint");
  }
  const int BX = p->nx < 16 ? p->nx : 16, BY = p->ny < 16 ? p->ny : 16, BZ = p->nz < 32 ? p->nz : 32;
  const int ntiles = (p->nx / BX) * (p->ny / BY) * (p->nz / BZ);
  const int cap = ((BX + rx - 1) / rx) * ((BY + ry - 1) / ry) * ((BZ + rz - 1) / rz);
  const size_t need = (size_t)ntiles * cap;
  if (p->esc_cap < need || p->esc_tiles < (size_t)ntiles) {
    if (p->esc_list) (void)hipFree(p->esc_list);
    if (p->esc_count) (void)hipFree(p->esc_count);
    p->esc_list = nullptr;
    p->esc_count = nullptr;
    p->esc_cap = 0;
    p->esc_tiles = 0;
    hipError_t e = hipMalloc((void**)&p->esc_list, need * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc((void**)&p->esc_count, (size_t)ntiles * sizeof(unsigned));
    if (e != hipSuccess) return pmx_set_error(4, std::string("escape list: ") + hipGetErrorString(e));
    p->esc_cap = need;
    p->esc_tiles = ntiles;
  }
  p->esc_ntiles_last = ntiles;
  hipLaunchKernelGGL(k_paint_tiles, dim3(ntiles), dim3(256), 0, s, a, w, w0, mesh, p->nx, p->ny, p->nz, rx, ry, rz, BX,
                     BY, BZ, PAINT_HALO, accumulate, p->esc_count, p->esc_list, cap);
  int rc = pmx_check_launch("paint_tiles");
  if (rc) return rc;
  hipLaunchKernelGGL(k_paint_escapes, dim3(ntiles), dim3(256), 0, s, a, w, w0, mesh, p->nx, p->ny, p->nz, rx, ry, rz,
                     p->esc_count, p->esc_list, cap);
  return pmx_check_launch("paint_escapes");
}

int pmx_launch_readout(pmx_plan_t* p, const PartArgs& a, const float* mesh, int ncomp, float* out, float ck,
                       float cd_post, int what, hipStream_t s) {
  int rx, ry, rz;
  ratios(p, a, rx, ry, rz);
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  if (ncomp == 3)
    hipLaunchKernelGGL(k_readout<3>, dim3(grid_particles(np)), dim3(256), 0, s, a, mesh, out, ck, cd_post, what, p->nx,
                       p->ny, p->nz, rx, ry, rz);
  else
    hipLaunchKernelGGL(k_readout<1>, dim3(grid_particles(np)), dim3(256), 0, s, a, mesh, out, ck, cd_post, what & 1,
                       p->nx, p->ny, p->nz, rx, ry, rz);
  return pmx_check_launch("readout");
}

int pmx_launch_healpix(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, const double box[3],
                       const double obs[3], double rmin, double rmax, int nside, int scheme, double* acc,
                       hipStream_t s) {
  int rx, ry, rz;
  ratios(p, a, rx, ry, rz);
  ShellArgs sa;
  for (int c = 0; c < 3; ++c) {
    sa.box[c] = box[c];
    sa.obs[c] = obs[c];
  }
  sa.rmin = rmin;
  sa.rmax = rmax;
  sa.scheme = scheme;
  HpGeom g;
  g.nside = nside;
  g.npix = 12L * nside * nside;
  g.ncap = 2L * nside * (nside - 1);
  g.fact2 = 4.0 / (double)g.npix;
  g.fact1 = (double)(2 * nside) * g.fact2;
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  hipLaunchKernelGGL(k_paint_healpix, dim3(grid_particles(np)), dim3(256), 0, s, a, w, w0, sa, g, acc, p->nx, p->ny,
                     p->nz, rx, ry, rz);
  return pmx_check_launch("paint_healpix");
}

int pmx_launch_finalize_map(const double* acc, float* map, long n, double scale, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_map, dim3(grid_for(n, 256)), dim3(256), 0, s, acc, map, n, scale, accumulate);
  return pmx_check_launch("finalize_map");
}

int pmx_launch_flat(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, int fnx, int fny, float zlo, float zhi,
                    double* acc, hipStream_t s) {
  int rx, ry, rz;
  ratios(p, a, rx, ry, rz);
  FlatArgs fa;
  fa.fnx = fnx;
  fa.fny = fny;
  fa.sxf = (float)fnx / (float)p->nx;
  fa.syf = (float)fny / (float)p->ny;
  fa.zlo = zlo;
  fa.zhi = zhi;
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  hipLaunchKernelGGL(k_paint_flat, dim3(grid_particles(np)), dim3(256), 0, s, a, w, w0, fa, acc, rx, ry, rz);
  return pmx_check_launch("paint_flat");
}

int pmx_launch_philox_raw(uint32_t* out, size_t nblocks, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_philox_raw, dim3(grid_for(nblocks, 256)), dim3(256), 0, s, out, nblocks, (uint32_t)seed,
                     (uint32_t)(seed >> 32));
  return pmx_check_launch("philox_raw");
}

int pmx_launch_white_noise(float* out, size_t n, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_white_noise, dim3(grid_for((n + 3) / 4, 256)), dim3(256), 0, s, out, n, (uint32_t)seed,
                     (uint32_t)(seed >> 32));
  return pmx_check_launch("white_noise");
}

int pmx_launch_lpt1(const float* F, float* dx, float* pv, size_t n, float D, float c, hipStream_t s) {
  hipLaunchKernelGGL(k_lpt1, dim3(grid_for(n, 256)), dim3(256), 0, s, F, dx, pv, n, D, c);
  return pmx_check_launch("lpt1");
}
