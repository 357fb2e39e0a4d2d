// Particle-side kernels: CIC paint (owner-computes tiles + escape pass, or one-pass atomic scatter),
// CIC readout fused with the kick and the next drift, HEALPix and flat-sky shell painting,
// counter-based white noise.
//
// Every particle kernel walks the particle grid in its Lagrangian (i, j, k) order, so a wavefront
// touches a compact patch of the mesh and the 8-corner gathers/scatters stay in L2. Index math is
// 32-bit (meshes up to 1024^3 cells, x3 components, fit) and the lattice coordinates of a thread's
// particles are advanced incrementally instead of being divided out of the flat index.
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

// The 8 CIC corners of one particle: mesh cell offsets and weights, x-fastest order
// (0,0,0),(1,0,0),(0,1,0),(1,1,0),(0,0,1),...
struct Cic {
  unsigned off[8];
  float w[8];
};

__device__ __forceinline__ Cic cic_from(Corner1D cx, Corner1D cy, Corner1D cz, int NY, int NZ) {
  const unsigned nyz = (unsigned)NY * NZ;
  const unsigned x0 = cx.i0 * nyz, x1 = cx.i1 * nyz;
  const unsigned y0 = (unsigned)cy.i0 * NZ, y1 = (unsigned)cy.i1 * NZ;
  Cic q;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = (o >> 2) & 1;
    q.off[o] = (ox ? x1 : x0) + (oy ? y1 : y0) + (unsigned)(oz ? cz.i1 : cz.i0);
    q.w[o] = ((ox ? cx.w1 : cx.w0) * (oy ? cy.w1 : cy.w0)) * (oz ? cz.w1 : cz.w0);
  }
  return q;
}

// Lattice coordinates (i, j, k) of a thread's current particle, advanced by a fixed stride.
struct LatticeWalk {
  int i, j, k, di, dj, dk, npy, npz;
  __device__ __forceinline__ void init(unsigned p, unsigned stride, int ny, int nz) {
    npy = ny;
    npz = nz;
    k = (int)(p % (unsigned)nz);
    const unsigned q = p / (unsigned)nz;
    j = (int)(q % (unsigned)ny);
    i = (int)(q / (unsigned)ny);
    dk = (int)(stride % (unsigned)nz);
    const unsigned s = stride / (unsigned)nz;
    dj = (int)(s % (unsigned)ny);
    di = (int)(s / (unsigned)ny);
  }
  __device__ __forceinline__ void step() {
    k += dk;
    int c = 0;
    if (k >= npz) { k -= npz; c = 1; }
    j += dj + c;
    c = 0;
    if (j >= npy) { j -= npy; c = 1; }
    i += di + c;
  }
};

__device__ __forceinline__ Cic particle_cic(const PartArgs& a, const LatticeWalk& lw, float3 x, int NX, int NY, int NZ,
                                            int rx, int ry, int rz) {
  if (a.mode == 0)
    return cic_from(corner_rel(lw.i * rx, x.x, NX), corner_rel(lw.j * ry, x.y, NY), corner_rel(lw.k * rz, x.z, NZ), NY,
                    NZ);
  return cic_from(corner_abs(x.x, NX), corner_abs(x.y, NY), corner_abs(x.z, NZ), NY, NZ);
}

__device__ __forceinline__ float3 ld3(const float* a, unsigned p) {
  return make_float3(a[3 * (size_t)p], a[3 * (size_t)p + 1], a[3 * (size_t)p + 2]);
}
__device__ __forceinline__ void st3(float* a, unsigned p, float3 v) {
  a[3 * (size_t)p] = v.x;
  a[3 * (size_t)p + 1] = v.y;
  a[3 * (size_t)p + 2] = v.z;
}
__device__ __forceinline__ float3 axpy3(float3 x, float3 v, float c) {  // x + v*c, two roundings
  return make_float3(x.x + v.x * c, x.y + v.y * c, x.z + v.z * c);
}

// Contiguous particle chunk of this workgroup. Hardware dispatches workgroup b to XCD b % 8; the
// logical id remaps so each XCD walks one contiguous slab of the Lagrangian grid and the mesh
// planes it gathers from / scatters to stay in that XCD's L2.
struct Chunk {
  unsigned begin, end;
};
__device__ __forceinline__ unsigned xcd_logical(unsigned b, unsigned G) {
  return (G % 8u == 0u) ? (b % 8u) * (G / 8u) + b / 8u : b;
}
__device__ __forceinline__ Chunk my_chunk(unsigned n) {
  const unsigned G = gridDim.x;
  const unsigned per = (n + G - 1) / G;
  const unsigned beg = min(n, xcd_logical(blockIdx.x, G) * per);
  return Chunk{beg, min(n, beg + per)};
}

__global__ void k_fill(float* __restrict__ a, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}
__global__ void k_fill_d(double* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0;
}

// One-pass CIC scatter with global atomics (absolute mode, or relative mode on request).
__global__ void __launch_bounds__(256) k_paint(PartArgs a, const float* __restrict__ w, float w0,
                                               float* __restrict__ mesh, int NX, int NY, int NZ, int rx, int ry,
                                               int rz) {
  const unsigned np = (unsigned)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  LatticeWalk lw;
  lw.init(ch.begin + threadIdx.x, blockDim.x, a.npy, a.npz);
  for (unsigned p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x, lw.step()) {
    const float3 x = ld3(a.x, p);
    const float wt = w ? w[p] : w0;
    const Cic q = particle_cic(a, lw, x, NX, NY, NZ, rx, ry, rz);
#pragma unroll
    for (int o = 0; o < 8; ++o) unsafeAtomicAdd(mesh + q.off[o], q.w[o] * wt);
  }
}

// Owner-computes CIC paint for relative mode (particles on their Lagrangian lattice): one
// workgroup per mesh tile (BX, BY, BZ) accumulates in LDS every deposit that lands in its tile from
// the particles whose lattice lies within the tile grown by H + 1 cells, then writes the tile with
// plain stores -- every mesh cell is written exactly once, no global atomics and no zero fill.
// A particle with |floor(dx)| > H along any axis ("escape") is skipped by all tiles; its home tile
// records it in its own segment of the escape list (LDS counter, no global atomics) and
// k_paint_escapes deposits those with atomics afterwards. Halo particles whose two corners per axis
// miss the tile leave after one bounds test; particles whose 8 corners are all inside the tile
// deposit without per-corner checks.
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

// Fixed-point LDS accumulation: deposits are added as int64 multiples of 2^-32 (exact for CIC
// weights, order independent, so the painted mesh is bitwise deterministic) -- integer LDS atomics
// also run several times faster than float ones. Before issuing, each particle's upper-z corner
// values are handed to the next lane when that lane is its z-neighbour on the lattice with the
// same base (x, y) and base z + 1, halving the LDS atomics on smooth displacement fields.
constexpr float PT_FIX = 4294967296.0f;  // 2^32
__device__ __forceinline__ unsigned long long fix64(float v) { return (unsigned long long)__float2ll_rn(v * PT_FIX); }

template <bool R1>
__global__ void __launch_bounds__(512) k_paint_tiles(PartArgs a, const float* __restrict__ w, float w0,
                                                     float* __restrict__ mesh, int NX, int NY, int NZ, int rx_,
                                                     int ry_, int rz_, int BX, int BY, int BZ, int H, int accumulate,
                                                     unsigned* __restrict__ esc_count, unsigned* __restrict__ esc_list,
                                                     int esc_cap) {
  __shared__ unsigned long long tile[PT_MAX];
  const int rx = R1 ? 1 : rx_, ry = R1 ? 1 : ry_, rz = R1 ? 1 : rz_;
  const int nty = NY / BY, ntz = NZ / BZ;
  const unsigned L = xcd_logical(blockIdx.x, gridDim.x);
  const int tz = L % ntz, ty = (L / ntz) % nty, tx = L / (ntz * nty);
  const int T0x = tx * BX, T0y = ty * BY, T0z = tz * BZ;
  const int ncell = BX * BY * BZ;
  for (int i = threadIdx.x; i < ncell; i += blockDim.x) tile[i] = 0ull;
  if (threadIdx.x == 0) esc_count[L] = 0u;
  __syncthreads();
  // lattice i*r in [T0 - 1 - H, T0 + B - 1 + H] (unwrapped)
  const int ilo = -floordiv(-(T0x - 1 - H), rx), ihi = floordiv(T0x + BX - 1 + H, rx);
  const int jlo = -floordiv(-(T0y - 1 - H), ry), jhi = floordiv(T0y + BY - 1 + H, ry);
  const int klo = -floordiv(-(T0z - 1 - H), rz), khi = floordiv(T0z + BZ - 1 + H, rz);
  const int ci = ihi - ilo + 1, cj = jhi - jlo + 1, ck = khi - klo + 1;
  const float fH = (float)H;
  const int BYZ = BY * BZ;
  const int lane = __lane_id();
  // division-free walk over (ii, jj, kk), kk fastest, stride blockDim.x
  const int stride = blockDim.x;
  const int dk = stride % ck, drow = stride / ck;
  const int dj = drow % cj, di = drow / cj;
  int kk = threadIdx.x % ck;
  int jj = (threadIdx.x / ck) % cj;
  int ii = threadIdx.x / (ck * cj);
  unsigned* my_esc = esc_list + (size_t)L * esc_cap;
  // every lane of a wave runs the same number of iterations (shuffles below need full waves)
  const int rows_total = ci * cj * ck;
  const int wave0 = threadIdx.x - lane;
  const int iters = (rows_total - wave0 + PT_UNROLL * stride - 1) / (PT_UNROLL * stride);
  for (int it = 0; it < iters; ++it) {
    unsigned pidx[PT_UNROLL];
    int li[PT_UNROLL], lj[PT_UNROLL], lk[PT_UNROLL];
    bool ok[PT_UNROLL];
#pragma unroll
    for (int u = 0; u < PT_UNROLL; ++u) {
      ok[u] = ii < ci;
      li[u] = ilo + ii;
      lj[u] = jlo + jj;
      lk[u] = klo + kk;
      pidx[u] = ok[u] ? ((unsigned)wrap_near(li[u], a.npx) * a.npy + wrap_near(lj[u], a.npy)) * a.npz +
                            wrap_near(lk[u], a.npz)
                      : 0u;
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
      const float3 x = xv[u];
      const float fx = floorf(x.x), fy = floorf(x.y), fz = floorf(x.z);
      const bool esc = ok[u] && !(fx >= -fH && fx <= fH && fy >= -fH && fy <= fH && fz >= -fH && fz <= fH);
      const int Lx = li[u] * rx - T0x, Ly = lj[u] * ry - T0y, Lz = lk[u] * rz - T0z;
      if (esc && Lx >= 0 && Lx < BX && Ly >= 0 && Ly < BY && Lz >= 0 && Lz < BZ)
        my_esc[atomicAdd(esc_count + L, 1u)] = pidx[u];
      const bool use = ok[u] && !esc;
      const int bx = Lx + (use ? (int)fx : 0), by = Ly + (use ? (int)fy : 0), bz = Lz + (use ? (int)fz : 0);
      const bool touch = use && bx >= -1 && bx < BX && by >= -1 && by < BY && bz >= -1 && bz < BZ;
      const bool inside = touch && bx >= 0 && bx < BX - 1 && by >= 0 && by < BY - 1 && bz >= 0 && bz < BZ - 1;
      const float wt = (touch && w) ? w[pidx[u]] : w0;
      const float frx = x.x - fx, fry = x.y - fy, frz = x.z - fz;
      const float wx0 = 1.0f - fabsf(frx), wx1 = 1.0f - fabsf(frx - 1.0f);
      const float wy0 = 1.0f - fabsf(fry), wy1 = 1.0f - fabsf(fry - 1.0f);
      const float wz0 = 1.0f - fabsf(frz), wz1 = 1.0f - fabsf(frz - 1.0f);
      const float wxy[4] = {wx0 * wy0, wx1 * wy0, wx0 * wy1, wx1 * wy1};
      float lo[4], hi[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        lo[c] = (wxy[c] * wz0) * wt;
        hi[c] = (wxy[c] * wz1) * wt;
      }
      // z-neighbour merge among fast-path lanes (full wave participates in the shuffles)
      const int base = inside ? (bx * BY + by) * BZ + bz : -4;
      const int prev_base = __shfl_up(base, 1);
      float prev_hi[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) prev_hi[c] = __shfl_up(hi[c], 1);
      const bool took = inside && lane > 0 && prev_base >= 0 && prev_base + 1 == base;
      const int took_i = took ? 1 : 0;
      const bool next_took = (__shfl_down(took_i, 1) != 0) && lane < 63;
      if (inside) {
        unsigned long long* t = tile + base;
        const int offs[4] = {0, BYZ, BZ, BYZ + BZ};
#pragma unroll
        for (int c = 0; c < 4; ++c) atomicAdd(t + offs[c], fix64(took ? lo[c] + prev_hi[c] : lo[c]));
        if (!next_took) {
#pragma unroll
          for (int c = 0; c < 4; ++c) atomicAdd(t + offs[c] + 1, fix64(hi[c]));
        }
      } else if (touch) {
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          const int ix = bx + (o & 1), iy = by + ((o >> 1) & 1), iz = bz + ((o >> 2) & 1);
          if (ix >= 0 && ix < BX && iy >= 0 && iy < BY && iz >= 0 && iz < BZ)
            atomicAdd(&tile[(ix * BY + iy) * BZ + iz], fix64((o >> 2) ? hi[o & 3] : lo[o & 3]));
        }
      }
    }
  }
  __syncthreads();
  // rows of BZ contiguous cells
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int iz = c % BZ, iy = (c / BZ) % BY, ix = c / BYZ;
    float* dst = mesh + ((size_t)(T0x + ix) * NY + (T0y + iy)) * NZ + (T0z + iz);
    const float v = (float)((double)(long long)tile[c] * 2.3283064365386962890625e-10);
    *dst = accumulate ? *dst + v : v;
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
    const unsigned p = seg[e];
    const float3 x = ld3(a.x, p);
    const float wt = w ? w[p] : w0;
    LatticeWalk lw;
    lw.init(p, 1u, a.npy, a.npz);
    const Cic q = particle_cic(a, lw, x, NX, NY, NZ, rx, ry, rz);
#pragma unroll
    for (int o = 0; o < 8; ++o) unsafeAtomicAdd(mesh + q.off[o], q.w[o] * wt);
  }
}

// x <- x + v*c
__global__ void __launch_bounds__(256) k_drift(float* __restrict__ x, const float* __restrict__ v, unsigned n, float c) {
  const Chunk ch = my_chunk(n);
  for (unsigned i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) x[i] = x[i] + v[i] * c;
}

// CIC readout of an NC-component interleaved mesh. what: bit0 store F in out, bit1 kick
// v <- v + F*ck (NC == 3), bit2 drift x <- x + v*cd after the kick. Corners are summed
// sequentially in x-fastest order.
template <int NC>
__global__ void __launch_bounds__(256) k_readout(PartArgs a, const float* __restrict__ mesh, float* __restrict__ out,
                                                 float ck, float cd, int what, int NX, int NY, int NZ, int rx,
                                                 int ry, int rz) {
  const unsigned np = (unsigned)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  LatticeWalk lw;
  lw.init(ch.begin + threadIdx.x, blockDim.x, a.npy, a.npz);
  for (unsigned p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x, lw.step()) {
    float3 x = ld3(a.x, p);
    const Cic q = particle_cic(a, lw, x, NX, NY, NZ, rx, ry, rz);
    float f[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const float* m = mesh + (size_t)q.off[o] * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = f[c] + m[c] * q.w[o];
    }
    if (what & 1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) out[(size_t)p * NC + c] = f[c];
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

// particle position in grid units, lattice added in relative mode, wrapped into the periodic box
__device__ __forceinline__ double grid_pos(int mode, int lat, float xc, int n) {
  double g = (double)xc + (mode == 0 ? (double)lat : 0.0);
  return g - floor(g / n) * n;
}

__global__ void __launch_bounds__(256) k_paint_healpix(PartArgs a, const float* __restrict__ w, float w0, ShellArgs sa,
                                                       HpGeom g, float* __restrict__ acc, int NX, int NY, int NZ,
                                                       int rx, int ry, int rz) {
  const unsigned np = (unsigned)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  LatticeWalk lw;
  lw.init(ch.begin + threadIdx.x, blockDim.x, a.npy, a.npz);
  for (unsigned p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x, lw.step()) {
    const float3 x = ld3(a.x, p);
    const double X = grid_pos(a.mode, lw.i * rx, x.x, NX) * sa.box[0] / NX - sa.obs[0];
    const double Y = grid_pos(a.mode, lw.j * ry, x.y, NY) * sa.box[1] / NY - sa.obs[1];
    const double Z = grid_pos(a.mode, lw.k * rz, x.z, NZ) * sa.box[2] / NZ - sa.obs[2];
    const double r = sqrt(X * X + Y * Y + Z * Z);
    if (!(r > sa.rmin && r <= sa.rmax)) continue;
    const double wt = w ? (double)w[p] : (double)w0;
    const double phi0 = atan2(Y, X);
    double phi = phi0 < 0 ? phi0 + 2.0 * M_PI : phi0;
    if (!(phi < 2.0 * M_PI)) phi = 0.0;  // -tiny + 2 pi can round up to 2 pi
    if (sa.scheme == 0) {
      const long pix = hp_ang2pix(g, Z / r, phi);
      unsafeAtomicAdd(acc + pix, (float)wt);
    } else {
      const double theta = atan2(sqrt(X * X + Y * Y), Z);
      long pix[4];
      double wg[4];
      hp_interpol(g, theta, phi, pix, wg);
#pragma unroll
      for (int m = 0; m < 4; ++m) unsafeAtomicAdd(acc + pix[m], (float)(wt * wg[m]));
    }
  }
}

__global__ void k_finalize_map(const float* __restrict__ acc, float* __restrict__ map, long n, double scale,
                               int accumulate) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float v = (float)((double)acc[i] * scale);
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
                                                    float* __restrict__ acc, int rx, int ry, int rz) {
  const unsigned np = (unsigned)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  LatticeWalk lw;
  lw.init(ch.begin + threadIdx.x, blockDim.x, a.npy, a.npz);
  for (unsigned p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x, lw.step()) {
    float3 x = ld3(a.x, p);
    if (a.mode == 0) x = make_float3(x.x + (float)(lw.i * rx), x.y + (float)(lw.j * ry), x.z + (float)(lw.k * rz));
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
      unsafeAtomicAdd(acc + (size_t)(ox ? cx.i1 : cx.i0) * fa.fny + (oy ? cy.i1 : cy.i0), wv);
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
    hipLaunchKernelGGL(k_drift, dim3(grid_particles(3 * np)), dim3(256), 0, s, a.x, a.v, (unsigned)(3 * np), cd_pre);
    int rc = pmx_check_launch("drift");
    if (rc) return rc;
    a.v = nullptr;
  }
  if (a.mode != 0 || atomic) {
    if (!accumulate) {
      int rc = pmx_launch_fill(mesh, p->ncell, 0.f, s);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_paint, dim3(grid_particles(np)), dim3(256), 0, s, a, w, w0, mesh, p->nx, p->ny, p->nz, rx, ry,
                       rz);
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
  if (rx == 1 && ry == 1 && rz == 1)
    hipLaunchKernelGGL(k_paint_tiles<true>, dim3(ntiles), dim3(512), 0, s, a, w, w0, mesh, p->nx, p->ny, p->nz, rx, ry,
                       rz, BX, BY, BZ, PAINT_HALO, accumulate, p->esc_count, p->esc_list, cap);
  else
    hipLaunchKernelGGL(k_paint_tiles<false>, dim3(ntiles), dim3(512), 0, s, a, w, w0, mesh, p->nx, p->ny, p->nz, rx, ry,
                       rz, BX, BY, BZ, PAINT_HALO, accumulate, p->esc_count, p->esc_list, cap);
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
                       const double obs[3], double rmin, double rmax, int nside, int scheme, float* acc,
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

int pmx_launch_finalize_map(const float* acc, float* map, long n, double scale, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_map, dim3(grid_for(n, 256)), dim3(256), 0, s, acc, map, n, scale, accumulate);
  return pmx_check_launch("finalize_map");
}

int pmx_launch_flat(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, int fnx, int fny, float zlo, float zhi,
                    float* acc, hipStream_t s) {
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
