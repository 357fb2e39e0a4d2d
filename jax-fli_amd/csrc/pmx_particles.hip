// Particle-side kernels: CIC paint (owner-computes tiles + escape pass, or one-pass atomic scatter),
// CIC readout fused with the kick and the next drift, flat-sky slab painting, the adjoint gathers,
// counter-based white noise. HEALPix painting and Born lensing are in pmx_sky.hip.
//
// Every particle kernel walks the particle grid in its Lagrangian (i, j, k) order, so a wavefront
// touches a compact patch of the mesh and the 8-corner gathers/scatters stay in L2. Index math is
// 32-bit (meshes up to 1024^3 cells, x3 components, fit) and the lattice coordinates of a thread's
// particles are advanced incrementally instead of being divided out of the flat index.
// FP contraction is disabled here so x + v*c and the CIC weights round exactly like the
// unfused float32 expressions of the reference (two roundings, no FMA).
#include "pmx_internal.h"
#include "pmx_device.h"
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

__device__ __forceinline__ float3 axpy3(float3 x, float3 v, float c) {  // x + v*c, two roundings
  return make_float3(x.x + v.x * c, x.y + v.y * c, x.z + v.z * c);
}

// Contiguous particle chunk of this workgroup. Hardware dispatches workgroup b to XCD b % 8; the
// logical id remaps so each XCD walks one contiguous slab of the Lagrangian grid and the mesh
// planes it gathers from / scatters to stay in that XCD's L2.
struct Chunk {
  unsigned begin, end;
};
__device__ __forceinline__ Chunk my_chunk(unsigned n) {
  const unsigned G = gridDim.x;
  const unsigned per = (n + G - 1) / G;
  const unsigned beg = min(n, xcd_logical(blockIdx.x, G) * per);
  return Chunk{beg, min(n, beg + per)};
}

// One atomicMax per workgroup of the workgroup's largest m (non-negative floats order like their
// bits): same-address atomics from every wave of a large grid serialise.
__device__ __forceinline__ void block_max_to(float m, unsigned* bits) {
  __shared__ float wm[8];
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (__lane_id() == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t = fmaxf(t, wm[i]);
    if (t > 0.f) atomicMax(bits, __float_as_uint(t));
  }
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

// ------------------------------------------------------------------ sorted CIC paint
// Counting sort of the particles by the mesh tile that holds their base cell (floor of the
// position), then one workgroup per tile accumulates the 8 corners of its particles in an LDS copy
// of the tile grown by one cell on each upper face, writes the tile's own cells with plain stores
// and its upper face cells to a face buffer; a last pass adds every tile's face cells to the cells
// of its upper neighbours. Every mesh cell is written by its owner, in a fixed order: no global
// float atomics, no zero fill, bitwise deterministic.
//   k_sort_count    [x += v c]; per-tile particle counts (LDS window histogram -> global counts)
//   k_sort_scan     exclusive prefix of the counts -> segment offsets and reservation cursors
//   k_sort_scatter  particle indices -> their tile's segment (per-workgroup bulk reservation)
//   k_paint_deposit per tile: LDS int64 fixed-point accumulation, tile cells + face cells out
//   k_paint_faces   tile cells on lower faces += the neighbours' face cells
// Deposits are added as int64 multiples of 2^-k, k chosen so the largest possible deposit is
// 2^30 (exact, order-independent sums); each deposit is truncated to that grid.
//
// Particles are read in Lagrangian sub-blocks of 8 x 8 x 64 lattice sites; a sub-block's tiles lie
// within a window of 5 x 5 x 8 tiles around the tile of its first site (-2..+2 in x and y, -2..+5
// in z), so they are histogrammed in an LDS window and reserved with one global atomic per window
// slot. Particles outside the window (displaced by more than about two tiles) count and reserve one
// by one with global atomics.
constexpr int SB_X = 8, SB_Y = 8, SB_Z = 64;  // particle sub-block (4096 sites, 256 threads x 16)
// window of tiles around the tile of a sub-block's first site: SW_* tiles per axis, SL_* of them
// below it. z is longer and reaches further up because a sub-block is 64 sites long in z (two tiles
// at one particle per cell, four at the 2-cell spacing of a particle sub-lattice).
constexpr int SW_X = 5, SW_Y = 5, SW_Z = 8, SL_X = 2, SL_Y = 2, SL_Z = 2;
constexpr int SORT_SLOTS = SW_X * SW_Y * SW_Z;
constexpr int TILE_PAD_MAX = 17 * 17 * 33;    // padded tile cells (int64 LDS: 76 KiB)

// Division by a runtime constant: q = floor(n / d) for all 32-bit n (round-up multiplier).
struct FDiv {
  unsigned m, d;
  int s1, s2;
};
__device__ __forceinline__ unsigned fdiv(unsigned n, const FDiv& f) {
  const unsigned t = __umulhi(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

struct SortGeom {
  int NX, NY, NZ;        // mesh
  int sx, sy, sz;        // log2 tile dims
  int ntx, nty, ntz;     // tiles per axis
  int rx, ry, rz;        // mesh / particle grid (relative mode)
  int nbx, nby, nbz;     // particle sub-blocks per axis
  int lpz, lpy;          // log2 of npz, npy when both are powers of two, else -1
  FDiv div_npz, div_npy;
  int wmask;             // NX - 1 when NX, NY, NZ are all powers of two (wrap by masking), else -1
};

// periodic wrap, one compare per side for the usual |b| < N, exact modulo otherwise
__device__ __forceinline__ int wrap_cell(int b, int n) {
  int c = b + (b < 0 ? n : 0);
  c -= (c >= n ? n : 0);
  if ((unsigned)c >= (unsigned)n) {
    c = b % n;
    c += c < 0 ? n : 0;
  }
  return c;
}

// (q / e, q % e) for small box and face indices, without an integer division:
// q = a e + r, 0 <= r < e, puts (q + 1/2) / e at least 1/(2e) away from an integer, and the float
// product (q + 1/2) * fl(1/e) is within about 2^-23 (q + 1) / e of it, inside that margin for any
// q < 2^20 (the boxes hold at most a few thousand cells).
__device__ __forceinline__ int div_small(int q, float inv_e, int e, int& rem) {
  const int a = (int)(((float)q + 0.5f) * inv_e);
  rem = q - a * e;
  return a;
}

// lattice coordinates of particle p
__device__ __forceinline__ void lattice_of(const PartArgs& a, const SortGeom& g, unsigned p, int& i, int& j, int& k) {
  if (g.lpz >= 0) {
    k = (int)(p & (unsigned)(a.npz - 1));
    const unsigned q = p >> g.lpz;
    j = (int)(q & (unsigned)(a.npy - 1));
    i = (int)(q >> g.lpy);
  } else {
    const unsigned q = fdiv(p, g.div_npz);
    k = (int)(p - q * (unsigned)a.npz);
    const unsigned ii = fdiv(q, g.div_npy);
    j = (int)(q - ii * (unsigned)a.npy);
    i = (int)ii;
  }
}

// Wrapped base cell of a particle at lattice (i, j, k) with displacement / position x, and the
// fractional offsets. Relative mode: lattice * r + floor(dx), weights from dx - floor(dx) (the
// displacement form of the reference's cic_paint_dx); absolute mode: floor(x).
struct Base {
  int cx, cy, cz;
  float fx, fy, fz;
};
__device__ __forceinline__ Base base_of(const PartArgs& a, const SortGeom& g, int i, int j, int k, float3 x) {
  const float flx = floorf(x.x), fly = floorf(x.y), flz = floorf(x.z);
  int bx = (int)flx, by = (int)fly, bz = (int)flz;
  if (a.mode == 0) {
    bx += __mul24(i, g.rx);
    by += __mul24(j, g.ry);
    bz += __mul24(k, g.rz);
  }
  Base b;
  if (g.wmask >= 0) {  // two's complement: b & (N - 1) is b mod N for any b
    b.cx = bx & (g.NX - 1);
    b.cy = by & (g.NY - 1);
    b.cz = bz & (g.NZ - 1);
  } else {
    b.cx = wrap_cell(bx, g.NX);
    b.cy = wrap_cell(by, g.NY);
    b.cz = wrap_cell(bz, g.NZ);
  }
  b.fx = x.x - flx;
  b.fy = x.y - fly;
  b.fz = x.z - flz;
  return b;
}

// wrapped tile difference d in [-(n/2), n - n/2): a unique representative for every tile
__device__ __forceinline__ int tdiff(int t, int c, int n) {
  int d = t - c;
  const int lo = -(n >> 1);
  if (d < lo) d += n;
  else if (d >= n + lo) d -= n;
  return d;
}

// Sub-block b (logical) -> its origin on the particle grid.
__device__ __forceinline__ void subblock_origin(const SortGeom& g, unsigned b, int& i0, int& j0, int& k0) {
  const int bk = (int)(b % (unsigned)g.nbz);
  const unsigned q = b / (unsigned)g.nbz;
  const int bj = (int)(q % (unsigned)g.nby);
  const int bi = (int)(q / (unsigned)g.nby);
  i0 = bi * SB_X;
  j0 = bj * SB_Y;
  k0 = bk * SB_Z;
}

// Window centre: tile of the sub-block's first site (relative mode: its lattice cell; absolute
// mode: that particle's base cell). Identical in k_sort_count and k_sort_scatter.
__device__ __forceinline__ int3 window_centre(const PartArgs& a, const SortGeom& g, int i0, int j0, int k0) {
  int bx, by, bz;
  if (a.mode == 0) {
    bx = i0 * g.rx;
    by = j0 * g.ry;
    bz = k0 * g.rz;
  } else {
    const unsigned p = ((unsigned)i0 * a.npy + j0) * a.npz + k0;
    const Base b = base_of(a, g, 0, 0, 0, ld3(a.x, p));
    bx = b.cx;
    by = b.cy;
    bz = b.cz;
  }
  return make_int3(bx >> g.sx, by >> g.sy, bz >> g.sz);
}

__device__ __forceinline__ unsigned slot_tile(const SortGeom& g, int3 c, int s) {
  const int dz = s % SW_Z - SL_Z, dy = (s / SW_Z) % SW_Y - SL_Y, dx = s / (SW_Y * SW_Z) - SL_X;
  const int tx = wrap_idx(c.x + dx, g.ntx), ty = wrap_idx(c.y + dy, g.nty), tz = wrap_idx(c.z + dz, g.ntz);
  return ((unsigned)tx * g.nty + ty) * g.ntz + tz;
}

// Visit the particles of this thread in sub-block (i0, j0, k0) (4 loads in flight):
// fn(valid, p, tile, slot) with slot = window slot of the tile or -1; called by every lane, so fn
// may use wave-wide operations.
template <class Fn>
__device__ __forceinline__ void for_particles(const PartArgs& a, const SortGeom& g, int i0, int j0, int k0, int3 c,
                                              Fn&& fn) {
  constexpr int PER = (SB_X * SB_Y * SB_Z) / 256, U = 4;
  for (int r0 = 0; r0 < PER; r0 += U) {
    unsigned ps[U];
    bool vs[U];
    int is[U], js[U], ks[U];
    float3 xs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int l = (r0 + u) * 256 + (int)threadIdx.x;
      is[u] = i0 + (l >> 9);
      js[u] = j0 + ((l >> 6) & 7);
      ks[u] = k0 + (l & 63);
      vs[u] = is[u] < a.npx && js[u] < a.npy && ks[u] < a.npz;
      ps[u] = vs[u] ? ((unsigned)is[u] * a.npy + js[u]) * a.npz + ks[u] : 0u;
      xs[u] = vs[u] ? ld3(a.x, ps[u]) : make_float3(0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const Base b = base_of(a, g, is[u], js[u], ks[u], xs[u]);
      const int tx = b.cx >> g.sx, ty = b.cy >> g.sy, tz = b.cz >> g.sz;
      const unsigned tile = ((unsigned)tx * g.nty + ty) * g.ntz + tz;
      const int dx = tdiff(tx, c.x, g.ntx), dy = tdiff(ty, c.y, g.nty), dz = tdiff(tz, c.z, g.ntz);
      const bool inw = (unsigned)(dx + SL_X) < (unsigned)SW_X && (unsigned)(dy + SL_Y) < (unsigned)SW_Y &&
                       (unsigned)(dz + SL_Z) < (unsigned)SW_Z;
      fn(vs[u], ps[u], tile, inw ? ((dx + SL_X) * SW_Y + (dy + SL_Y)) * SW_Z + (dz + SL_Z) : -1);
    }
  }
}

constexpr unsigned SORT_FAR = 0xFFFFFFFFu;  // pack entry of a particle outside its sub-block's window

// Pass 1 for relative mode: the lattice cell plus floor(displacement) gives an unwrapped base cell
// whose tile, unwrapped, differs from the window centre (the sub-block's first lattice tile) by a
// plain subtraction; no periodic wrap or canonical difference is needed, because a slot names the
// tile wrap(centre + d) and two slots naming one tile simply reserve two ranges of it. Each lane's
// lattice coordinates follow from its index and the (unrolled) round: i = i0 + r / 2,
// j = j0 + 4 (r % 2) + wave, k = k0 + lane. Particles outside the window take the far path with
// the wrapped tile, as in for_particles.
__device__ __forceinline__ void count_relative(const PartArgs& a, const SortGeom& g, int i0, int j0, int k0, int3 c,
                                               unsigned* lcnt, unsigned* __restrict__ count,
                                               unsigned* __restrict__ pack, unsigned& nfar) {
  constexpr int PER = (SB_X * SB_Y * SB_Z) / 256, U = 4;
  const int wv = (int)threadIdx.x >> 6;
  const int k = k0 + ((int)threadIdx.x & 63);
  const bool kin = k < a.npz;
  const int bz0 = __mul24(k, g.rz);
#pragma unroll
  for (int r0 = 0; r0 < PER; r0 += U) {
    unsigned ps[U];
    bool vs[U];
    int is[U], js[U];
    float3 xs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + u;
      is[u] = i0 + (r >> 1);
      js[u] = j0 + (r & 1) * 4 + wv;
      vs[u] = kin && is[u] < a.npx && js[u] < a.npy;
      ps[u] = vs[u] ? ((unsigned)is[u] * a.npy + js[u]) * a.npz + k : 0u;
      xs[u] = vs[u] ? ld3(a.x, ps[u]) : make_float3(0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int bx = __mul24(is[u], g.rx) + (int)floorf(xs[u].x);
      const int by = __mul24(js[u], g.ry) + (int)floorf(xs[u].y);
      const int bz = bz0 + (int)floorf(xs[u].z);
      const int dx = (bx >> g.sx) - c.x + SL_X, dy = (by >> g.sy) - c.y + SL_Y, dz = (bz >> g.sz) - c.z + SL_Z;
      const bool inw = (unsigned)dx < (unsigned)SW_X && (unsigned)dy < (unsigned)SW_Y && (unsigned)dz < (unsigned)SW_Z;
      const int slot = (dx * SW_Y + dy) * SW_Z + dz;
      const bool on = vs[u];
      // rank within the slot's segment: one LDS atomic per lane (lanes of a wave that share a slot
      // are serialised by the LDS unit, which costs less than aggregating them per wave)
      const unsigned rk = (on && inw) ? atomicAdd(&lcnt[slot], 1u) : 0u;
      if (on) pack[ps[u]] = inw ? ((unsigned)slot << 16) | rk : SORT_FAR;
      if (on && !inw) {
        const Base bb = base_of(a, g, is[u], js[u], k, xs[u]);
        atomicAdd(count + (((unsigned)(bb.cx >> g.sx) * g.nty + (bb.cy >> g.sy)) * g.ntz + (bb.cz >> g.sz)), 1u);
        ++nfar;
      }
    }
  }
}

// Pass 1: optional pre-drift x += v * cd (written back), then per-tile particle counts. The window
// histogram of each sub-block is kept in wgcnt and each particle's (window slot, rank within the
// slot) in pack for the scatter pass. far: particles outside windows. With weights (wstride
// values per particle) the largest |weight| goes to wmax_bits.

__global__ void __launch_bounds__(256) k_sort_count(PartArgs a, SortGeom g, float cd, unsigned* __restrict__ count,
                                                    unsigned* __restrict__ wgcnt, unsigned* __restrict__ far,
                                                    const float* __restrict__ w, int wstride,
                                                    unsigned* __restrict__ wmax_bits, unsigned* __restrict__ pack,
                                                    Slots sl) {
  __shared__ unsigned lcnt[SORT_SLOTS];
  __shared__ int3 cen;
  const unsigned z = blockIdx.z;
  if (!slot_live(sl, z)) return;
  a.x = at_slot(a.x, sl.stride, z);
  a.v = at_slot(a.v, sl.stride, z);
  count = at_slot(count, sl.stride, z);
  wgcnt = at_slot(wgcnt, sl.stride, z);
  far = at_slot(far, sl.stride, z);
  w = at_slot(w, sl.stride, z);
  wmax_bits = at_slot(wmax_bits, sl.stride, z);
  pack = at_slot(pack, sl.stride, z);
  const unsigned b = xcd_logical(blockIdx.x, gridDim.x);
  int i0, j0, k0;
  subblock_origin(g, b, i0, j0, k0);
  for (int s = threadIdx.x; s < SORT_SLOTS; s += blockDim.x) lcnt[s] = 0u;
  if (a.v) {  // drift this sub-block (each thread its own particles) before anything reads them
    for (int r = 0; r < (SB_X * SB_Y * SB_Z) / 256; ++r) {
      const int l = r * 256 + (int)threadIdx.x;
      const int i = i0 + (l >> 9), j = j0 + ((l >> 6) & 7), k = k0 + (l & 63);
      if (i >= a.npx || j >= a.npy || k >= a.npz) continue;
      const unsigned p = ((unsigned)i * a.npy + j) * a.npz + k;
      st3(a.x, p, axpy3(ld3(a.x, p), ld3(a.v, p), cd));
    }
  }
  if (threadIdx.x == 0) cen = window_centre(a, g, i0, j0, k0);
  __syncthreads();
  const int3 c = cen;
  unsigned nfar = 0;
  if (a.mode == 0) {
    count_relative(a, g, i0, j0, k0, c, lcnt, count, pack, nfar);
  } else {
    for_particles(a, g, i0, j0, k0, c, [&](bool on, unsigned p, unsigned tile, int slot) {
      const unsigned r = (on && slot >= 0) ? atomicAdd(&lcnt[slot], 1u) : 0u;  // rank within the slot's segment
      if (on) pack[p] = slot >= 0 ? ((unsigned)slot << 16) | r : SORT_FAR;
      if (on && slot < 0) {
        atomicAdd(count + tile, 1u);
        ++nfar;
      }
    });
  }
  if (w) {  // largest |weight| for the fixed-point scale (non-negative floats order like their bits)
    float m = 0.f;
    for (int r = 0; r < (SB_X * SB_Y * SB_Z) / 256; ++r) {
      const int l = r * 256 + (int)threadIdx.x;
      const int i = i0 + (l >> 9), j = j0 + ((l >> 6) & 7), k = k0 + (l & 63);
      if (i >= a.npx || j >= a.npy || k >= a.npz) continue;
      const size_t q = (size_t)(((unsigned)i * a.npy + j) * a.npz + k) * wstride;
      for (int c = 0; c < wstride; ++c) m = fmaxf(m, fabsf(w[q + c]));
    }
    block_max_to(m, wmax_bits);
  }
  if (nfar) atomicAdd(far, nfar);
  __syncthreads();
  for (int s = threadIdx.x; s < SORT_SLOTS; s += blockDim.x) {
    const unsigned n = lcnt[s];
    wgcnt[(size_t)b * SORT_SLOTS + s] = n;
    if (n) atomicAdd(count + slot_tile(g, c, s), n);
  }
}

// Pass 2 (one workgroup): offsets = exclusive prefix of count, cursor = offsets, tcount = count
// (count re-zeroed); the fixed-point scale 2^k (largest deposit 2^30) from the weight bound; far
// counter moved to far_last.
__global__ void __launch_bounds__(1024) k_sort_scan(unsigned* __restrict__ count, unsigned* __restrict__ tcount,
                                                    unsigned* __restrict__ offset,
                                                    unsigned* __restrict__ cursor, int ntiles, float wmax_host,
                                                    unsigned* __restrict__ wmax_bits, float* __restrict__ scale,
                                                    unsigned* __restrict__ far, unsigned* __restrict__ far_last,
                                                    Slots sl) {
  __shared__ unsigned part[1024];
  const unsigned z = blockIdx.z;
  if (!slot_live(sl, z)) return;
  count = at_slot(count, sl.stride, z);
  tcount = at_slot(tcount, sl.stride, z);
  offset = at_slot(offset, sl.stride, z);
  cursor = at_slot(cursor, sl.stride, z);
  wmax_bits = at_slot(wmax_bits, sl.stride, z);
  scale = at_slot(scale, sl.stride, z);
  far = at_slot(far, sl.stride, z);
  far_last = at_slot(far_last, sl.stride, z);
  const int T = blockDim.x, t = threadIdx.x;
  const int per = (ntiles + T - 1) / T;
  const int beg = min(ntiles, t * per), end = min(ntiles, beg + per);
  unsigned sum = 0;
  for (int i = beg; i < end; ++i) sum += count[i];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < T; o <<= 1) {  // Hillis-Steele inclusive scan of the per-thread sums
    const unsigned v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  unsigned run = part[t] - sum;
  for (int i = beg; i < end; ++i) {
    const unsigned c = count[i];
    offset[i] = run;
    cursor[i] = run;
    tcount[i] = c;
    count[i] = 0u;  // the next paint's count pass starts from zero
    run += c;
  }
  if (t == 0) {
    const float m = wmax_host >= 0.f ? wmax_host : __uint_as_float(*wmax_bits);
    int e = 0;
    if (m > 0.f) (void)frexpf(m, &e);  // m < 2^e
    e = max(-90, min(e, 90));
    scale[0] = ldexpf(1.0f, 30 - e);
    scale[1] = ldexpf(1.0f, e - 30);
    *wmax_bits = 0u;
    *far_last = *far;
    *far = 0u;
  }
}

// Pass 3: each particle's index into its tile's segment: the segment base the sub-block reserved for
// the particle's window slot plus the rank the count pass gave it (no re-binning); particles
// outside the window take one global cursor increment on their tile.
__global__ void __launch_bounds__(256) k_sort_scatter(PartArgs a, SortGeom g, unsigned* __restrict__ cursor,
                                                      const unsigned* __restrict__ wgcnt,
                                                      const unsigned* __restrict__ pack, unsigned* __restrict__ seg,
                                                      Slots sl) {
  __shared__ unsigned lbase[SORT_SLOTS];
  __shared__ int3 cen;
  const unsigned z = blockIdx.z;
  if (!slot_live(sl, z)) return;
  a.x = at_slot(a.x, sl.stride, z);
  cursor = at_slot(cursor, sl.stride, z);
  wgcnt = at_slot(wgcnt, sl.stride, z);
  pack = at_slot(pack, sl.stride, z);
  seg = at_slot(seg, sl.stride, z);
  const unsigned b = xcd_logical(blockIdx.x, gridDim.x);
  int i0, j0, k0;
  subblock_origin(g, b, i0, j0, k0);
  if (threadIdx.x == 0) cen = window_centre(a, g, i0, j0, k0);
  __syncthreads();
  const int3 c = cen;
  for (int s = threadIdx.x; s < SORT_SLOTS; s += blockDim.x) {
    const unsigned n = wgcnt[(size_t)b * SORT_SLOTS + s];
    lbase[s] = n ? atomicAdd(cursor + slot_tile(g, c, s), n) : 0u;
  }
  __syncthreads();
  for (int r = 0; r < (SB_X * SB_Y * SB_Z) / 256; ++r) {
    const int l = r * 256 + (int)threadIdx.x;
    const int i = i0 + (l >> 9), j = j0 + ((l >> 6) & 7), k = k0 + (l & 63);
    if (i >= a.npx || j >= a.npy || k >= a.npz) continue;
    const unsigned p = ((unsigned)i * a.npy + j) * a.npz + k;
    const unsigned q = pack[p];
    if (q != SORT_FAR) {
      seg[lbase[q >> 16] + (q & 0xFFFFu)] = p;
    } else {
      const Base bb = base_of(a, g, i, j, k, ld3(a.x, p));
      const unsigned tile = ((unsigned)(bb.cx >> g.sx) * g.nty + (bb.cy >> g.sy)) * g.ntz + (bb.cz >> g.sz);
      seg[atomicAdd(cursor + tile, 1u)] = p;
    }
  }
}

// Face cells of a tile grown by one cell on its upper faces (padded local coordinates with at
// least one coordinate equal to the tile length), enumerated as: x face (ly, lz in [0, B]),
// y face (lx in [0, BX), lz in [0, BZ]), z face (lx in [0, BX), ly in [0, BY)).
__device__ __forceinline__ int face_index(int lx, int ly, int lz, int BX, int BY, int BZ) {
  if (lx == BX) return ly * (BZ + 1) + lz;
  if (ly == BY) return (BY + 1) * (BZ + 1) + lx * (BZ + 1) + lz;
  return (BY + 1) * (BZ + 1) + BX * (BZ + 1) + lx * BY + ly;
}

// Pass 4: one workgroup per tile: deposit the 8 corners of the tile's particles into the padded LDS
// tile, write the tile's cells to the mesh (accumulate: add) and its face cells to faces. Particle
// p carries w[p * wstride] * wmul (w != null) or w0. Reads only the sort's output, so several
// deposits (e.g. the components of a vector) can follow one sort.
template <bool SIGNED>
__global__ void __launch_bounds__(512) k_paint_deposit(PartArgs a, SortGeom g, const float* __restrict__ w,
                                                       int wstride, float wmul, float w0, int ncomp,
                                                       const unsigned* __restrict__ tcount,
                                                       const unsigned* __restrict__ offset,
                                                       const unsigned* __restrict__ seg,
                                                       const float* __restrict__ scale, float* __restrict__ mesh,
                                                       int mstride, size_t mcomp, float* __restrict__ faces,
                                                       int face_stride, size_t fcomp, int accumulate, Slots sl) {
  __shared__ __align__(16) unsigned long long tile[TILE_PAD_MAX];
  const unsigned zs = blockIdx.z;
  if (!slot_live(sl, zs)) return;
  a.x = at_slot(a.x, sl.stride, zs);
  w = at_slot(w, sl.stride, zs);
  tcount = at_slot(tcount, sl.stride, zs);
  offset = at_slot(offset, sl.stride, zs);
  seg = at_slot(seg, sl.stride, zs);
  scale = at_slot(scale, sl.stride, zs);
  mesh = at_slot(mesh, sl.stride, zs);
  faces = at_slot(faces, sl.stride, zs);
  const int BX = 1 << g.sx, BY = 1 << g.sy, BZ = 1 << g.sz;
  const int PY = BY + 1, PZ = BZ + 1;
  const int npad = (BX + 1) * PY * PZ;
  const unsigned L = xcd_logical(blockIdx.x, gridDim.x);
  const int tz = (int)(L % (unsigned)g.ntz), ty = (int)((L / (unsigned)g.ntz) % (unsigned)g.nty),
            tx = (int)(L / ((unsigned)g.ntz * g.nty));
  const int T0x = tx << g.sx, T0y = ty << g.sy, T0z = tz << g.sz;
  const unsigned n = tcount[L], beg = offset[L];
  const float S = scale[0];
  const float wS = w0 * S, mS = wmul * S;
  const double inv = (double)scale[1];
  // byte offsets of the +1 neighbours in the padded tile
  const int oz = 8, oy = PZ * 8, ox = PY * PZ * 8;
  char* lds = reinterpret_cast<char*>(tile);
  const int ncell = BX * BY * BZ;
  const int nface = PY * PZ + BX * PZ + BX * BY;
  // one component per workgroup (blockIdx.y) when the grid has a y extent, else one after the
  // other through the same LDS tile (the entries stay in L1/L2)
  const int c_lo = gridDim.y > 1 ? (int)blockIdx.y : 0, c_hi = gridDim.y > 1 ? c_lo + 1 : ncomp;
  for (int comp = c_lo; comp < c_hi; ++comp) {
    {  // zero the padded tile with 16-byte stores
      ulonglong2* t2 = reinterpret_cast<ulonglong2*>(tile);
      for (int i = threadIdx.x; i < npad / 2; i += blockDim.x) t2[i] = make_ulonglong2(0ull, 0ull);
      if ((npad & 1) && threadIdx.x == 0) tile[npad - 1] = 0ull;
    }
    __syncthreads();
    constexpr int U = 4;
    unsigned pn[U];  // the particle indices of the thread's next group, loaded one group ahead
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned e = threadIdx.x + u * blockDim.x;
      pn[u] = e < n ? seg[beg + e] : 0u;
    }
    for (unsigned e0 = threadIdx.x; e0 < n; e0 += U * blockDim.x) {
      unsigned ps[U];
      float3 xs[U];
#pragma unroll
      for (int u = 0; u < U; ++u) ps[u] = pn[u];
#pragma unroll
      for (int u = 0; u < U; ++u) xs[u] = (e0 + u * blockDim.x < n) ? ld3(a.x, ps[u]) : make_float3(0.f, 0.f, 0.f);
      // the particle weights in flight with the positions (loaded inside the corner loop they cost one
      // memory round trip per particle: the 3-component adjoint deposit 49 -> 45 us at 128^3)
      float wv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) wv[u] = (w && e0 + u * blockDim.x < n) ? w[(size_t)ps[u] * wstride + comp] : 0.f;
      // the next group's particle indices in flight with this group's positions (99.9 -> 97.0 us at 256^3)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const unsigned e = e0 + (U + u) * blockDim.x;
        pn[u] = e < n ? seg[beg + e] : 0u;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e0 + u * blockDim.x >= n) continue;
        const unsigned p = ps[u];
        int i = 0, j = 0, k = 0;
        if (a.mode == 0) lattice_of(a, g, p, i, j, k);
        const Base b = base_of(a, g, i, j, k, xs[u]);
        const int base = (((b.cx - T0x) * PY + (b.cy - T0y)) * PZ + (b.cz - T0z)) * 8;
        const float c = w ? wv[u] * mS : wS;
        const float wx[2] = {1.0f - fabsf(b.fx), 1.0f - fabsf(b.fx - 1.0f)};
        const float wy[2] = {1.0f - fabsf(b.fy), 1.0f - fabsf(b.fy - 1.0f)};
        // c is folded into the z weights: exact when c is a power of two (scalar weight 1)
        const float wz[2] = {(1.0f - fabsf(b.fz)) * c, (1.0f - fabsf(b.fz - 1.0f)) * c};
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          const int cx = o & 1, cy = (o >> 1) & 1, cz = o >> 2;
          const float v = (wx[cx] * wy[cy]) * wz[cz];
          const unsigned long long d =
              SIGNED ? (unsigned long long)(long long)__float2int_rz(v) : (unsigned long long)__float2uint_rz(v);
          atomicAdd(reinterpret_cast<unsigned long long*>(lds + base + cx * ox + cy * oy + cz * oz), d);
        }
      }
    }
    __syncthreads();
    float* mc = mesh + comp * mcomp;
    if (mstride == 1 && BZ >= 4 && (g.NZ & 3) == 0 && (reinterpret_cast<uintptr_t>(mc) & 15u) == 0) {
      // four consecutive z cells per thread: one 16-byte store (105 -> 98 us at 256^3 with the zeroing)
      for (int q = threadIdx.x; 4 * q < ncell; q += blockDim.x) {
        const int q4 = 4 * q;
        const int iz = q4 & (BZ - 1), iy = (q4 >> g.sz) & (BY - 1), ix = q4 >> (g.sz + g.sy);
        float4* dst = reinterpret_cast<float4*>(mc + ((size_t)(T0x + ix) * g.NY + (T0y + iy)) * g.NZ + (T0z + iz));
        const unsigned long long* t = tile + (ix * PY + iy) * PZ + iz;
        float4 v = make_float4((float)((double)(long long)t[0] * inv), (float)((double)(long long)t[1] * inv),
                               (float)((double)(long long)t[2] * inv), (float)((double)(long long)t[3] * inv));
        if (accumulate) {
          const float4 o = *dst;
          v = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
        }
        *dst = v;
      }
    } else {
      for (int q = threadIdx.x; q < ncell; q += blockDim.x) {
        const int iz = q & (BZ - 1), iy = (q >> g.sz) & (BY - 1), ix = q >> (g.sz + g.sy);
        float* dst = mc + (((size_t)(T0x + ix) * g.NY + (T0y + iy)) * g.NZ + (T0z + iz)) * mstride;
        const float v = (float)((double)(long long)tile[(ix * PY + iy) * PZ + iz] * inv);
        *dst = accumulate ? *dst + v : v;
      }
    }
    float* f = faces + comp * fcomp + (size_t)L * face_stride;
    const float ipz = __builtin_amdgcn_rcpf((float)PZ);
    for (int q = threadIdx.x; q < nface; q += blockDim.x) {
      int lx, ly, lz;
      if (q < PY * PZ) {
        lx = BX;
        ly = div_small(q, ipz, PZ, lz);
      } else if (q < PY * PZ + BX * PZ) {
        lx = div_small(q - PY * PZ, ipz, PZ, lz);
        ly = BY;
      } else {
        const int r = q - PY * PZ - BX * PZ;
        lx = r >> g.sy;
        ly = r & (BY - 1);
        lz = BZ;
      }
      f[q] = (float)((double)(long long)tile[(lx * PY + ly) * PZ + lz] * inv);
    }
    __syncthreads();
  }
}

// Pass 5: cells on a tile's lower faces receive the face cells of its lower neighbours (x, y, z,
// xy, xz, yz, xyz, in that order). One thread per (tile, lower-shell cell).
__global__ void __launch_bounds__(256) k_paint_faces(SortGeom g, const float* __restrict__ faces, int face_stride,
                                                     size_t fcomp, float* __restrict__ mesh, int mstride, size_t mcomp,
                                                     int ncomp, int shell, Slots sl) {
  const unsigned zs = blockIdx.z;
  if (!slot_live(sl, zs)) return;
  faces = at_slot(faces, sl.stride, zs);
  mesh = at_slot(mesh, sl.stride, zs);
  const int BX = 1 << g.sx, BY = 1 << g.sy, BZ = 1 << g.sz;
  const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t ntiles = (size_t)g.ntx * g.nty * g.ntz;
  if (gid >= ntiles * shell) return;
  const unsigned L = (unsigned)(gid / shell);
  const int q = (int)(gid % shell);
  // lower shell cell q: lx == 0 plane (BY x BZ), then ly == 0 (lx >= 1), then lz == 0 (lx, ly >= 1)
  int lx, ly, lz;
  if (q < BY * BZ) {
    lx = 0;
    ly = q / BZ;
    lz = q % BZ;
  } else if (q < BY * BZ + (BX - 1) * BZ) {
    const int r = q - BY * BZ;
    lx = 1 + r / BZ;
    ly = 0;
    lz = r % BZ;
  } else {
    const int r = q - BY * BZ - (BX - 1) * BZ;
    lx = 1 + r / (BY - 1);
    ly = 1 + r % (BY - 1);
    lz = 0;
  }
  const int tz = (int)(L % (unsigned)g.ntz), ty = (int)((L / (unsigned)g.ntz) % (unsigned)g.nty),
            tx = (int)(L / ((unsigned)g.ntz * g.nty));
  size_t fi[7];
  int nsrc = 0;
#pragma unroll
  for (int m = 1; m < 8; ++m) {
    const int ex = m & 1, ey = (m >> 1) & 1, ez = m >> 2;
    if ((ex && lx != 0) || (ey && ly != 0) || (ez && lz != 0)) continue;
    const int nx = ex ? (tx == 0 ? g.ntx - 1 : tx - 1) : tx;
    const int ny = ey ? (ty == 0 ? g.nty - 1 : ty - 1) : ty;
    const int nz = ez ? (tz == 0 ? g.ntz - 1 : tz - 1) : tz;
    const unsigned nb = ((unsigned)nx * g.nty + ny) * g.ntz + nz;
    fi[nsrc++] = (size_t)nb * face_stride + face_index(ex ? BX : lx, ey ? BY : ly, ez ? BZ : lz, BX, BY, BZ);
  }
  const size_t cell = ((size_t)((tx << g.sx) + lx) * g.NY + ((ty << g.sy) + ly)) * g.NZ + ((tz << g.sz) + lz);
  for (int c = 0; c < ncomp; ++c) {
    float sum = 0.f;
    for (int m = 0; m < nsrc; ++m) sum += faces[c * fcomp + fi[m]];
    float* dst = mesh + c * mcomp + cell * mstride;
    *dst = *dst + sum;
  }
}

// ------------------------------------------------------------------ adjoint (VJP) kernels
// Sign of the reference's CIC weight derivatives: d(1 - |f|)/df = -sign(f), d(1 - |f - 1|)/df =
// -sign(f - 1), with sign(0) = 0.
__device__ __forceinline__ float sgnf(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }

__device__ __forceinline__ float absmax4(float m, float4 v) {
  return fmaxf(fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y))), fmaxf(fabsf(v.z), fabsf(v.w)));
}
__device__ __forceinline__ bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Largest |a[i]| into *bits: 16-byte loads, four in flight per thread (a scalar tail, and a scalar
// walk for unaligned arrays).
__global__ void __launch_bounds__(256) k_absmax(const float* __restrict__ a, size_t n, unsigned* __restrict__ bits) {
  float m = 0.f;
  const size_t T = (size_t)gridDim.x * blockDim.x, tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t n4 = aligned16(a) ? n / 4 : 0;
  const float4* a4 = reinterpret_cast<const float4*>(a);
  size_t i = tid;
  for (; i + 3 * T < n4; i += 4 * T) {
    const float4 v0 = a4[i], v1 = a4[i + T], v2 = a4[i + 2 * T], v3 = a4[i + 3 * T];
    m = absmax4(absmax4(absmax4(absmax4(m, v0), v1), v2), v3);
  }
  for (; i < n4; i += T) m = absmax4(m, a4[i]);
  for (size_t j = 4 * n4 + tid; j < n; j += T) m = fmaxf(m, fabsf(a[j]));
  block_max_to(m, bits);
}

// Fixed-point paint scale for deposits of at most |mul| * (*bits as float); *bits re-zeroed.
__global__ void k_set_scale(unsigned* __restrict__ bits, float mul, float* __restrict__ scale) {
  const float m = __uint_as_float(*bits) * fabsf(mul);
  int e = 0;
  if (m > 0.f) (void)frexpf(m, &e);
  e = max(-90, min(e, 90));
  scale[0] = ldexpf(1.0f, 30 - e);
  scale[1] = ldexpf(1.0f, e - 30);
  *bits = 0u;
}

// out = S^T U = -sum_d S_d U_d, S_d the 4-point stencil 8 (u[+1] - u[-1]) - (u[+2] - u[-2]) along d;
// component d of U at U[d] + cell * ustride. One workgroup per (x, y) row.
__global__ void __launch_bounds__(256) k_div4(const float* __restrict__ U0, const float* __restrict__ U1,
                                              const float* __restrict__ U2, int ustride, float* __restrict__ out,
                                              int NX, int NY, int NZ) {
  const unsigned G = gridDim.x, b = blockIdx.x;
  const unsigned row = (G % 8u == 0u) ? (b % 8u) * (G / 8u) + b / 8u : b;
  const int x = (int)(row / (unsigned)NY), y = (int)(row % (unsigned)NY);
  auto wr = [](int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
  const size_t r0 = (size_t)row * NZ;
  const size_t xp1 = ((size_t)wr(x + 1, NX) * NY + y) * NZ, xm1 = ((size_t)wr(x - 1, NX) * NY + y) * NZ;
  const size_t xp2 = ((size_t)wr(x + 2, NX) * NY + y) * NZ, xm2 = ((size_t)wr(x - 2, NX) * NY + y) * NZ;
  const size_t yp1 = ((size_t)x * NY + wr(y + 1, NY)) * NZ, ym1 = ((size_t)x * NY + wr(y - 1, NY)) * NZ;
  const size_t yp2 = ((size_t)x * NY + wr(y + 2, NY)) * NZ, ym2 = ((size_t)x * NY + wr(y - 2, NY)) * NZ;
  const size_t st = (size_t)ustride;
  for (int z = threadIdx.x; z < NZ; z += blockDim.x) {
    const float gx = 8.0f * (U0[(xp1 + z) * st] - U0[(xm1 + z) * st]) - (U0[(xp2 + z) * st] - U0[(xm2 + z) * st]);
    const float gy = 8.0f * (U1[(yp1 + z) * st] - U1[(ym1 + z) * st]) - (U1[(yp2 + z) * st] - U1[(ym2 + z) * st]);
    const float gz = 8.0f * (U2[(r0 + wr(z + 1, NZ)) * st] - U2[(r0 + wr(z - 1, NZ)) * st]) -
                     (U2[(r0 + wr(z + 2, NZ)) * st] - U2[(r0 + wr(z - 2, NZ)) * st]);
    out[r0 + z] = -(gx + gy + gz);
  }
}

// k_div4 for separate component meshes (ustride 1), marching along x like the gradient stencil: a
// workgroup takes yb consecutive rows (x, y0 .. y0 + yb - 1), tpr = NZ / 4 threads per row with 4 cells
// each, and walks xb x planes. U0[x - 2 .. x + 2] of a thread's cells stay in registers; U1's plane
// rows y0 - 2 .. y0 + vrows + 1 and U2's own rows sit in LDS for the y and z neighbours. The same
// expressions as k_div4 (bitwise the same output).
constexpr int DIV4_ROWS_FLOATS = 5120;  // (yb + 4) * NZ with yb = 256 / tpr, for every NZ up to 1024
__global__ void __launch_bounds__(256) k_div4m(const float* __restrict__ U0, const float* __restrict__ U1,
                                               const float* __restrict__ U2, float* __restrict__ out, int NX, int NY,
                                               int NZ, int tpr, int yb, int xb, int nyb) {
  __shared__ float4 p1[DIV4_ROWS_FLOATS / 4];
  __shared__ float4 p2[1024 / 4];  // yb * NZ = 4 * tpr * yb <= 1024
  const float* r1 = reinterpret_cast<const float*>(p1);
  const float* r2 = reinterpret_cast<const float*>(p2);
  const unsigned Lb = xcd_logical(blockIdx.x, gridDim.x);
  const int by = (int)(Lb % (unsigned)nyb), bx = (int)(Lb / (unsigned)nyb);
  const int me = (int)threadIdx.x, nt = (int)blockDim.x;
  const int lr = me / tpr, t = me - lr * tpr, z0 = 4 * t;
  const int y0 = by * yb, x0 = bx * xb;
  const int vrows = min(yb, NY - y0);
  const bool valid = lr < vrows;
  const int y = valid ? y0 + lr : y0;
  const int xend = min(x0 + xb, NX);
  auto wr = [](int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
  auto ld = [&](const float* U, int xx, int yy) {
    return *reinterpret_cast<const float4*>(U + ((size_t)xx * NY + yy) * NZ + z0);
  };
  float4 w0 = ld(U0, wr(x0 - 2, NX), y), w1 = ld(U0, wr(x0 - 1, NX), y), w2 = ld(U0, x0, y),
         w3 = ld(U0, wr(x0 + 1, NX), y);
  const int hq = 4 * tpr;
  for (int x = x0; x < xend; ++x) {
    const float4 w4 = ld(U0, wr(x + 2, NX), y);
    if (valid) {
      p1[((lr + 2) * NZ + z0) / 4] = ld(U1, x, y);
      p2[(lr * NZ + z0) / 4] = ld(U2, x, y);
    }
    for (int q = me; q < hq; q += nt) {  // U1 halo rows y0 - 2, y0 - 1, y0 + vrows, y0 + vrows + 1
      const int h = q / tpr, c = q - h * tpr;
      const int ly = h < 2 ? h : vrows + h;
      const int yy = wr(y0 - 2 + ly, NY);
      p1[(ly * NZ) / 4 + c] = *reinterpret_cast<const float4*>(U1 + ((size_t)x * NY + yy) * NZ + 4 * c);
    }
    __syncthreads();
    const float* a1 = r1 + (lr + 2) * NZ;
    const float* a2 = r2 + lr * NZ;
    float o[4];
    const float ax1[4] = {w3.x, w3.y, w3.z, w3.w}, bx1[4] = {w1.x, w1.y, w1.z, w1.w};
    const float ax2[4] = {w4.x, w4.y, w4.z, w4.w}, bx2[4] = {w0.x, w0.y, w0.z, w0.w};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int z = z0 + m;
      const float gx = 8.0f * (ax1[m] - bx1[m]) - (ax2[m] - bx2[m]);
      const float gy = 8.0f * (a1[z + NZ] - a1[z - NZ]) - (a1[z + 2 * NZ] - a1[z - 2 * NZ]);
      const float gz = 8.0f * (a2[wr(z + 1, NZ)] - a2[wr(z - 1, NZ)]) - (a2[wr(z + 2, NZ)] - a2[wr(z - 2, NZ)]);
      o[m] = -(gx + gy + gz);
    }
    if (valid) *reinterpret_cast<float4*>(out + ((size_t)x * NY + y) * NZ + z0) = make_float4(o[0], o[1], o[2], o[3]);
    __syncthreads();  // p1 and p2 are rewritten for the next plane
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = w4;
  }
}

// block sum of a double into part[blockIdx.x] (k_finish_sum adds the partials in a fixed order:
// deterministic, and no same-address atomics from thousands of workgroups)
__device__ __forceinline__ void block_sum_store(double v, double* part_out) {
  block_sum_to(v, part_out + blockIdx.x);
}

// *dst += sum of part[0..n) (one workgroup of 256)
__global__ void __launch_bounds__(256) k_finish_sum(const double* __restrict__ part, int n, double* __restrict__ dst) {
  __shared__ double s[256];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) t += part[i];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *dst += s[0];
}

// Position cotangent of readout and/or paint at the particles:
//   xbar_p (+)= sum_c grad w_c(x_p) * ( u_p . M[c]  +  wt_p * rhob[c] )
// with u_p = umul * u[p] (NC components, HAS_U: readout of the NC-component mesh M), rhob the
// density cotangent of a paint with particle weights wt (HAS_RHO). WBAR: wbar_p = sum_c w_c rhob[c]
// (the weight cotangent). dot (HAS_U, non-null): += sum_p u[p] . F_p with F_p = sum_c w_c M[c].
template <int NC, bool HAS_U, bool HAS_RHO, bool WBAR>
__global__ void __launch_bounds__(256) k_vjp_gather(PartArgs a, const float* __restrict__ u, float umul,
                                                    const float* __restrict__ M, const float* __restrict__ rhob,
                                                    const float* __restrict__ w, float w0, float* __restrict__ xbar,
                                                    int accumulate, float* __restrict__ wbar, double* __restrict__ part,
                                                    float* __restrict__ vkick, float ckick,
                                                    int NX, int NY, int NZ, int rx, int ry, int rz) {
  const unsigned np = (unsigned)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  LatticeWalk lw;
  lw.init(ch.begin + threadIdx.x, blockDim.x, a.npy, a.npz);
  double local = 0.0;
  const bool kick = HAS_U && NC == 3 && vkick != nullptr;
  for (unsigned p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x, lw.step()) {
    const float3 x = ld3(a.x, p);
    // loaded with x, ahead of the corner gathers: one memory round trip fewer per particle, for 78
    // VGPRs against 63 (512^3 particles on a 1024^3 mesh: forward + VJP 4.08-4.17 -> 3.83-3.87 s)
    const float3 v_in = kick ? ld3(vkick, p) : make_float3(0.f, 0.f, 0.f);
    const float3 xb_in = accumulate ? ld3(xbar, p) : make_float3(0.f, 0.f, 0.f);
    Corner1D cx, cy, cz;
    float fx, fy, fz;
    if (a.mode == 0) {
      cx = corner_rel(lw.i * rx, x.x, NX);
      cy = corner_rel(lw.j * ry, x.y, NY);
      cz = corner_rel(lw.k * rz, x.z, NZ);
      fx = x.x - floorf(x.x);
      fy = x.y - floorf(x.y);
      fz = x.z - floorf(x.z);
    } else {
      cx = corner_abs(x.x, NX);
      cy = corner_abs(x.y, NY);
      cz = corner_abs(x.z, NZ);
      fx = x.x - floorf(x.x);
      fy = x.y - floorf(x.y);
      fz = x.z - floorf(x.z);
    }
    const float dwx[2] = {-sgnf(fx), -sgnf(fx - 1.0f)}, dwy[2] = {-sgnf(fy), -sgnf(fy - 1.0f)},
                dwz[2] = {-sgnf(fz), -sgnf(fz - 1.0f)};
    const float wxv[2] = {cx.w0, cx.w1}, wyv[2] = {cy.w0, cy.w1}, wzv[2] = {cz.w0, cz.w1};
    const Cic q = cic_from(cx, cy, cz, NY, NZ);
    float ur[NC > 0 ? NC : 1] = {}, uu[NC > 0 ? NC : 1] = {};
    if (HAS_U) {
#pragma unroll
      for (int d = 0; d < NC; ++d) {
        ur[d] = u[(size_t)p * NC + d];
        uu[d] = ur[d] * umul;
      }
    }
    const float wt = HAS_RHO ? (w ? w[p] : w0) : 0.f;
    float gx = 0.f, gy = 0.f, gz = 0.f, ws = 0.f, Fd = 0.f;
    float Fv[NC > 0 ? NC : 1] = {};
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      float sv = 0.f;
      if (HAS_U) {
        float um = 0.f, uf = 0.f;
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          const float m = M[(size_t)q.off[o] * NC + d];
          um += uu[d] * m;
          uf += ur[d] * m;
          Fv[d] = Fv[d] + m * q.w[o];
        }
        sv += um;
        Fd += q.w[o] * uf;
      }
      if (HAS_RHO) {
        const float r = rhob[q.off[o]];
        sv += wt * r;
        if (WBAR) ws += q.w[o] * r;
      }
      gx += sv * ((dwx[ox] * wyv[oy]) * wzv[oz]);
      gy += sv * ((wxv[ox] * dwy[oy]) * wzv[oz]);
      gz += sv * ((wxv[ox] * wyv[oy]) * dwz[oz]);
    }
    float3 g = make_float3(gx, gy, gz);
    if (accumulate) g = make_float3(xb_in.x + g.x, xb_in.y + g.y, xb_in.z + g.z);
    st3(xbar, p, g);
    if (kick)  // v += F ck, F summed in the readout's corner order
      st3(vkick, p, axpy3(v_in, make_float3(Fv[0], Fv[1], Fv[2 % NC]), ckick));
    if (WBAR) wbar[p] = ws;
    if (HAS_U) local += (double)Fd;
  }
  if (HAS_U && part != nullptr) block_sum_store(local, part);
}

// Reverse-adjoint preparation of a KKD step (v += F ck; x += v cd) before the state is reversed:
// dot += <xbar, v>, vbar += cd * xbar, *bits = max |vbar| (for the deposit scale; skipped when bits
// is null). Also the transpose of any drift x += v cd on its own.
__global__ void __launch_bounds__(256) k_adj_prep(const float* __restrict__ xbar, float* __restrict__ vbar,
                                                  const float* __restrict__ v, size_t n, float cd,
                                                  double* __restrict__ part, unsigned* __restrict__ bits) {
  double local = 0.0;
  float m = 0.f;
  const size_t T = (size_t)gridDim.x * blockDim.x, tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  // 16-byte lanes, two in flight per array and thread, when all three arrays allow it; scalar tail
  const size_t n4 = (aligned16(xbar) && aligned16(vbar) && aligned16(v)) ? n / 4 : 0;
  const float4* x4 = reinterpret_cast<const float4*>(xbar);
  const float4* v4 = reinterpret_cast<const float4*>(v);
  float4* b4 = reinterpret_cast<float4*>(vbar);
  auto one4 = [&](const float4 xb, const float4 vv, const float4 bb) {
    local += (double)xb.x * (double)vv.x;
    local += (double)xb.y * (double)vv.y;
    local += (double)xb.z * (double)vv.z;
    local += (double)xb.w * (double)vv.w;
    const float4 nb = make_float4(bb.x + cd * xb.x, bb.y + cd * xb.y, bb.z + cd * xb.z, bb.w + cd * xb.w);
    m = absmax4(m, nb);
    return nb;
  };
  size_t i = tid;
  for (; i + T < n4; i += 2 * T) {
    const float4 xa = x4[i], xc = x4[i + T], va = v4[i], vc = v4[i + T], ba = b4[i], bc = b4[i + T];
    b4[i] = one4(xa, va, ba);
    b4[i + T] = one4(xc, vc, bc);
  }
  for (; i < n4; i += T) b4[i] = one4(x4[i], v4[i], b4[i]);
  for (size_t j = 4 * n4 + tid; j < n; j += T) {
    const float xb = xbar[j];
    local += (double)xb * (double)v[j];
    const float vb = vbar[j] + cd * xb;
    vbar[j] = vb;
    m = fmaxf(m, fabsf(vb));
  }
  if (bits) block_max_to(m, bits);
  block_sum_store(local, part);
}

// out <- x + v*c (out may alias x), as the fused step's drift computes it
__global__ void __launch_bounds__(256) k_axpy(float* out, const float* x, const float* __restrict__ v, size_t n,
                                              float c) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = x[i] + v[i] * c;
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
  const bool kick = NC == 3 && (what & 2);
  for (unsigned p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x, lw.step()) {
    float3 x = ld3(a.x, p);
    // the velocity loads with x, ahead of the corner gathers (one memory round trip fewer)
    const float3 v_in = kick ? ld3(a.v, p) : make_float3(0.f, 0.f, 0.f);
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
      if (kick) {
        float3 v = v_in;
        v = axpy3(v, make_float3(f[0], f[1], f[2]), ck);
        st3(a.v, p, v);
        if (what & 4) {
          x = axpy3(x, v, cd);
          st3(a.x, p, x);
        }
      } else if (what & 8) {  // x += F * ck (a displacement correction)
        st3(a.x, p, axpy3(x, make_float3(f[0], f[1], f[2]), ck));
      }
    }
  }
}

// Readout staged through LDS. A workgroup takes an RB_I x RB_J x RB_K Lagrangian block of particles
// and bounds the cells of their CIC corners (unwrapped cell coordinates). When that box holds at
// most RO_CAP cells it is loaded into LDS with z-contiguous reads (each mesh cell once per
// workgroup instead of once per corner that touches it) and the 8 corners are gathered from LDS;
// larger boxes (shell crossings, wide spreads) gather from global memory as k_readout does. The
// weights and the summation order are those of k_readout: the results are bitwise identical.
constexpr int RB_I = 4, RB_J = 8, RB_K = 8, RO_CAP = 1664;

__device__ __forceinline__ int base_cell(int mode, int lattice, float xc) {
  return (mode == 0 ? lattice : 0) + (int)floorf(xc);
}

// Wave-wide min / max: four DPP steps fused into v_min/v_max (quad swaps, half-row and row mirrors;
// bound_ctrl is harmless as every lane has a source) leave each 16-lane row's extreme in all of its
// lanes; the four rows are then combined from lanes 0, 16, 32, 48 on the scalar unit. All lanes
// must be active.
template <bool MAX>
__device__ __forceinline__ int wave_minmax(int v) {
  auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
  v = op(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));   // quad_perm [1,0,3,2]
  v = op(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));   // quad_perm [2,3,0,1]
  v = op(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));  // row_half_mirror
  v = op(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true));  // row_mirror
  const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
  const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
  return op(op(r0, r1), op(r2, r3));
}

// base cell (unwrapped) and the two CIC weights along one axis, with the expressions of corner_rel
// (relative mode) and corner_abs (absolute mode)
struct Axis1D {
  int b;
  float w0, w1;
};
__device__ __forceinline__ Axis1D axis_of(int mode, int lattice, float xc) {
  const float fl = floorf(xc);
  Axis1D r;
  if (mode == 0) {
    const float fr = xc - fl;
    r.b = lattice + (int)fl;
    r.w0 = 1.0f - fabsf(fr);
    r.w1 = 1.0f - fabsf(fr - 1.0f);
  } else {
    r.b = (int)fl;
    r.w0 = 1.0f - fabsf(xc - fl);
    r.w1 = 1.0f - fabsf(xc - (fl + 1.0f));
  }
  return r;
}

template <int NC>
__global__ void __launch_bounds__(256) k_readout_lds(PartArgs a, const float* __restrict__ mesh, float* __restrict__ out,
                                                     float ck, float cd, int what, int NX, int NY, int NZ, int rx,
                                                     int ry, int rz, int nbj, int nbk, FDiv dj, FDiv dk, Slots sl) {
  __shared__ float sm[RO_CAP * NC];
  __shared__ int red[6][4];
  const unsigned zs = blockIdx.z;
  if (!slot_live(sl, zs)) return;
  a.x = at_slot(a.x, sl.stride, zs);
  a.v = at_slot(a.v, sl.stride, zs);
  mesh = at_slot(mesh, sl.stride, zs);
  out = at_slot(out, sl.stride, zs);
  if (sl.coef) {
    ck = sl.coef[4 * zs + 1];
    cd = sl.coef[4 * zs + 2];
  }
  const unsigned b = xcd_logical(blockIdx.x, gridDim.x);
  const unsigned bq = fdiv(b, dk), bi_ = fdiv(bq, dj);
  const int bk = (int)(b - bq * (unsigned)nbk), bj = (int)(bq - bi_ * (unsigned)nbj), bi = (int)bi_;
  const int t = threadIdx.x;
  const int i = bi * RB_I + t / (RB_J * RB_K), j = bj * RB_J + (t / RB_K) % RB_J, k = bk * RB_K + t % RB_K;
  const unsigned p = ((unsigned)i * a.npy + j) * a.npz + k;
  float3 x = ld3(a.x, p);
  const Axis1D ax = axis_of(a.mode, i * rx, x.x), ay = axis_of(a.mode, j * ry, x.y), az = axis_of(a.mode, k * rz, x.z);
  const int wv = t >> 6;
  {
    const int r0 = wave_minmax<false>(ax.b), r1 = wave_minmax<false>(ay.b), r2 = wave_minmax<false>(az.b);
    const int r3 = wave_minmax<true>(ax.b), r4 = wave_minmax<true>(ay.b), r5 = wave_minmax<true>(az.b);
    if ((t & 63) == 0) {
      red[0][wv] = r0;
      red[1][wv] = r1;
      red[2][wv] = r2;
      red[3][wv] = r3;
      red[4][wv] = r4;
      red[5][wv] = r5;
    }
  }
  __syncthreads();
  int mn[3], mx[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    mn[c] = __builtin_amdgcn_readfirstlane(min(min(red[c][0], red[c][1]), min(red[c][2], red[c][3])));
    mx[c] = __builtin_amdgcn_readfirstlane(
                max(max(red[3 + c][0], red[3 + c][1]), max(red[3 + c][2], red[3 + c][3]))) + 1;  // upper corner
  }
  // box extents; each is checked against the cap before the product so that 32 bits suffice
  const unsigned ex = (unsigned)(mx[0] - mn[0] + 1), ey = (unsigned)(mx[1] - mn[1] + 1),
                 ez = (unsigned)(mx[2] - mn[2] + 1);
  const bool fits = ex <= (unsigned)RO_CAP && ey <= (unsigned)RO_CAP && ez <= (unsigned)RO_CAP &&
                    ex * ey <= (unsigned)RO_CAP && ex * ey * ez <= (unsigned)RO_CAP;
  float f[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) f[c] = 0.f;
  if (fits) {
    const int EY = (int)ey, EZ = (int)ez, ncell = (int)(ex * ey * ez);
    // v_rcp is within 1 ulp of 1/e: div_small's margin holds for q < 2^21
    const float iy = __builtin_amdgcn_rcpf((float)EY), iz = __builtin_amdgcn_rcpf((float)EZ);
    // the box origin wrapped once; inside the box one conditional subtraction wraps each axis when
    // the box is no longer than the mesh, an exact modulo otherwise (meshes smaller than the box)
    const int ox = wrap_cell(mn[0], NX), oy = wrap_cell(mn[1], NY), oz = wrap_cell(mn[2], NZ);
    auto cell_of = [&](int q, bool exact) {
      int qz, qy;
      const int qx = div_small(div_small(q, iz, EZ, qz), iy, EY, qy);
      int gx = ox + qx, gy = oy + qy, gz = oz + qz;
      if (!exact) {
        gx -= gx >= NX ? NX : 0;
        gy -= gy >= NY ? NY : 0;
        gz -= gz >= NZ ? NZ : 0;
      } else {
        gx %= NX;
        gy %= NY;
        gz %= NZ;
      }
      return mesh + (size_t)(((unsigned)gx * NY + (unsigned)gy) * NZ + (unsigned)gz) * NC;
    };
    if (ex <= (unsigned)NX && ey <= (unsigned)NY && ez <= (unsigned)NZ) {
      // two cells per thread and trip, both loads issued before either LDS store
      int q = t;
      for (; q + 256 < ncell; q += 512) {
        const float* m0 = cell_of(q, false);
        const float* m1 = cell_of(q + 256, false);
        float a0[NC], a1[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          a0[c] = m0[c];
          a1[c] = m1[c];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          sm[q * NC + c] = a0[c];
          sm[(q + 256) * NC + c] = a1[c];
        }
      }
      if (q < ncell) {
        const float* m0 = cell_of(q, false);
#pragma unroll
        for (int c = 0; c < NC; ++c) sm[q * NC + c] = m0[c];
      }
    } else {
      for (int q = t; q < ncell; q += 256) {
        const float* m = cell_of(q, true);
#pragma unroll
        for (int c = 0; c < NC; ++c) sm[q * NC + c] = m[c];
      }
    }
    __syncthreads();
    const int lx = ax.b - mn[0], ly = ay.b - mn[1], lz = az.b - mn[2];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox_ = o & 1, oy_ = (o >> 1) & 1, oz_ = (o >> 2) & 1;
      const float w = ((ox_ ? ax.w1 : ax.w0) * (oy_ ? ay.w1 : ay.w0)) * (oz_ ? az.w1 : az.w0);
      const float* m = sm + (((lx + ox_) * EY + (ly + oy_)) * EZ + (lz + oz_)) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = f[c] + m[c] * w;
    }
  } else {
    const int x0 = wrap_idx(ax.b, NX), x1 = wrap_idx(ax.b + 1, NX), y0 = wrap_idx(ay.b, NY),
              y1 = wrap_idx(ay.b + 1, NY), z0 = wrap_idx(az.b, NZ), z1 = wrap_idx(az.b + 1, NZ);
    const unsigned nyz = (unsigned)NY * NZ;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox_ = o & 1, oy_ = (o >> 1) & 1, oz_ = (o >> 2) & 1;
      const float w = ((ox_ ? ax.w1 : ax.w0) * (oy_ ? ay.w1 : ay.w0)) * (oz_ ? az.w1 : az.w0);
      const unsigned off = (unsigned)(ox_ ? x1 : x0) * nyz + (unsigned)(oy_ ? y1 : y0) * NZ + (unsigned)(oz_ ? z1 : z0);
      const float* m = mesh + (size_t)off * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) f[c] = f[c] + m[c] * w;
    }
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
    } else if (what & 8) {  // x += F * ck (a displacement correction)
      st3(a.x, p, axpy3(x, make_float3(f[0], f[1], f[2]), ck));
    }
  }
}

// k_vjp_gather on the readout's Lagrangian blocks (RB_I x RB_J x RB_K particles, one per thread): the
// block's corner box of M (NC floats per cell) and rhob (one) is staged in LDS when it holds at most
// RO_CAP cells, else the block gathers from global memory. Same weights, corner order and sums as
// k_vjp_gather: bitwise the same results. Blocks are walked grid-stride so that part[] keeps one
// partial per workgroup (at most kRedMax).
template <int NC, bool HAS_U, bool HAS_RHO, bool WBAR>
__global__ void __launch_bounds__(256) k_vjp_gather_lds(PartArgs a, const float* __restrict__ u, float umul,
                                                        const float* __restrict__ M, const float* __restrict__ rhob,
                                                        const float* __restrict__ w, float w0,
                                                        float* __restrict__ xbar, int accumulate,
                                                        float* __restrict__ wbar, double* __restrict__ part,
                                                        float* __restrict__ vkick, float ckick, int NX, int NY,
                                                        int NZ, int rx, int ry, int rz, int nbj, int nbk,
                                                        unsigned nblocks, FDiv dj, FDiv dk) {
  constexpr int S = (HAS_U ? NC : 0) + (HAS_RHO ? 1 : 0);  // floats staged per cell: M components, then rhob
  __shared__ float sm[RO_CAP * S];
  __shared__ int red[6][4];
  const int t = threadIdx.x, wv = t >> 6;
  double local = 0.0;
  for (unsigned b = xcd_logical(blockIdx.x, gridDim.x); b < nblocks; b += gridDim.x) {
    const unsigned bq = fdiv(b, dk), bi_ = fdiv(bq, dj);
    const int bk = (int)(b - bq * (unsigned)nbk), bj = (int)(bq - bi_ * (unsigned)nbj), bi = (int)bi_;
    const int i = bi * RB_I + t / (RB_J * RB_K), j = bj * RB_J + (t / RB_K) % RB_J, k = bk * RB_K + t % RB_K;
    const unsigned p = ((unsigned)i * a.npy + j) * a.npz + k;
    const float3 x = ld3(a.x, p);
    const Axis1D ax = axis_of(a.mode, i * rx, x.x), ay = axis_of(a.mode, j * ry, x.y), az = axis_of(a.mode, k * rz, x.z);
    {
      const int r0 = wave_minmax<false>(ax.b), r1 = wave_minmax<false>(ay.b), r2 = wave_minmax<false>(az.b);
      const int r3 = wave_minmax<true>(ax.b), r4 = wave_minmax<true>(ay.b), r5 = wave_minmax<true>(az.b);
      if ((t & 63) == 0) {
        red[0][wv] = r0;
        red[1][wv] = r1;
        red[2][wv] = r2;
        red[3][wv] = r3;
        red[4][wv] = r4;
        red[5][wv] = r5;
      }
    }
    __syncthreads();
    int mn[3], mx[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      mn[c] = __builtin_amdgcn_readfirstlane(min(min(red[c][0], red[c][1]), min(red[c][2], red[c][3])));
      mx[c] = __builtin_amdgcn_readfirstlane(
                  max(max(red[3 + c][0], red[3 + c][1]), max(red[3 + c][2], red[3 + c][3]))) + 1;
    }
    const unsigned ex = (unsigned)(mx[0] - mn[0] + 1), ey = (unsigned)(mx[1] - mn[1] + 1),
                   ez = (unsigned)(mx[2] - mn[2] + 1);
    const bool fits = ex <= (unsigned)RO_CAP && ey <= (unsigned)RO_CAP && ez <= (unsigned)RO_CAP &&
                      ex * ey <= (unsigned)RO_CAP && ex * ey * ez <= (unsigned)RO_CAP;
    const int EY = (int)ey, EZ = (int)ez;
    if (fits) {
      const int ncell = (int)(ex * ey * ez);
      const float iy = __builtin_amdgcn_rcpf((float)EY), iz = __builtin_amdgcn_rcpf((float)EZ);
      const int gx0 = wrap_cell(mn[0], NX), gy0 = wrap_cell(mn[1], NY), gz0 = wrap_cell(mn[2], NZ);
      const bool exact = !(ex <= (unsigned)NX && ey <= (unsigned)NY && ez <= (unsigned)NZ);
      for (int q = t; q < ncell; q += 256) {
        int qz, qy;
        const int qx = div_small(div_small(q, iz, EZ, qz), iy, EY, qy);
        int gx = gx0 + qx, gy = gy0 + qy, gz = gz0 + qz;
        if (!exact) {
          gx -= gx >= NX ? NX : 0;
          gy -= gy >= NY ? NY : 0;
          gz -= gz >= NZ ? NZ : 0;
        } else {
          gx %= NX;
          gy %= NY;
          gz %= NZ;
        }
        const size_t cell = ((size_t)gx * NY + gy) * NZ + gz;
        if constexpr (HAS_U) {
#pragma unroll
          for (int d = 0; d < NC; ++d) sm[q * S + d] = M[cell * NC + d];
        }
        if constexpr (HAS_RHO) sm[q * S + S - 1] = rhob[cell];
      }
      __syncthreads();
    }
    const int lx = ax.b - mn[0], ly = ay.b - mn[1], lz = az.b - mn[2];
    const int x0 = wrap_idx(ax.b, NX), x1 = wrap_idx(ax.b + 1, NX), y0 = wrap_idx(ay.b, NY),
              y1 = wrap_idx(ay.b + 1, NY), z0 = wrap_idx(az.b, NZ), z1 = wrap_idx(az.b + 1, NZ);
    const unsigned nyz = (unsigned)NY * NZ;
    const float fx = x.x - floorf(x.x), fy = x.y - floorf(x.y), fz = x.z - floorf(x.z);
    const float dwx[2] = {-sgnf(fx), -sgnf(fx - 1.0f)}, dwy[2] = {-sgnf(fy), -sgnf(fy - 1.0f)},
                dwz[2] = {-sgnf(fz), -sgnf(fz - 1.0f)};
    const float wxv[2] = {ax.w0, ax.w1}, wyv[2] = {ay.w0, ay.w1}, wzv[2] = {az.w0, az.w1};
    float ur[NC > 0 ? NC : 1] = {}, uu[NC > 0 ? NC : 1] = {};
    if (HAS_U) {
#pragma unroll
      for (int d = 0; d < NC; ++d) {
        ur[d] = u[(size_t)p * NC + d];
        uu[d] = ur[d] * umul;
      }
    }
    const float wt = HAS_RHO ? (w ? w[p] : w0) : 0.f;
    float gxs = 0.f, gys = 0.f, gzs = 0.f, ws = 0.f, Fd = 0.f;
    float Fv[NC > 0 ? NC : 1] = {};
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      const float wgt = ((ox ? ax.w1 : ax.w0) * (oy ? ay.w1 : ay.w0)) * (oz ? az.w1 : az.w0);
      const int lc = ((lx + ox) * EY + (ly + oy)) * EZ + (lz + oz);
      const size_t gc = (size_t)(ox ? x1 : x0) * nyz + (unsigned)(oy ? y1 : y0) * NZ + (unsigned)(oz ? z1 : z0);
      float sv = 0.f;
      if (HAS_U) {
        float um = 0.f, uf = 0.f;
#pragma unroll
        for (int d = 0; d < NC; ++d) {
          const float m = fits ? sm[lc * S + d] : M[gc * NC + d];
          um += uu[d] * m;
          uf += ur[d] * m;
          Fv[d] = Fv[d] + m * wgt;
        }
        sv += um;
        Fd += wgt * uf;
      }
      if (HAS_RHO) {
        const float r = fits ? sm[lc * S + S - 1] : rhob[gc];
        sv += wt * r;
        if (WBAR) ws += wgt * r;
      }
      gxs += sv * ((dwx[ox] * wyv[oy]) * wzv[oz]);
      gys += sv * ((wxv[ox] * dwy[oy]) * wzv[oz]);
      gzs += sv * ((wxv[ox] * wyv[oy]) * dwz[oz]);
    }
    float3 g = make_float3(gxs, gys, gzs);
    if (accumulate) {
      const float3 o = ld3(xbar, p);
      g = make_float3(o.x + g.x, o.y + g.y, o.z + g.z);
    }
    st3(xbar, p, g);
    if (HAS_U && NC == 3 && vkick != nullptr)  // v += F ck, F summed in the readout's corner order
      st3(vkick, p, axpy3(ld3(vkick, p), make_float3(Fv[0], Fv[1], Fv[2 % NC]), ckick));
    if (WBAR) wbar[p] = ws;
    if (HAS_U) local += (double)Fd;
    __syncthreads();  // red and sm are rewritten by the next block
  }
  if (HAS_U && part != nullptr) block_sum_store(local, part);
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

// Position cotangent of k_paint_flat for the map cotangent map_bar (the finalised map's, so the
// 1 / pixel-volume scale enters here): pos_bar (+)= scale * wt * d/dpos sum_c w_c map_bar[c] over
// the 4 corners, the weight derivatives -sgn(p - corner) as in k_vjp_gather, x and y chained through
// the flat-grid scaling; the slab selection is a step and z gets nothing.
__global__ void __launch_bounds__(256) k_paint_flat_vjp(PartArgs a, const float* __restrict__ w, float w0, FlatArgs fa,
                                                        const float* __restrict__ map_bar, float scale,
                                                        float* __restrict__ pos_bar, int accumulate, int rx, int ry,
                                                        int rz) {
  const unsigned np = (unsigned)a.npx * a.npy * a.npz;
  const Chunk ch = my_chunk(np);
  LatticeWalk lw;
  lw.init(ch.begin + threadIdx.x, blockDim.x, a.npy, a.npz);
  for (unsigned p = ch.begin + threadIdx.x; p < ch.end; p += blockDim.x, lw.step()) {
    float3 x = ld3(a.x, p);
    if (a.mode == 0) x = make_float3(x.x + (float)(lw.i * rx), x.y + (float)(lw.j * ry), x.z + (float)(lw.k * rz));
    float gx = 0.f, gy = 0.f;
    if (x.z > fa.zlo && x.z <= fa.zhi) {
      const float wt = (w ? w[p] : w0) * scale;
      float px = x.x * fa.sxf, py = x.y * fa.syf;
      px = px - floorf(px / fa.fnx) * fa.fnx;
      py = py - floorf(py / fa.fny) * fa.fny;
      const Corner1D cx = corner_abs(px, fa.fnx), cy = corner_abs(py, fa.fny);
      const float fx = px - floorf(px), fy = py - floorf(py);
      const float dwx[2] = {-sgnf(fx), -sgnf(fx - 1.0f)}, dwy[2] = {-sgnf(fy), -sgnf(fy - 1.0f)};
      const float wxv[2] = {cx.w0, cx.w1}, wyv[2] = {cy.w0, cy.w1};
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        const int ox = o & 1, oy = o >> 1;
        const float m = map_bar[(size_t)(ox ? cx.i1 : cx.i0) * fa.fny + (oy ? cy.i1 : cy.i0)];
        gx += dwx[ox] * wyv[oy] * m;
        gy += wxv[ox] * dwy[oy] * m;
      }
      gx *= wt * fa.sxf;
      gy *= wt * fa.syf;
    }
    float3 g = make_float3(gx, gy, 0.f);
    if (accumulate) {
      const float3 o = ld3(pos_bar, p);
      g = make_float3(o.x + g.x, o.y + g.y, o.z);
    }
    st3(pos_bar, p, g);
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
// dx = D F at the particle lattice (cell (i rx, j ry, k rz) of the force mesh), pv = c dx
__global__ void k_lpt1(const float* __restrict__ F, float* __restrict__ dx, float* __restrict__ pv, int npx, int npy,
                       int npz, int rx, int ry, int rz, int NY, int NZ, float D, float c) {
  const size_t n = (size_t)npx * npy * npz;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const unsigned k = (unsigned)(p % (unsigned)npz), j = (unsigned)((p / (unsigned)npz) % (unsigned)npy),
                   i = (unsigned)(p / ((size_t)npz * npy));
    const size_t cell = ((size_t)i * rx * NY + (size_t)j * ry) * NZ + (size_t)k * rz;
    for (int d = 0; d < 3; ++d) {
      const float v = F[3 * cell + d] * D;
      dx[3 * p + d] = v;
      pv[3 * p + d] = v * c;
    }
  }
}

// 2LPT source delta2 = sum_{i<j} (H_ii H_jj - H_ij^2) from the interleaved Hessian components
__global__ void k_delta2(const float* __restrict__ H, float* __restrict__ rho, size_t n, int offdiag) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const float h0 = H[3 * q], h1 = H[3 * q + 1], h2 = H[3 * q + 2];
    if (offdiag)
      rho[q] = rho[q] - (h0 * h0 + h1 * h1 + h2 * h2);
    else
      rho[q] = h1 * h0 + h2 * (h0 + h1);
  }
}

// One LPT term at the lattice: dx (+)= D F, pv (+)= c (D F). (D, c) are columns 2 term and 2 term + 1
// of g.coef: constants (ntab == 1) or linear interpolation in the particle's comoving distance
// r = |lattice position * cell - observer| on the uniform grid of ntab points over [0, r_max].
template <bool ACC>
__global__ void k_lpt_apply(const float* __restrict__ F, float* __restrict__ dx, float* __restrict__ pv, int npx,
                            int npy, int npz, int rx, int ry, int rz, int NY, int NZ, LptGeom g, int term) {
  const size_t n = (size_t)npx * npy * npz;
  const double* cD = g.coef + (size_t)(2 * term) * g.ntab;
  const double* cV = g.coef + (size_t)(2 * term + 1) * g.ntab;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const unsigned k = (unsigned)(p % (unsigned)npz), j = (unsigned)((p / (unsigned)npz) % (unsigned)npy),
                   i = (unsigned)(p / ((size_t)npz * npy));
    const size_t cell = ((size_t)i * rx * NY + (size_t)j * ry) * NZ + (size_t)k * rz;
    float D, c;
    if (g.ntab == 1) {
      D = (float)cD[0];
      c = (float)cV[0];
    } else {
      const double px_ = (double)(i * rx) * g.cell[0] - g.obs[0];
      const double py_ = (double)(j * ry) * g.cell[1] - g.obs[1];
      const double pz_ = (double)(k * rz) * g.cell[2] - g.obs[2];
      const double r = sqrt(px_ * px_ + py_ * py_ + pz_ * pz_);
      double t = r / g.r_max * (double)(g.ntab - 1);
      t = t < 0.0 ? 0.0 : (t > (double)(g.ntab - 1) ? (double)(g.ntab - 1) : t);
      int it = (int)t;
      if (it > g.ntab - 2) it = g.ntab - 2;
      const double fr = t - (double)it;
      D = (float)(cD[it] + fr * (cD[it + 1] - cD[it]));
      c = (float)(cV[it] + fr * (cV[it + 1] - cV[it]));
    }
    for (int d = 0; d < 3; ++d) {
      const float v = F[3 * cell + d] * D;
      if (ACC) {
        dx[3 * p + d] += v;
        pv[3 * p + d] += v * c;
      } else {
        dx[3 * p + d] = v;
        pv[3 * p + d] = v * c;
      }
    }
  }
}

// transpose of k_lpt1 onto the force mesh: G[cell] = D (dx_bar + c pv_bar) of the particle on that
// lattice cell, 0 on the cells between lattice sites (every cell is written)
__global__ void k_lpt1_vjp(const float* __restrict__ dxb, const float* __restrict__ pvb, float* __restrict__ G,
                           int NX, int NY, int NZ, int rx, int ry, int rz, int npy, int npz, float D, float c) {
  const size_t n = (size_t)NX * NY * NZ;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const unsigned z = (unsigned)(q % (unsigned)NZ), y = (unsigned)((q / (unsigned)NZ) % (unsigned)NY),
                   x = (unsigned)(q / ((size_t)NZ * NY));
    float g[3] = {0.f, 0.f, 0.f};
    if (x % rx == 0 && y % ry == 0 && z % rz == 0) {
      const size_t p = ((size_t)(x / rx) * npy + y / ry) * npz + z / rz;
      for (int d = 0; d < 3; ++d) {
        const float a = dxb ? dxb[3 * p + d] : 0.f, b = pvb ? pvb[3 * p + d] : 0.f;
        g[d] = D * (a + c * b);
      }
    }
    for (int d = 0; d < 3; ++d) G[3 * q + d] = g[d];
  }
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

static FDiv make_fdiv(unsigned d) {
  FDiv f;
  f.d = d;
  int l = 0;
  while ((1ull << l) < d) ++l;  // ceil(log2 d)
  f.m = (unsigned)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  f.s1 = l > 0 ? 1 : 0;
  f.s2 = l > 0 ? l - 1 : 0;
  return f;
}

static int ilog2(int n) {
  int l = 0;
  while ((1 << (l + 1)) <= n) ++l;
  return l;
}

// Paint tiles of a plan: (min(nx,16), min(ny,16), min(nz,32)) cells; every supported axis length
// is a multiple of its tile length.
void pmx_paint_tiles(const pmx_plan_t* p, int& sx, int& sy, int& sz, int& ntiles) {
  sx = ilog2(p->nx < 16 ? p->nx : 16);
  sy = ilog2(p->ny < 16 ? p->ny : 16);
  sz = ilog2(p->nz < 32 ? p->nz : 32);
  ntiles = (p->nx >> sx) * (p->ny >> sy) * (p->nz >> sz);
}
int pmx_paint_face_cells(const pmx_plan_t* p) {
  int sx, sy, sz, ntiles;
  pmx_paint_tiles(p, sx, sy, sz, ntiles);
  const int BX = 1 << sx, BY = 1 << sy, BZ = 1 << sz;
  return (BY + 1) * (BZ + 1) + BX * (BZ + 1) + BX * BY;
}

static SortGeom sort_geom(const pmx_plan_t* p, const PartArgs& a, int& ntiles) {
  int rx, ry, rz;
  ratios(const_cast<pmx_plan_t*>(p), a, rx, ry, rz);
  SortGeom g;
  g.NX = p->nx;
  g.NY = p->ny;
  g.NZ = p->nz;
  pmx_paint_tiles(p, g.sx, g.sy, g.sz, ntiles);
  g.ntx = p->nx >> g.sx;
  g.nty = p->ny >> g.sy;
  g.ntz = p->nz >> g.sz;
  g.rx = a.mode == 0 ? rx : 1;
  g.ry = a.mode == 0 ? ry : 1;
  g.rz = a.mode == 0 ? rz : 1;
  g.nbx = (a.npx + SB_X - 1) / SB_X;
  g.nby = (a.npy + SB_Y - 1) / SB_Y;
  g.nbz = (a.npz + SB_Z - 1) / SB_Z;
  g.div_npz = make_fdiv((unsigned)a.npz);
  g.div_npy = make_fdiv((unsigned)a.npy);
  const bool pow2 = (a.npz & (a.npz - 1)) == 0 && (a.npy & (a.npy - 1)) == 0;
  g.lpz = pow2 ? ilog2(a.npz) : -1;
  g.lpy = pow2 ? ilog2(a.npy) : -1;
  auto is_pow2 = [](int n) { return (n & (n - 1)) == 0; };
  g.wmask = (is_pow2(p->nx) && is_pow2(p->ny) && is_pow2(p->nz)) ? p->nx - 1 : -1;
  return g;
}

// Window-count entries the sort keeps for a particle grid (one per window slot and sub-block).
size_t pmx_sort_wg_entries(int npx, int npy, int npz) {
  return (size_t)((npx + SB_X - 1) / SB_X) * ((npy + SB_Y - 1) / SB_Y) * ((npz + SB_Z - 1) / SB_Z) * SORT_SLOTS;
}

// Size the sort buffers for a particle grid of npx x npy x npz (grow only; synchronous). Adds the
// bytes newly allocated to `grown`.
int pmx_sort_reserve(pmx_plan_t* p, int npx, int npy, int npz, size_t& grown) {
  const size_t np = (size_t)npx * npy * npz;
  const size_t nsb = (size_t)((npx + SB_X - 1) / SB_X) * ((npy + SB_Y - 1) / SB_Y) * ((npz + SB_Z - 1) / SB_Z);
  if (p->ps_wg_cap < nsb * SORT_SLOTS) {
    if (p->ps_wgcnt) (void)hipFree(p->ps_wgcnt);
    grown -= p->ps_wg_cap * sizeof(unsigned);
    p->ps_wgcnt = nullptr;
    p->ps_wg_cap = 0;
    hipError_t e = hipMalloc((void**)&p->ps_wgcnt, nsb * SORT_SLOTS * sizeof(unsigned));
    if (e != hipSuccess) return pmx_set_error(4, std::string("paint sort counts: ") + hipGetErrorString(e));
    p->ps_wg_cap = nsb * SORT_SLOTS;
    grown += p->ps_wg_cap * sizeof(unsigned);
  }
  if (p->ps_seg_cap < np) {
    if (p->ps_seg) (void)hipFree(p->ps_seg);
    if (p->ps_pack) (void)hipFree(p->ps_pack);
    grown -= 2 * p->ps_seg_cap * sizeof(unsigned);
    p->ps_seg = p->ps_pack = nullptr;
    p->ps_seg_cap = 0;
    hipError_t e = hipMalloc((void**)&p->ps_seg, np * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc((void**)&p->ps_pack, np * sizeof(unsigned));
    if (e != hipSuccess) {
      if (p->ps_seg) (void)hipFree(p->ps_seg);
      p->ps_seg = nullptr;
      return pmx_set_error(4, std::string("paint sort segments: ") + hipGetErrorString(e));
    }
    p->ps_seg_cap = np;
    grown += 2 * np * sizeof(unsigned);
  }
  return 0;
}

// Sort the particles of `a` by paint tile ([a.v] pre-drift x += v * cd_pre first). The fixed-point
// scale for the following deposits comes from |w0| or, with w (wstride values per particle), from
// the largest |w|.
int pmx_paint_sort(pmx_plan_t* p, const PartArgs& a, float cd_pre, const float* w, int wstride, float w0,
                   hipStream_t s) {
  int ntiles;
  const SortGeom g = sort_geom(p, a, ntiles);
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  const size_t nsb = (size_t)g.nbx * g.nby * g.nbz;
  if (p->ps_wg_cap < nsb * SORT_SLOTS || p->ps_seg_cap < np)
    return pmx_set_error(4, "paint sort buffers hold " + std::to_string(p->ps_seg_cap) + " particles / " +
                                std::to_string(p->ps_wg_cap / SORT_SLOTS) + " sub-blocks, this grid needs " +
                                std::to_string(np) + " / " + std::to_string(nsb) +
                                ": call pmx_plan_reserve with its particle grid first");
  unsigned* wmax_bits = p->ps_misc;
  unsigned* far = p->ps_misc + 1;
  unsigned* far_last = p->ps_misc + 2;
  const unsigned S = (unsigned)p->nslots;
  if (S > 1 && (a.v || w)) return pmx_set_error(2, "a batched paint takes no pre-drift and no per-particle weights");
  hipLaunchKernelGGL(k_sort_count, dim3((unsigned)nsb, 1, S), dim3(256), 0, s, a, g, cd_pre, p->ps_count, p->ps_wgcnt,
                     far, w, wstride, wmax_bits, p->ps_pack, p->slots);
  int rc = pmx_check_launch("sort_count");
  if (rc) return rc;
  PartArgs b = a;
  b.v = nullptr;
  hipLaunchKernelGGL(k_sort_scan, dim3(1, 1, S), dim3(1024), 0, s, p->ps_count, p->ps_tcount, p->ps_offset,
                     p->ps_cursor, ntiles, w ? -1.0f : fabsf(w0), wmax_bits, p->ps_scale, far, far_last, p->slots);
  if ((rc = pmx_check_launch("sort_scan"))) return rc;
  hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned)nsb, 1, S), dim3(256), 0, s, b, g, p->ps_cursor, p->ps_wgcnt,
                     p->ps_pack, p->ps_seg, p->slots);
  return pmx_check_launch("sort_scatter");
}

// Deposit the last sorted particles into `mesh`: particle p carries w[p * wstride] * wmul (w != null)
// or w0; `is_signed` when any carried value may be negative.
int pmx_paint_deposit(pmx_plan_t* p, const PartArgs& a_in, const float* w, int wstride, float wmul, float w0,
                      int is_signed, float* mesh, int mstride, int accumulate, hipStream_t s, int ncomp,
                      size_t mcomp) {
  PartArgs a = a_in;
  a.v = nullptr;
  int ntiles;
  const SortGeom g = sort_geom(p, a, ntiles);
  const int BX = 1 << g.sx, BY = 1 << g.sy, BZ = 1 << g.sz;
  const int face_stride = pmx_paint_face_cells(p);
  const size_t fcomp = (size_t)ntiles * face_stride;
  const unsigned S = (unsigned)p->nslots;
  // components in parallel workgroups when the tiles alone are too few to fill the GPU (128^3: 256
  // tiles; one workgroup per component cuts the 3-component deposit 69 -> 53 us); with enough tiles,
  // one workgroup walks the components so their reads of the tile's particles stay in cache
  const unsigned gy = (ncomp > 1 && (size_t)ntiles * S < 1024) ? (unsigned)ncomp : 1u;
  if (is_signed)
    hipLaunchKernelGGL(k_paint_deposit<true>, dim3(ntiles, gy, S), dim3(512), 0, s, a, g, w, wstride, wmul, w0, ncomp,
                       p->ps_tcount, p->ps_offset, p->ps_seg, p->ps_scale, mesh, mstride, mcomp, p->ps_faces,
                       face_stride, fcomp, accumulate, p->slots);
  else
    hipLaunchKernelGGL(k_paint_deposit<false>, dim3(ntiles, gy, S), dim3(512), 0, s, a, g, w, wstride, wmul, w0, ncomp,
                       p->ps_tcount, p->ps_offset, p->ps_seg, p->ps_scale, mesh, mstride, mcomp, p->ps_faces,
                       face_stride, fcomp, accumulate, p->slots);
  int rc = pmx_check_launch("paint_deposit");
  if (rc) return rc;
  const int shell = BX * BY * BZ - (BX - 1) * (BY - 1) * (BZ - 1);
  const size_t nthreads = (size_t)ntiles * shell;
  hipLaunchKernelGGL(k_paint_faces, dim3((unsigned)((nthreads + 255) / 256), 1, S), dim3(256), 0, s, g, p->ps_faces,
                     face_stride, fcomp, mesh, mstride, mcomp, ncomp, shell, p->slots);
  return pmx_check_launch("paint_faces");
}

// Paint into `mesh` (accumulate=0 overwrites it). Default: the sorted owner-computes paint, both
// modes; `atomic`: one-pass scatter with global atomics (after a separate pre-drift pass).
// a.v != null applies the pre-drift x += v * cd_pre first (written back to x).
int pmx_launch_paint(pmx_plan_t* p, const PartArgs& a_in, const float* w, float w0, float cd_pre, float* mesh,
                     int accumulate, int atomic, hipStream_t s) {
  int rx, ry, rz;
  PartArgs a = a_in;
  ratios(p, a, rx, ry, rz);
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  if (atomic) {
    if (p->nslots > 1) return pmx_set_error(2, "a batched plan paints with the sorted deposit only");
    if (a.v != nullptr) {
      hipLaunchKernelGGL(k_drift, dim3(grid_particles(3 * np)), dim3(256), 0, s, a.x, a.v, (unsigned)(3 * np),
                         cd_pre);
      int rc = pmx_check_launch("drift");
      if (rc) return rc;
      a.v = nullptr;
    }
    if (!accumulate) {
      int rc = pmx_launch_fill(mesh, p->ncell, 0.f, s);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_paint, dim3(grid_particles(np)), dim3(256), 0, s, a, w, w0, mesh, p->nx, p->ny, p->nz, rx, ry,
                       rz);
    return pmx_check_launch("paint");
  }
  int rc = pmx_paint_sort(p, a, cd_pre, w, 1, w0, s);
  if (rc) return rc;
  return pmx_paint_deposit(p, a, w, 1, 1.0f, w0, w != nullptr || w0 < 0.f, mesh, 1, accumulate, s);
}

int pmx_launch_readout(pmx_plan_t* p, const PartArgs& a, const float* mesh, int ncomp, float* out, float ck,
                       float cd_post, int what, hipStream_t s) {
  int rx, ry, rz;
  ratios(p, a, rx, ry, rz);
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  // LDS staging pays when a block's corner box is small: whole blocks, and a lattice spacing of one
  // cell (at two cells apart an 4x8x8 block alone spans 8x16x16 cells)
  const bool staged = a.npx % RB_I == 0 && a.npy % RB_J == 0 && a.npz % RB_K == 0 &&
                      (RB_I * rx + 2) * (RB_J * ry + 2) * (RB_K * rz + 2) * 2 <= RO_CAP;
  if (staged) {
    const int nbj = a.npy / RB_J, nbk = a.npz / RB_K;
    const FDiv dj = make_fdiv((unsigned)nbj), dk = make_fdiv((unsigned)nbk);
    const unsigned G = (unsigned)(np / (RB_I * RB_J * RB_K));
    const unsigned S = (unsigned)p->nslots;
    if (ncomp == 3)
      hipLaunchKernelGGL(k_readout_lds<3>, dim3(G, 1, S), dim3(256), 0, s, a, mesh, out, ck, cd_post, what, p->nx,
                         p->ny, p->nz, rx, ry, rz, nbj, nbk, dj, dk, p->slots);
    else
      hipLaunchKernelGGL(k_readout_lds<1>, dim3(G, 1, S), dim3(256), 0, s, a, mesh, out, ck, cd_post, what & 1,
                         p->nx, p->ny, p->nz, rx, ry, rz, nbj, nbk, dj, dk, p->slots);
    return pmx_check_launch("readout");
  }
  if (p->nslots > 1) return pmx_set_error(2, "a batched readout needs the staged (LDS) form");
  if (ncomp == 3)
    hipLaunchKernelGGL(k_readout<3>, dim3(grid_particles(np)), dim3(256), 0, s, a, mesh, out, ck, cd_post, what, p->nx,
                       p->ny, p->nz, rx, ry, rz);
  else
    hipLaunchKernelGGL(k_readout<1>, dim3(grid_particles(np)), dim3(256), 0, s, a, mesh, out, ck, cd_post, what & 1,
                       p->nx, p->ny, p->nz, rx, ry, rz);
  return pmx_check_launch("readout");
}

int pmx_launch_finalize_map(const float* acc, float* map, long n, double scale, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_map, dim3(grid_for(n, 256)), dim3(256), 0, s, acc, map, n, scale, accumulate);
  return pmx_check_launch("finalize_map");
}

int pmx_launch_flat_vjp(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, int fnx, int fny, float zlo,
                        float zhi, const float* map_bar, float scale, float* pos_bar, int accumulate, hipStream_t s) {
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
  hipLaunchKernelGGL(k_paint_flat_vjp, dim3(grid_particles(np)), dim3(256), 0, s, a, w, w0, fa, map_bar, scale, pos_bar,
                     accumulate, rx, ry, rz);
  return pmx_check_launch("paint_flat_vjp");
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

int pmx_launch_lpt1(pmx_plan_t* p, const float* F, float* dx, float* pv, int npx, int npy, int npz, float D, float c,
                    hipStream_t s) {
  const size_t n = (size_t)npx * npy * npz;
  hipLaunchKernelGGL(k_lpt1, dim3(grid_for(n, 256)), dim3(256), 0, s, F, dx, pv, npx, npy, npz, p->nx / npx,
                     p->ny / npy, p->nz / npz, p->ny, p->nz, D, c);
  return pmx_check_launch("lpt1");
}
int pmx_launch_delta2(const float* H, float* rho, size_t ncell, int offdiag, hipStream_t s) {
  hipLaunchKernelGGL(k_delta2, dim3(grid_for(ncell, 256)), dim3(256), 0, s, H, rho, ncell, offdiag);
  return pmx_check_launch("delta2");
}
int pmx_launch_lpt_apply(pmx_plan_t* p, const float* F, float* dx, float* pv, int npx, int npy, int npz,
                         const LptGeom& g, int term, int accumulate, hipStream_t s) {
  const size_t n = (size_t)npx * npy * npz;
  if (accumulate)
    hipLaunchKernelGGL(k_lpt_apply<true>, dim3(grid_for(n, 256)), dim3(256), 0, s, F, dx, pv, npx, npy, npz,
                       p->nx / npx, p->ny / npy, p->nz / npz, p->ny, p->nz, g, term);
  else
    hipLaunchKernelGGL(k_lpt_apply<false>, dim3(grid_for(n, 256)), dim3(256), 0, s, F, dx, pv, npx, npy, npz,
                       p->nx / npx, p->ny / npy, p->nz / npz, p->ny, p->nz, g, term);
  return pmx_check_launch("lpt_apply");
}
int pmx_launch_lpt1_vjp(pmx_plan_t* p, const float* dxb, const float* pvb, int npx, int npy, int npz, float D, float c,
                        float* G, hipStream_t s) {
  hipLaunchKernelGGL(k_lpt1_vjp, dim3(grid_for(p->ncell, 256)), dim3(256), 0, s, dxb, pvb, G, p->nx, p->ny, p->nz,
                     p->nx / npx, p->ny / npy, p->nz / npz, npy, npz, D, c);
  return pmx_check_launch("lpt1_vjp");
}

// ------------------------------------------------------------------ adjoint launchers
int pmx_launch_absmax(const float* a, size_t n, unsigned* bits, hipStream_t s) {
  hipLaunchKernelGGL(k_absmax, dim3(grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024), dim3(256), 0, s, a, n, bits);
  return pmx_check_launch("absmax");
}
int pmx_launch_set_scale(unsigned* bits, float mul, float* scale, hipStream_t s) {
  hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(1), 0, s, bits, mul, scale);
  return pmx_check_launch("set_scale");
}
int pmx_launch_div4(pmx_plan_t* p, const float* U0, const float* U1, const float* U2, int ustride, float* out,
                    hipStream_t s) {
  const bool vec = ustride == 1 && p->nz % 4 == 0 && p->nz <= 1024 &&
                   ((reinterpret_cast<uintptr_t>(U0) | reinterpret_cast<uintptr_t>(U1) | reinterpret_cast<uintptr_t>(U2) |
                     reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
  if (vec) {
    const int tpr = p->nz / 4, yb = 256 / tpr;
    const int nyb = (p->ny + yb - 1) / yb;
    const int xb = p->nx < 4 ? p->nx : 4;
    const int nxb = (p->nx + xb - 1) / xb;
    hipLaunchKernelGGL(k_div4m, dim3((unsigned)(nxb * nyb)), dim3(tpr * yb), 0, s, U0, U1, U2, out, p->nx, p->ny,
                       p->nz, tpr, yb, xb, nyb);
    return pmx_check_launch("div4");
  }
  const unsigned rows = (unsigned)p->nx * p->ny;
  const int threads = p->nz < 256 ? p->nz : 256;
  hipLaunchKernelGGL(k_div4, dim3(rows), dim3(threads), 0, s, U0, U1, U2, ustride, out, p->nx, p->ny, p->nz);
  return pmx_check_launch("div4");
}
int pmx_launch_vjp_gather(pmx_plan_t* p, const PartArgs& a, int ncomp, const float* u, float umul, const float* M,
                          const float* rhob, const float* w, float w0, float* xbar, int accumulate, float* wbar,
                          double* dot, float* vkick, float ckick, hipStream_t s) {
  int rx, ry, rz;
  ratios(p, a, rx, ry, rz);
  const size_t np = (size_t)a.npx * a.npy * a.npz;
  double* part = dot ? p->red : nullptr;
  // the readout's staged form when its blocks tile the particle grid and a block's box is small
  const bool staged = a.npx % RB_I == 0 && a.npy % RB_J == 0 && a.npz % RB_K == 0 &&
                      (RB_I * rx + 2) * (RB_J * ry + 2) * (RB_K * rz + 2) * 2 <= RO_CAP;
  const unsigned nblocks = (unsigned)(np / (RB_I * RB_J * RB_K));
  const int nbj = a.npy / RB_J, nbk = a.npz / RB_K;
  const FDiv dj = make_fdiv((unsigned)(nbj > 0 ? nbj : 1)), dk = make_fdiv((unsigned)(nbk > 0 ? nbk : 1));
  const dim3 G(staged ? (nblocks < (unsigned)kRedMax ? nblocks : (unsigned)kRedMax) : (unsigned)grid_particles(np)),
      B(256);
#define PMX_G(NC, HU, HR, WB)                                                                                         \
  do {                                                                                                                \
    if (staged)                                                                                                       \
      hipLaunchKernelGGL((k_vjp_gather_lds<NC, HU, HR, WB>), G, B, 0, s, a, u, umul, M, rhob, w, w0, xbar,          \
                         accumulate, wbar, part, vkick, ckick, p->nx, p->ny, p->nz, rx, ry, rz, nbj, nbk, nblocks, dj, \
                         dk);                                                                                         \
    else                                                                                                              \
      hipLaunchKernelGGL((k_vjp_gather<NC, HU, HR, WB>), G, B, 0, s, a, u, umul, M, rhob, w, w0, xbar, accumulate,  \
                         wbar, part, vkick, ckick, p->nx, p->ny, p->nz, rx, ry, rz);                                  \
  } while (0)
  if (u && rhob) {
    if (ncomp == 3) PMX_G(3, true, true, false);
    else PMX_G(1, true, true, false);
  } else if (u) {
    if (ncomp == 3) PMX_G(3, true, false, false);
    else PMX_G(1, true, false, false);
  } else if (wbar) {
    PMX_G(1, false, true, true);
  } else {
    PMX_G(1, false, true, false);
  }
#undef PMX_G
  int rc = pmx_check_launch("vjp_gather");
  if (rc || !dot || !u) return rc;
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(256), 0, s, p->red, (int)G.x, dot);
  return pmx_check_launch("finish_sum");
}
int pmx_launch_axpy(float* out, const float* x, const float* v, float c, size_t n, hipStream_t s) {
  const size_t g = grid_for(n, 256);
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)(g < 8192 ? g : 8192)), dim3(256), 0, s, out, x, v, n, c);
  return pmx_check_launch("axpy");
}
int pmx_launch_adj_prep(pmx_plan_t* p, const float* xbar, float* vbar, const float* v, size_t n, float cd, double* dot,
                        unsigned* bits, hipStream_t s) {
  const int G = grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024;
  hipLaunchKernelGGL(k_adj_prep, dim3(G), dim3(256), 0, s, xbar, vbar, v, n, cd, p->red, bits);
  int rc = pmx_check_launch("adj_prep");
  if (rc) return rc;
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(256), 0, s, p->red, G, dot);
  return pmx_check_launch("finish_sum");
}
