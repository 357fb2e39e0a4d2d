// Hand-written 3D real FFT for gfx950: LDS-resident Stockham radix-8/4/3/2 line transforms,
// arranged as five bandwidth-bound passes whose stores transpose through LDS so every load and
// store of every pass is a contiguous run.
//
// Forward (r2c) chain, per mesh R[x][y][z] (nx, ny, nz), nzh = nz/2 + 1:
//   P1  r2c along z (nz/2-point complex FFT + split)   R[x][y][z]   -> A[kz][x][y]
//   P2  c2c along y                                    A[kz][x][y]  -> B[kz][ky][x]
//   P3  c2c along x  + k-space operator + inverse x     B[kz][ky][x] -> C[c][kz][x][ky]
// Inverse chain:
//   P4  inverse c2c along y                            C[c][kz][x][ky] -> D[c][x][y][kz]
//   P5  c2r along z (pre-merge + nz/2-point FFT)       D[c][x][y][kz]  -> out[x][y][z][c]
// P3 holds the whole spectral line along x for one (ky, kz): the Green's function, the gradient and
// the optional CIC compensation are applied there and the three force spectra are transformed
// back along x before they ever leave LDS. Transforms are unnormalised; the 1/N of the inverse is
// folded into the P3 operator.
#include "pmx_internal.h"
#include "pmx_device.h"

namespace {

// ---------------------------------------------------------------- radix plans
template <int N> struct FP;
#define PMX_FP(NN, SS, R0, R1, R2, R3)                        \
  template <> struct FP<NN> {                                 \
    static constexpr int S = SS;                              \
    static constexpr int R[4] = {R0, R1, R2, R3};             \
  };
PMX_FP(4, 1, 4, 0, 0, 0)
PMX_FP(8, 1, 8, 0, 0, 0)
PMX_FP(16, 2, 4, 4, 0, 0)
PMX_FP(32, 2, 8, 4, 0, 0)
PMX_FP(48, 3, 4, 4, 3, 0)
PMX_FP(64, 2, 8, 8, 0, 0)
PMX_FP(96, 3, 8, 4, 3, 0)
PMX_FP(128, 3, 8, 4, 4, 0)
PMX_FP(192, 3, 8, 8, 3, 0)
PMX_FP(256, 3, 8, 8, 4, 0)
PMX_FP(512, 3, 8, 8, 8, 0)
PMX_FP(1024, 4, 8, 8, 4, 4)
#undef PMX_FP

// LDS layout of one line of N complex values: element e sits at slot px<N>(e), and the lines of a
// workgroup are LSP<N> slots apart. Both come from a model of gfx950's LDS -- 64 banks of 4 bytes, a
// wave's 8-byte accesses served 32 lanes at a time -- which reproduces the bank-conflict counter of
// the FFT passes, applied to every Stockham stage's strided reads and writes and to the line loads and
// transposed stores: for 256, 512 and 1024 points an XOR of the low five bits with bits 3..7 makes the
// stages conflict-free (half the LDS cycles of a skipped slot every 16 values, e + e/16), and the other
// XORs below cut 15-50% for the shorter lines; 8 and 128 points keep e + e/16. Every map is a
// permutation of the line's slots (the XOR forms) or injective (e + e/16) within LSP<N>.
template <int N> __device__ __forceinline__ int px(int e) {
  if constexpr (N == 256 || N == 512 || N == 1024) return e ^ ((e >> 3) & 31);
  else if constexpr (N == 96 || N == 192) return e ^ ((e >> 5) & 3);
  else if constexpr (N == 64) return e ^ (((e >> 3) & 7) << 2);
  else if constexpr (N == 32) return e ^ ((e >> 3) & 3);
  else if constexpr (N == 16) return e ^ ((e >> 2) & 3);
  else return e + (e >> 4);
}
template <int N> constexpr int LSP() {
  switch (N) {
    case 4: return 5;
    case 8: return 9;
    case 16: return 20;
    case 32: return 36;
    case 48: return 59;
    case 64: return 69;
    case 96: return 108;
    case 128: return 137;
    case 192: return 216;
    case 256: return 260;
    case 512: return 520;
    case 1024: return 1040;
    default: return N + (N >> 4) + 1;
  }
}

// threads per line and lines per workgroup
template <int N> constexpr int TPL() { return N / FP<N>::R[0]; }
constexpr int pow2_floor(int v) { int p = 1; while (p * 2 <= v) p *= 2; return p; }
template <int N> constexpr int LPW() {
  int l = pow2_floor(256 / TPL<N>());
  return l < 1 ? 1 : (l > 64 ? 64 : l);
}

// Workgroups are dispatched to the 8 XCDs round-robin by linear id, i.e. by blockIdx.x when
// gridDim.x is a multiple of 8. Remapping x so that each XCD takes one contiguous range of line
// groups puts the workgroups whose transposed stores share cache lines on one L2, which merges them.
__device__ __forceinline__ unsigned xcd_x() {
  return (gridDim.x % 8u == 0u) ? xcd_logical(blockIdx.x, gridDim.x) : blockIdx.x;
}

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------- butterflies
template <int R, bool INV> struct Bfly;
template <bool INV> struct Bfly<2, INV> {
  __device__ static __forceinline__ void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <bool INV> struct Bfly<3, INV> {
  __device__ static __forceinline__ void run(float2* v) {
    const float s = 0.86602540378443864676f;
    float2 t = cadd(v[1], v[2]);
    float2 d = mul_mi<INV>(csub(v[1], v[2]));
    float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    v[0] = cadd(v[0], t);
    v[1] = make_float2(m.x + s * d.x, m.y + s * d.y);
    v[2] = make_float2(m.x - s * d.x, m.y - s * d.y);
  }
};
template <bool INV> struct Bfly<4, INV> {
  __device__ static __forceinline__ void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};
template <bool INV> struct Bfly<8, INV> {
  __device__ static __forceinline__ void run(float2* v) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<4, INV>::run(e);
    Bfly<4, INV>::run(o);
    // o[k] *= w^k, w = exp(-+2 pi i / 8)
    float2 o1 = INV ? make_float2((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h)
                    : make_float2((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    float2 o2 = mul_mi<INV>(o[2]);
    float2 o3 = INV ? make_float2((-o[3].x - o[3].y) * h, (o[3].x - o[3].y) * h)
                    : make_float2((o[3].y - o[3].x) * h, (-o[3].x - o[3].y) * h);
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// ---------------------------------------------------------------- Stockham stage
// One radix-R stage of an N-point transform held in LDS, NS = product of the radices already
// applied. Butterfly j reads a[j + r N/R], twiddles by w^(r (j mod NS) N/(NS R)), writes
// a[(j/NS) NS R + j mod NS + r NS] (self-sorting: natural order out after the last stage). The stage
// runs in place in one line buffer: each thread reads all of its butterflies' inputs into registers,
// the workgroup synchronises, and the outputs are written over the inputs -- half the LDS of a
// two-buffer (ping-pong) form, so twice the workgroups per CU, for one more barrier per stage. Every
// thread of the workgroup must call; the result is in `a`, visible to the whole workgroup on return.
template <int N, int R, int NS, bool INV>
__device__ __forceinline__ void stage_inplace(float2* a, int t, const float2* __restrict__ tw) {
  constexpr int T = TPL<N>();
  constexpr int NB = N / R, NJ = (NB + T - 1) / T;
  float2 v[NJ][R];
#pragma unroll
  for (int jj = 0; jj < NJ; ++jj) {
    const int j = t + jj * T;
    if ((NB % T == 0) || j < NB) {
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) v[jj][r] = a[px<N>(j + r * NB)];
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const float2 w = tw[r * k * (N / (NS * R))];
          v[jj][r] = INV ? cmulc(v[jj][r], w) : cmul(v[jj][r], w);
        }
      }
      Bfly<R, INV>::run(v[jj]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < NJ; ++jj) {
    const int j = t + jj * T;
    if ((NB % T == 0) || j < NB) {
      const int k = j % NS;
      const int d = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) a[px<N>(d + r * NS)] = v[jj][r];
    }
  }
  __syncthreads();
}
template <int N, bool INV, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_inplace(float2* a, int t, const float2* __restrict__ tw) {
  if constexpr (S < FP<N>::S) {
    constexpr int R = FP<N>::R[S];
    stage_inplace<N, R, NS, INV>(a, t, tw);
    fft_inplace<N, INV, S + 1, NS * R>(a, t, tw);
  }
}

// ---------------------------------------------------------------- P1: r2c along z
template <int NZ>
__global__ void __launch_bounds__(256) k_r2c_z(const float* __restrict__ in, float2* __restrict__ outA,
                                               int NX, int NY, const float2* __restrict__ twM,
                                               const float2* __restrict__ twN, Slots sl) {
  constexpr int M = NZ / 2, T = TPL<M>(), L = LPW<M>(), LS = LSP<M>(), NT = L * T;
  __shared__ float2 s0[L * LS];
  const unsigned zs = blockIdx.z;
  if (!slot_live(sl, zs)) return;
  in = at_slot(in, sl.stride, zs);
  outA = at_slot(outA, sl.stride, zs);
  __shared__ float2 stw[M];  // the line twiddles, read from LDS by every stage
  const int x = blockIdx.y, y0 = (int)xcd_x() * L, tid = threadIdx.x;
  const int nl = min(L, NY - y0);
  const float2* src = reinterpret_cast<const float2*>(in + ((size_t)x * NY + y0) * NZ);
  for (int i = tid; i < M; i += NT) stw[i] = twM[i];
  for (int i = tid; i < nl * M; i += NT) {
    const int l = i / M, k = i - l * M;
    s0[l * LS + px<M>(k)] = src[i];
  }
  __syncthreads();
  const int l = tid / T, t = tid - l * T;
  fft_inplace<M, false>(s0 + l * LS, t, stw);
  // split: X_k = E_k + W^k O_k,  E = (Z_k + conj Z_{M-k})/2,  O = -i (Z_k - conj Z_{M-k})/2, in place:
  // one thread takes the pair (p, M - p) (Z_M is Z_0), reads both inputs and writes both outputs
  constexpr int NP = M / 2 + 1;
  for (int i = tid; i < L * NP; i += NT) {
    const int ll = i / NP, p = i - ll * NP, q = M - p;
    float2* line = s0 + ll * LS;
    const float2 a = line[px<M>(p)], b = line[px<M>(q == M ? 0 : q)];
    const float2 ea = cscale(cadd(a, cconj(b)), 0.5f), oa = cscale(mul_mi<false>(csub(a, cconj(b))), 0.5f);
    const float2 eb = cscale(cadd(b, cconj(a)), 0.5f), ob = cscale(mul_mi<false>(csub(b, cconj(a))), 0.5f);
    const float2 xp = cadd(ea, cmul(twN[p], oa)), xq = cadd(eb, cmul(twN[q], ob));
    line[px<M>(p)] = xp;
    if (q != p) line[px<M>(q)] = xq;
  }
  __syncthreads();
  float2* dst = outA + (size_t)x * NY + y0;
  const size_t osk = (size_t)NX * NY;
  for (int i = tid; i < (M + 1) * L; i += NT) {
    const int k = i / L, ll = i - k * L;
    if (ll < nl) dst[k * osk + ll] = s0[ll * LS + px<M>(k)];
  }
}

// ---------------------------------------------------------------- P2 / P4: c2c lines
// line (by, bz, l) starts at in + by*isy + bz*isz + l*isl (elements contiguous); element k of line l
// is stored at out + by*osy + bz*osz + k*osk + l.
template <int N, bool INV>
__global__ void __launch_bounds__(256) k_c2c_lines(const float2* __restrict__ in, float2* __restrict__ out,
                                                   long isy, long isl, long isz, long osy, long osk, long osz,
                                                   int nlines, const float2* __restrict__ tw, int nzb, Slots sl) {
  constexpr int T = TPL<N>(), L = LPW<N>(), LS = LSP<N>(), NT = L * T;
  __shared__ float2 s0[L * LS];
  const unsigned zs = blockIdx.z / (unsigned)nzb;  // grid z: slot * nzb + the pass's own z block
  const int by = blockIdx.y, bz = (int)(blockIdx.z - zs * (unsigned)nzb), l0 = (int)xcd_x() * L, tid = threadIdx.x;
  if (l0 >= nlines) return;  // a padded grid's extra workgroups exit whole
  if (!slot_live(sl, zs)) return;
  in = at_slot(in, sl.stride, zs);
  out = at_slot(out, sl.stride, zs);
  __shared__ float2 stw[N];
  const int nl = min(L, nlines - l0);
  const float2* src = in + by * isy + bz * isz + (long)l0 * isl;
  for (int i = tid; i < N; i += NT) stw[i] = tw[i];
  for (int i = tid; i < nl * N; i += NT) {
    const int l = i / N, k = i - l * N;
    s0[l * LS + px<N>(k)] = src[l * isl + k];
  }
  __syncthreads();
  const int l = tid / T, t = tid - l * T;
  fft_inplace<N, INV>(s0 + l * LS, t, stw);
  const float2* res = s0;
  float2* dst = out + by * osy + bz * osz + l0;
  for (int i = tid; i < N * L; i += NT) {
    const int k = i / L, ll = i - k * L;
    if (ll < nl) dst[k * osk + ll] = res[ll * LS + px<N>(k)];
  }
}

// ---------------------------------------------------------------- P3: x lines + k-space operator

struct XTabs {
  const float *gx, *gy, *gz, *w2x, *w2y, *w2z, *dx, *dy, *dz;
};

__device__ __forceinline__ double interp_lin(double x, const double* __restrict__ xp, const double* __restrict__ fp,
                                             int n) {
  // ind = clip(searchsorted_left(xp, x) - 1, 0, n - 2); linear, extrapolating at both ends
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (xp[mid] < x) lo = mid + 1; else hi = mid;
  }
  int ind = lo - 1;
  ind = ind < 0 ? 0 : (ind > n - 2 ? n - 2 : ind);
  const double xi = xp[ind];
  return fp[ind] + (x - xi) * (fp[ind + 1] - fp[ind]) / (xp[ind + 1] - xi);
}

// Legendre polynomial P_l(mu) by the three-term recurrence n P_n = (2n-1) mu P_{n-1} - (n-1) P_{n-2}
__device__ __forceinline__ double legendre(double mu, int ell) {
  if (ell == 0) return 1.0;
  double p0 = 1.0, p1 = mu;
  for (int n = 2; n <= ell; ++n) {
    const double pn = ((2 * n - 1) * mu * p1 - (n - 1) * p0) / n;
    p0 = p1;
    p1 = pn;
  }
  return p1;
}

template <int N, int MODE>
__global__ void __launch_bounds__(256) k_xpass(const float2* __restrict__ in, float2* __restrict__ out, int NY,
                                               int NZ, long ocomp, KOp op, XTabs tb,
                                               const float2* __restrict__ tw, Slots sl) {
  constexpr int T = TPL<N>(), L = LPW<N>(), LS = LSP<N>(), NT = L * T;
  constexpr int NBUF = (MODE == XP_FORCES || MODE == XP_HESS) ? 2 : 1;
  __shared__ float2 sm[NBUF * L * LS];
  const unsigned zs = blockIdx.z;
  if (!slot_live(sl, zs)) return;
  in = at_slot(in, sl.stride, zs);
  out = at_slot(out, sl.stride, zs);
  if (sl.coef) op.scale = sl.coef[4 * zs + 3];
  __shared__ float2 stw[N];
  float2* s0 = sm;
  const int kz = blockIdx.y, ky0 = (int)xcd_x() * L, tid = threadIdx.x;
  const int nl = min(L, NY - ky0);
  const float2* src = in + ((size_t)kz * NY + ky0) * N;
  // XP_POT: the thread's R = L N / NT spectrum coefficients, from the tables while the lines load
  constexpr int R = (L * N) / NT;
  float pc[MODE == XP_POT ? R : 1];
  if constexpr (MODE == XP_POT) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * NT, ll = i / N, kx = i - ll * N;
      const int ky = min(ky0 + ll, NY - 1);
      const float k2 = tb.w2x[kx] + tb.w2y[ky] + tb.w2z[kz];
      float coef = (k2 > 0.f) ? op.scale / k2 : 0.f;
      if (op.flags & 1u) coef *= tb.dx[kx] * tb.dy[ky] * tb.dz[kz];
      if ((op.flags & 2u) && k2 > 0.f) coef *= expf(-op.pgd_kl2 / k2) * expf(-(k2 * k2) / op.pgd_ks4);
      if ((op.flags & 4u) && k2 > 0.f)
        coef *= op.mix_a + op.mix_b * (expf(-op.pgd_kl2 / k2) * expf(-(k2 * k2) / op.pgd_ks4));
      pc[r] = coef;
    }
  }
  for (int i = tid; i < N; i += NT) stw[i] = tw[i];
  for (int i = tid; i < nl * N; i += NT) {
    const int l = i / N, k = i - l * N;
    s0[l * LS + px<N>(k)] = src[i];
  }
  __syncthreads();
  const int l = tid / T, t = tid - l * T;
  constexpr bool FWD = (MODE != XP_INV);
  fft_inplace<N, !FWD>(s0 + l * LS, t, stw);
  float2* res = s0;

  if constexpr (MODE == XP_FWD) {
    float2* dst = out + ((size_t)kz * NY + ky0) * N;
    for (int i = tid; i < nl * N; i += NT) {
      const int ll = i / N, k = i - ll * N;
      dst[i] = res[ll * LS + px<N>(k)];
    }
  } else if constexpr (MODE == XP_INV) {
    float2* dst = out + (size_t)kz * N * NY + ky0;
    for (int i = tid; i < N * L; i += NT) {
      const int k = i / L, ll = i - k * L;
      if (ll < nl) dst[(size_t)k * NY + ll] = res[ll * LS + px<N>(k)];
    }
  } else if constexpr (MODE == XP_FORCES) {
    float2* s2 = sm + L * LS;
    for (int c = 0; c < 3; ++c) {
      for (int i = tid; i < L * N; i += NT) {
        const int ll = i / N, kx = i - ll * N;
        const int ky = min(ky0 + ll, NY - 1);
        const float k2 = tb.w2x[kx] + tb.w2y[ky] + tb.w2z[kz];
        const float g = (c == 0) ? tb.gx[kx] : (c == 1 ? tb.gy[ky] : tb.gz[kz]);
        float coef = (k2 > 0.f) ? g * op.scale / k2 : 0.f;
        if (op.flags & 1u) coef *= tb.dx[kx] * tb.dy[ky] * tb.dz[kz];
        const float2 d = res[ll * LS + px<N>(kx)];
        s2[ll * LS + px<N>(kx)] = make_float2(-coef * d.y, coef * d.x);  // i * coef * d
      }
      __syncthreads();
      fft_inplace<N, true>(s2 + l * LS, t, stw);
      const float2* r2 = s2;
      float2* dst = out + c * ocomp + (size_t)kz * N * NY + ky0;
      for (int i = tid; i < N * L; i += NT) {
        const int k = i / L, ll = i - k * L;
        if (ll < nl) dst[(size_t)k * NY + ll] = r2[ll * LS + px<N>(k)];
      }
      __syncthreads();
    }
  } else if constexpr (MODE == XP_HESS) {
    // H_ij = ifft(g_i g_j delta / k^2): the FastPM gradient kernel twice on the potential -delta/k^2,
    // (i g_i)(i g_j)(-delta / k^2); real coefficients, 0 at k = 0 and on Nyquist planes (g = 0 there)
    float2* s2 = sm + L * LS;
    const bool off = (op.flags & kHessOffDiag) != 0u;
    for (int c = 0; c < 3; ++c) {
      const int ci = off ? (c == 2 ? 1 : 0) : c, cj = off ? (c == 0 ? 1 : 2) : c;
      for (int i = tid; i < L * N; i += NT) {
        const int ll = i / N, kx = i - ll * N;
        const int ky = min(ky0 + ll, NY - 1);
        const float k2 = tb.w2x[kx] + tb.w2y[ky] + tb.w2z[kz];
        const float gi = (ci == 0) ? tb.gx[kx] : (ci == 1 ? tb.gy[ky] : tb.gz[kz]);
        const float gj = (cj == 0) ? tb.gx[kx] : (cj == 1 ? tb.gy[ky] : tb.gz[kz]);
        const float coef = (k2 > 0.f) ? gi * gj * op.scale / k2 : 0.f;
        const float2 d = res[ll * LS + px<N>(kx)];
        s2[ll * LS + px<N>(kx)] = make_float2(coef * d.x, coef * d.y);
      }
      __syncthreads();
      fft_inplace<N, true>(s2 + l * LS, t, stw);
      const float2* r2 = s2;
      float2* dst = out + c * ocomp + (size_t)kz * N * NY + ky0;
      for (int i = tid; i < N * L; i += NT) {
        const int k = i / L, ll = i - k * L;
        if (ll < nl) dst[(size_t)k * NY + ll] = r2[ll * LS + px<N>(k)];
      }
      __syncthreads();
    }
  } else if constexpr (MODE == XP_POT) {
    // potential for the real-space FastPM gradient: phi_hat = delta_hat * scale / k^2 (0 at k = 0),
    // the coefficients computed at the start
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * NT, ll = i / N, kx = i - ll * N;
      res[ll * LS + px<N>(kx)] = cscale(res[ll * LS + px<N>(kx)], pc[r]);
    }
    __syncthreads();
    fft_inplace<N, true>(res + l * LS, t, stw);
    const float2* r2 = res;
    float2* dst = out + (size_t)kz * N * NY + ky0;
    for (int i = tid; i < N * L; i += NT) {
      const int k = i / L, ll = i - k * L;
      if (ll < nl) dst[(size_t)k * NY + ll] = r2[ll * LS + px<N>(k)];
    }
  } else if constexpr (MODE == XP_COLOR) {
    for (int i = tid; i < L * N; i += NT) {
      const int ll = i / N, kx = i - ll * N;
      const int ky = min(ky0 + ll, NY - 1);
      const double kk = sqrt(op.kphys_x[kx] * op.kphys_x[kx] + op.kphys_y[ky] * op.kphys_y[ky] +
                             op.kphys_z[kz] * op.kphys_z[kz]);
      double pk = interp_lin(kk, op.ktab, op.ptab, op.ntab) * op.cnorm;
      const float amp = (float)(sqrt(pk > 0.0 ? pk : 0.0)) * op.scale;
      res[ll * LS + px<N>(kx)] = cscale(res[ll * LS + px<N>(kx)], amp);
    }
    __syncthreads();
    fft_inplace<N, true>(res + l * LS, t, stw);
    const float2* r2 = res;
    float2* dst = out + (size_t)kz * N * NY + ky0;
    for (int i = tid; i < N * L; i += NT) {
      const int k = i / L, ll = i - k * L;
      if (ll < nl) dst[(size_t)k * NY + ll] = r2[ll * LS + px<N>(k)];
    }
  } else if constexpr (MODE == XP_POWER) {
    // bins: 0 .. nedges (nedges + 1 entries), bin = number of edges <= |k|. The half spectrum holds
    // one of each pair (k, k') of an interior kz plane, k' the index-mirror of k: the spectrum there is
    // the conjugate, |k'| = |k|, and k' is -k except on the x/y Nyquist lines, whose fftfreq component
    // keeps its sign; mu' is taken from k' itself. Planes kz = 0 and kz = nz/2 hold every mode once.
    extern __shared__ double hist[];  // [npoles re][npoles im][k sum][count] x (nedges + 1)
    const int nb = op.nedges + 1, np = op.npoles;
    for (int i = tid; i < (2 * np + 2) * nb; i += NT) hist[i] = 0.0;
    __syncthreads();
    const bool interior = !(kz == 0 || 2 * kz == NZ);
    const double mult = interior ? 2.0 : 1.0;
    const double kzp = op.kphys_z[kz];
    int lmax = 0;
    for (int j = 0; j < np; ++j) lmax = max(lmax, op.poles[j]);
    for (int i = tid; i < nl * N; i += NT) {
      const int ll = i / N, kx = i - ll * N;
      const int ky = ky0 + ll;
      const double kxp = op.kphys_x[kx], kyp = op.kphys_y[ky];
      double s = 0.0;
      s = __dadd_rn(s, __dmul_rn(kxp, kxp));
      s = __dadd_rn(s, __dmul_rn(kyp, kyp));
      s = __dadd_rn(s, __dmul_rn(kzp, kzp));
      const double kk = __dsqrt_rn(s);
      int lo = 0, hi = op.nedges;  // upper_bound
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (op.kedges[mid] <= kk) lo = mid + 1; else hi = mid;
      }
      const float2 d = res[ll * LS + px<N>(kx)];
      double re, im = 0.0;
      if (op.spec2) {  // m1 conj(m2)
        const float2 e = op.spec2[((size_t)kz * NY + ky) * N + kx];
        re = (double)d.x * e.x + (double)d.y * e.y;
        im = (double)d.y * e.x - (double)d.x * e.y;
      } else {
        re = (double)d.x * d.x + (double)d.y * d.y;
      }
      if (lmax == 0) {
        atomicAdd(&hist[lo], re * mult);
        if (op.spec2 && !interior) atomicAdd(&hist[np * nb + lo], im);
      } else {
        const double inv = (kk == 0.0) ? 0.0 : 1.0 / kk;
        const double mu = (kk == 0.0) ? 0.0 : (kxp * op.los[0] + kyp * op.los[1] + kzp * op.los[2]) * inv;
        double mu2 = 0.0;
        if (interior) {
          const double kxq = op.kphys_x[kx ? N - kx : 0], kyq = op.kphys_y[ky ? NY - ky : 0];
          mu2 = (kk == 0.0) ? 0.0 : (kxq * op.los[0] + kyq * op.los[1] - kzp * op.los[2]) * inv;
        }
        for (int j = 0; j < np; ++j) {
          const int ell = op.poles[j];
          double w = legendre(mu, ell) * (2 * ell + 1);
          double wr = w, wi = w;
          if (interior) {
            const double w2 = legendre(mu2, ell) * (2 * ell + 1);
            wr = w + w2;
            wi = w - w2;
          }
          atomicAdd(&hist[j * nb + lo], re * wr);
          if (op.spec2) atomicAdd(&hist[(np + j) * nb + lo], im * wi);
        }
      }
      if (op.acc_k) {
        atomicAdd(&hist[2 * np * nb + lo], kk * mult);
        atomicAdd(&hist[(2 * np + 1) * nb + lo], mult);
      }
    }
    __syncthreads();
    for (int i = tid; i < np * nb; i += NT) {
      if (hist[i] != 0.0) atomicAdd(&op.pk_sum[i], hist[i]);
      if (op.spec2 && hist[np * nb + i] != 0.0) atomicAdd(&op.pk_im[i], hist[np * nb + i]);
    }
    if (op.acc_k) {
      for (int i = tid; i < nb; i += NT) {
        if (hist[2 * np * nb + i] != 0.0) atomicAdd(&op.k_sum[i], hist[2 * np * nb + i]);
        if (hist[(2 * np + 1) * nb + i] != 0.0) atomicAdd(&op.k_cnt[i], hist[(2 * np + 1) * nb + i]);
      }
    }
  }
}

// ---------------------------------------------------------------- P5: c2r along z
// lines (c, l): D[c][x][y0+l][0..nzh) (rows of NZHP complex) -> out[x][y0+l][z][c] (NC components
// interleaved)
template <int NZ, int NC>
__global__ void __launch_bounds__(384) k_c2r_z(const float2* __restrict__ inD, float* __restrict__ out, int NY,
                                               long icomp, const float2* __restrict__ twM,
                                               const float2* __restrict__ twN, Slots sl) {
  constexpr int M = NZ / 2, T = TPL<M>();
  constexpr int L = (NC == 1) ? LPW<M>() : (LPW<M>() / 2 > 0 ? LPW<M>() / 2 : 1);
  constexpr int LS = LSP<M>(), NL = NC * L, NT = NL * T, NZH = M + 1, NZHP = (NZH + 7) & ~7;
  __shared__ float2 s0[NL * LS];
  const unsigned zs = blockIdx.z;
  if (!slot_live(sl, zs)) return;
  inD = at_slot(inD, sl.stride, zs);
  out = at_slot(out, sl.stride, zs);
  __shared__ float2 stw[M];
  const int x = blockIdx.y, y0 = blockIdx.x * L, tid = threadIdx.x;
  const int nl = min(L, NY - y0);
  for (int i = tid; i < M; i += NT) stw[i] = twM[i];
  // load X[k], k = 0..M
  for (int i = tid; i < NC * nl * NZH; i += NT) {
    const int c = i / (nl * NZH), rem = i - c * (nl * NZH);
    const int ll = rem / NZH, k = rem - ll * NZH;
    s0[(c * L + ll) * LS + px<M>(k)] = inD[c * icomp + ((size_t)x * NY + y0 + ll) * NZHP + k];
  }
  __syncthreads();
  // merge: Z'_k = (X_k + conj X_{M-k}) + i (X_k - conj X_{M-k}) conj(W^k), in place: one thread takes
  // the pair (p, M - p), reads X_p and X_{M-p} and writes Z'_p and Z'_{M-p} (X_M only feeds Z'_0)
  constexpr int NP = M / 2 + 1;
  for (int i = tid; i < NL * NP; i += NT) {
    const int ln = i / NP, p = i - ln * NP, q = M - p;
    float2* line = s0 + ln * LS;
    const float2 xp = line[px<M>(p)], xq = line[px<M>(q)];
    const float2 ap = cadd(xp, cconj(xq)), bp = cmulc(csub(xp, cconj(xq)), twN[p]);
    line[px<M>(p)] = make_float2(ap.x - bp.y, ap.y + bp.x);
    if (p != 0 && q != p) {
      const float2 aq = cadd(xq, cconj(xp)), bq = cmulc(csub(xq, cconj(xp)), twN[q]);
      line[px<M>(q)] = make_float2(aq.x - bq.y, aq.y + bq.x);
    }
  }
  __syncthreads();
  const int ln = tid / T, t = tid - ln * T;
  fft_inplace<M, true>(s0 + ln * LS, t, stw);
  const float2* res = s0;
  float* dst = out + ((size_t)x * NY + y0) * NZ * NC;
  for (int i = tid; i < nl * NZ * NC; i += NT) {
    const int ll = i / (NZ * NC), rem = i - ll * (NZ * NC);
    const int z = rem / NC, c = rem - z * NC;
    const float2 v = res[(c * L + ll) * LS + px<M>((z >> 1))];
    dst[i] = (z & 1) ? v.y : v.x;
  }
}

// ---------------------------------------------------------------- force mesh from the potential
// F_d = 8 (phi[+e_d] - phi[-e_d]) - (phi[+2 e_d] - phi[-2 e_d]) per cell (periodic), written interleaved
// F[x][y][z][3]: the real-space form of the FastPM gradient i (8 sin k - sin 2k) / 6 (the 1/12 is
// folded into the potential). A workgroup takes yb consecutive rows (x, y0 .. y0 + yb - 1)
// of one y block (tpr = NZ / 4 threads per row, 4 cells each) and walks xb consecutive x planes.
// Each thread keeps phi[x - 2 .. x + 2] of its 4 cells in registers (one new 16-byte load per
// plane); the plane's rows y0 - 2 .. y0 + yb + 1 sit in LDS for the y and z neighbours (the own
// rows from registers, the 4 halo rows loaded by the workgroup). About two loads per output row
// instead of nine (72 -> 60 us at 256^3 against one workgroup per four rows of one plane). Outputs are
// staged through LDS and stored as contiguous 16-byte lanes.
constexpr int GRAD4_LDS_FLOATS = 5120;  // (yb + 4) * NZ <= 5120 for every supported NZ (yb = 256 / tpr)
__global__ void __launch_bounds__(256) k_grad4(const float* __restrict__ phi, float* __restrict__ F, int NX, int NY,
                                                int NZ, int tpr, int yb, int xb, int nyb, Slots sl) {
  // (yb + 4) * NZ floats, sized at launch (8 KB at NZ = 256: seven workgroups per CU, as the 68 VGPRs
  // allow, instead of the five a fixed 20 KB array left room for)
  extern __shared__ __align__(16) float4 plane4[];
  __shared__ float4 stage[3 * 256];
  float* plane = reinterpret_cast<float*>(plane4);
  const unsigned zs = blockIdx.z;
  if (!slot_live(sl, zs)) return;
  phi = at_slot(phi, sl.stride, zs);
  F = at_slot(F, sl.stride, zs);
  const unsigned Lb = xcd_logical(blockIdx.x, gridDim.x);
  const int by = (int)(Lb % (unsigned)nyb), bx = (int)(Lb / (unsigned)nyb);
  const int me = (int)threadIdx.x, nt = (int)blockDim.x;  // nt = yb * tpr
  const int lr = me / tpr, t = me - lr * tpr, z0 = 4 * t;
  const int y0 = by * yb, x0 = bx * xb;
  const int vrows = min(yb, NY - y0);  // rows of this block inside the mesh
  const bool valid = lr < vrows;
  const int y = valid ? y0 + lr : y0;  // past the edge: a real row, loaded and never stored
  const int xend = min(x0 + xb, NX);
  auto wr = [](int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
  auto ld = [&](int xx, int yy) { return *reinterpret_cast<const float4*>(phi + ((size_t)xx * NY + yy) * NZ + z0); };
  float4 w0 = ld(wr(x0 - 2, NX), y), w1 = ld(wr(x0 - 1, NX), y), w2 = ld(x0, y), w3 = ld(wr(x0 + 1, NX), y);
  const int hq = 4 * tpr;  // halo: 4 rows of tpr float4
  for (int x = x0; x < xend; ++x) {
    const float4 w4 = ld(wr(x + 2, NX), y);
    // the plane's rows in LDS at index (row - y0 + 2): the vrows own rows from w2, halo rows y0 - 2,
    // y0 - 1 and y0 + vrows, y0 + vrows + 1 (periodic) from memory, so LDS rows 0 .. vrows + 3 hold
    // consecutive rows of the plane
    if (valid) plane4[((lr + 2) * NZ + z0) / 4] = w2;
    for (int q = me; q < hq; q += nt) {
      const int h = q / tpr, c = q - h * tpr;
      const int ly = h < 2 ? h : vrows + h;  // LDS rows 0, 1, vrows + 2, vrows + 3
      const int yy = wr(y0 - 2 + ly, NY);    // y0 - 2 .. y0 + vrows + 1, within one period of [0, NY)
      plane4[(ly * NZ) / 4 + c] = *reinterpret_cast<const float4*>(phi + ((size_t)x * NY + yy) * NZ + 4 * c);
    }
    __syncthreads();
    const float* rw = plane + (lr + 2) * NZ;
    const float4 yp1 = plane4[((lr + 3) * NZ + z0) / 4], ym1 = plane4[((lr + 1) * NZ + z0) / 4];
    const float4 yp2 = plane4[((lr + 4) * NZ + z0) / 4], ym2 = plane4[(lr * NZ + z0) / 4];
    const float v[8] = {rw[z0 == 0 ? NZ - 2 : z0 - 2], rw[z0 == 0 ? NZ - 1 : z0 - 1], w2.x, w2.y, w2.z, w2.w,
                        rw[z0 + 4 == NZ ? 0 : z0 + 4], rw[z0 + 4 == NZ ? 1 : z0 + 5]};
    const float ax1[4] = {w3.x, w3.y, w3.z, w3.w}, bx1[4] = {w1.x, w1.y, w1.z, w1.w};
    const float ax2[4] = {w4.x, w4.y, w4.z, w4.w}, bx2[4] = {w0.x, w0.y, w0.z, w0.w};
    const float ay1[4] = {yp1.x, yp1.y, yp1.z, yp1.w}, by1[4] = {ym1.x, ym1.y, ym1.z, ym1.w};
    const float ay2[4] = {yp2.x, yp2.y, yp2.z, yp2.w}, by2[4] = {ym2.x, ym2.y, ym2.z, ym2.w};
    float o[12];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      o[3 * m] = 8.0f * (ax1[m] - bx1[m]) - (ax2[m] - bx2[m]);
      o[3 * m + 1] = 8.0f * (ay1[m] - by1[m]) - (ay2[m] - by2[m]);
      o[3 * m + 2] = 8.0f * (v[m + 3] - v[m + 1]) - (v[m + 4] - v[m]);
    }
    stage[3 * me] = make_float4(o[0], o[1], o[2], o[3]);
    stage[3 * me + 1] = make_float4(o[4], o[5], o[6], o[7]);
    stage[3 * me + 2] = make_float4(o[8], o[9], o[10], o[11]);
    __syncthreads();
    const int nq = 3 * vrows * tpr;
    float4* out = reinterpret_cast<float4*>(F + ((size_t)x * NY + y0) * NZ * 3);
    for (int q = me; q < nq; q += nt) out[q] = stage[q];
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = w4;
  }
}

// ---------------------------------------------------------------- host dispatch
#define PMX_FOR_AXIS(M_) \
  M_(8) M_(16) M_(32) M_(64) M_(96) M_(128) M_(192) M_(256) M_(512) M_(1024)

template <int NZ> int launch_r2c_z(pmx_plan_t* p, const float* in, float2* out, hipStream_t s) {
  constexpr int M = NZ / 2, L = LPW<M>(), T = TPL<M>();
  dim3 grid((p->ny + L - 1) / L, p->nx, p->nslots), block(L * T);
  hipLaunchKernelGGL(k_r2c_z<NZ>, grid, block, 0, s, in, out, p->nx, p->ny, p->tw_zh, p->tw_z, p->slots);
  return pmx_check_launch("r2c_z");
}

template <int N, bool INV>
int launch_c2c(const float2* in, float2* out, long isy, long isl, long isz, long osy, long osk, long osz,
               int nlines, int ny_blocks, int nz_blocks, const float2* tw, hipStream_t s, int nslots,
               const Slots& sl) {
  constexpr int L = LPW<N>(), T = TPL<N>();
  // runs of fewer than 8 lines store partial 64-byte sectors: pad the grid's x extent to a multiple
  // of 8 so the XCD remap gathers the workgroups sharing those sectors on one L2
  const int g0 = (nlines + L - 1) / L;
  const int gx = L < 8 ? (g0 + 7) / 8 * 8 : g0;
  dim3 grid(gx, ny_blocks, nz_blocks * nslots), block(L * T);
  hipLaunchKernelGGL((k_c2c_lines<N, INV>), grid, block, 0, s, in, out, isy, isl, isz, osy, osk, osz, nlines, tw,
                     nz_blocks, sl);
  return pmx_check_launch("c2c_lines");
}

template <int N, int MODE> int launch_xpass(pmx_plan_t* p, const KOp& op, const float2* in, float2* out, hipStream_t s) {
  constexpr int L = LPW<N>(), T = TPL<N>();
  dim3 grid((p->ny + L - 1) / L, p->nzh, p->nslots), block(L * T);
  XTabs tb{p->grad_x, p->grad_y, p->grad_z, p->w2_x, p->w2_y, p->w2_z, p->dec_x, p->dec_y, p->dec_z};
  size_t shm = (MODE == XP_POWER) ? (2 * (size_t)op.npoles + 2) * (op.nedges + 1) * sizeof(double) : 0;
  hipLaunchKernelGGL((k_xpass<N, MODE>), grid, block, shm, s, in, out, p->ny, p->nz, (long)p->nspec, op, tb, p->tw_x,
                     p->slots);
  return pmx_check_launch("xpass");
}

template <int NZ, int NC> int launch_c2r_z(pmx_plan_t* p, const float2* in, float* out, hipStream_t s) {
  constexpr int M = NZ / 2, T = TPL<M>();
  constexpr int L = (NC == 1) ? LPW<M>() : (LPW<M>() / 2 > 0 ? LPW<M>() / 2 : 1);
  dim3 grid((p->ny + L - 1) / L, p->nx, p->nslots), block(NC * L * T);
  hipLaunchKernelGGL((k_c2r_z<NZ, NC>), grid, block, 0, s, in, out, p->ny, (long)p->nspecd, p->tw_zh, p->tw_z,
                     p->slots);
  return pmx_check_launch("c2r_z");
}

}  // namespace

bool pmx_fft_len_ok(int n) {
  switch (n) {
#define M_(v) case v:
    PMX_FOR_AXIS(M_)
#undef M_
    return true;
    default:
      return false;
  }
}

int pmx_pass_r2c_z(pmx_plan_t* p, const float* in, float2* outA, hipStream_t s) {
  switch (p->nz) {
#define M_(v) case v: return launch_r2c_z<v>(p, in, outA, s);
    PMX_FOR_AXIS(M_)
#undef M_
  }
  return pmx_set_error(2, "unsupported nz");
}

int pmx_pass_c2c_y(pmx_plan_t* p, const float2* inA, float2* outB, hipStream_t s) {
  // lines: A[kz][x][*] (kz = block y, x = line), out B[kz][ky][x]
  const long NX = p->nx, NY = p->ny;
  switch (p->ny) {
#define M_(v)                                                                                                \
  case v:                                                                                                    \
    return launch_c2c<v, false>(inA, outB, NX * NY, NY, 0, NY * NX, NX, 0, p->nx, p->nzh, 1, p->tw_y, s, p->nslots, \
                                p->slots);
    PMX_FOR_AXIS(M_)
#undef M_
  }
  return pmx_set_error(2, "unsupported ny");
}

int pmx_fft_forward_to_B(pmx_plan_t* p, const float* in, hipStream_t s) {
  int rc = pmx_pass_r2c_z(p, in, p->specA, s);
  if (rc) return rc;
  return pmx_pass_c2c_y(p, p->specA, p->specB, s);
}

namespace {
template <int N> constexpr size_t xpass_static_lds() {
  return ((size_t)LPW<N>() * LSP<N>() + N) * sizeof(float2);  // one line buffer + the twiddles
}
}  // namespace

int pmx_xpass_power_poles(const pmx_plan_t* p, int nedges) {
  // per-workgroup LDS kept within 64 KiB: the x-line buffers plus the (2 npoles + 2) histograms
  size_t stat = 0;
  switch (p->nx) {
#define M_(v) case v: stat = xpass_static_lds<v>(); break;
    PMX_FOR_AXIS(M_)
#undef M_
    default: return 0;
  }
  const size_t budget = 65536 - stat, per = (size_t)(nedges + 1) * sizeof(double);
  const long np = (long)(budget / per) / 2 - 1;
  return np < 1 ? 0 : (np > kMaxPolesLaunch ? kMaxPolesLaunch : (int)np);
}

int pmx_xpass(pmx_plan_t* p, const KOp& op, const float2* inB, float2* out, hipStream_t s) {
  switch (p->nx) {
#define M_(v)                                                               \
  case v:                                                                   \
    switch (op.mode) {                                                      \
      case XP_FWD: return launch_xpass<v, XP_FWD>(p, op, inB, out, s);      \
      case XP_INV: return launch_xpass<v, XP_INV>(p, op, inB, out, s);      \
      case XP_FORCES: return launch_xpass<v, XP_FORCES>(p, op, inB, out, s); \
      case XP_COLOR: return launch_xpass<v, XP_COLOR>(p, op, inB, out, s);  \
      case XP_POWER: return launch_xpass<v, XP_POWER>(p, op, inB, out, s);  \
      case XP_POT: return launch_xpass<v, XP_POT>(p, op, inB, out, s);      \
      case XP_HESS: return launch_xpass<v, XP_HESS>(p, op, inB, out, s);    \
    }                                                                       \
    return pmx_set_error(2, "bad k-space mode");
    PMX_FOR_AXIS(M_)
#undef M_
  }
  return pmx_set_error(2, "unsupported nx");
}

int pmx_pass_c2c_y_inv(pmx_plan_t* p, int ncomp, const float2* inC, float2* outD, hipStream_t s) {
  // P4: lines C[c][kz][x][*] for consecutive kz at fixed x -> D[c][x][ky][kz], rows of nzhp complex so
  // a workgroup's runs of consecutive kz start on 64-byte boundaries
  const long NX = p->nx, NY = p->ny, NZH = p->nzhp;
  if (ncomp != 1 && ncomp != 3) return pmx_set_error(2, "ncomp must be 1 or 3");
  switch (p->ny) {
#define M_(v)                                                                                         \
  case v:                                                                                             \
    return launch_c2c<v, true>(inC, outD, NY, NX * NY, (long)p->nspec, NY * NZH, NZH, (long)p->nspecd, \
                               p->nzh, p->nx, ncomp, p->tw_y, s, p->nslots, p->slots);
    PMX_FOR_AXIS(M_)
#undef M_
  }
  return pmx_set_error(2, "unsupported ny");
}

int pmx_pass_c2r_z(pmx_plan_t* p, int ncomp, const float2* inD, float* out, hipStream_t s) {
  if (ncomp != 1 && ncomp != 3) return pmx_set_error(2, "ncomp must be 1 or 3");
  switch (p->nz) {
#define M_(v) \
  case v: return ncomp == 3 ? launch_c2r_z<v, 3>(p, inD, out, s) : launch_c2r_z<v, 1>(p, inD, out, s);
    PMX_FOR_AXIS(M_)
#undef M_
  }
  return pmx_set_error(2, "unsupported nz");
}

int pmx_grad4(pmx_plan_t* p, const float* phi, float* F, hipStream_t s) {
  // every supported axis length is a multiple of 4 and at most 1024 (tpr <= 256)
  const int tpr = p->nz / 4;
  if (p->nz % 4 != 0 || tpr > 256) return pmx_set_error(2, "grad4 needs nz a multiple of 4, at most 1024");
  if ((reinterpret_cast<uintptr_t>(phi) | reinterpret_cast<uintptr_t>(F)) & 15u)
    return pmx_set_error(2, "grad4 needs 16-byte aligned potential and force buffers");
  const int yb = 256 / tpr;                  // rows per workgroup: (yb + 4) * nz <= GRAD4_LDS_FLOATS
  const int nyb = (p->ny + yb - 1) / yb;
  const int xb = p->nx < 4 ? p->nx : 4;      // x planes each workgroup walks
  const int nxb = (p->nx + xb - 1) / xb;
  const size_t shm = (size_t)(yb + 4) * p->nz * sizeof(float);
  hipLaunchKernelGGL(k_grad4, dim3((unsigned)(nxb * nyb), 1, p->nslots), dim3(tpr * yb), shm, s, phi, F, p->nx, p->ny,
                     p->nz, tpr, yb, xb, nyb, p->slots);
  return pmx_check_launch("grad4");
}

int pmx_inverse_from_C(pmx_plan_t* p, int ncomp, const float2* inC, float* out, hipStream_t s) {
  int rc = pmx_pass_c2c_y_inv(p, ncomp, inC, p->specD, s);
  if (rc) return rc;
  return pmx_pass_c2r_z(p, ncomp, p->specD, out, s);
}
