// extern "C" entry points of libpmx: plans, argument checks, and the composition of the pass and
// particle kernels into the PM force and step pipelines.
#include "pmx_internal.h"
#include "../../include/pmx.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

// launchers from pmx_particles.hip
int pmx_launch_fill(float* a, size_t n, float v, hipStream_t s);
int pmx_launch_fill_d(double* a, size_t n, hipStream_t s);
int pmx_launch_finalize_map(const float* acc, float* map, long n, double scale, int accumulate, hipStream_t s);
int pmx_launch_flat_vjp(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, int fnx, int fny, float zlo,
                        float zhi, const float* map_bar, float scale, float* pos_bar, int accumulate, hipStream_t s);
int pmx_launch_flat(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, int fnx, int fny, float zlo, float zhi,
                    float* acc, hipStream_t s);
int pmx_launch_philox_raw(uint32_t* out, size_t nblocks, uint64_t seed, hipStream_t s);
int pmx_launch_white_noise(float* out, size_t n, uint64_t seed, hipStream_t s);
int pmx_launch_lpt1(pmx_plan_t* p, const float* F, float* dx, float* pv, int npx, int npy, int npz, float D, float c,
                    hipStream_t s);
// delta2 of 2LPT from three Hessian components (interleaved (x,y,z,3)): diagonal pass
// rho = H11 H00 + H22 (H00 + H11), off-diagonal pass rho -= H01^2 + H02^2 + H12^2
int pmx_launch_delta2(const float* H, float* rho, size_t ncell, int offdiag, hipStream_t s);
// dx (+)= D F, pv (+)= c D F at the lattice
int pmx_launch_lpt_apply(pmx_plan_t* p, const float* F, float* dx, float* pv, int npx, int npy, int npz,
                         const LptGeom& g, int term, int accumulate, hipStream_t s);
int pmx_launch_ud_grade(const float* in, int nside_in, float* out, int nside_out, double ratio, int ring_in,
                        int ring_out, int pess, hipStream_t s);
int pmx_launch_lpt1_vjp(pmx_plan_t* p, const float* dxb, const float* pvb, int npx, int npy, int npz, float D, float c,
                        float* G, hipStream_t s);
// launchers from pmx_sky.hip
int pmx_sky_max_tiles();
int pmx_born_max_shells();
int pmx_born_max_sources();
int pmx_launch_healpix(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, const double box[3],
                       const double obs[3], double rmin, double rmax, int nside, int scheme, int ntiles,
                       const int* rot, const double* shift, float* map, int accumulate, hipStream_t s,
                       const PmxDrift* drift = nullptr);
int pmx_launch_healpix_vjp(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, const double box[3],
                           const double obs[3], double rmin, double rmax, int nside, int scheme, int ntiles,
                           const int* rot, const double* shift, const float* map_bar, float* pos_bar, int accumulate,
                           hipStream_t s, const PmxDrift* drift = nullptr, float* vel_bar = nullptr);
int pmx_launch_image_positions(pmx_plan_t* p, const PartArgs& a, const double box[3], const double obs[3], int code,
                               const double shift[3], float* out, hipStream_t s, const PmxDrift* drift);
int pmx_launch_image_positions_vjp(pmx_plan_t* p, const PartArgs& a, const double box[3], const double obs[3],
                                   int code, const double shift[3], const float* gbar, float* xbar, float* vbar,
                                   int accumulate, hipStream_t s, const PmxDrift* drift);
int pmx_launch_born(pmx_plan_t* p, const float* maps, int S, long npix, const double* C, int K, float* kappa,
                    int accumulate, hipStream_t s);
int pmx_launch_born_vjp(pmx_plan_t* p, const float* maps, int S, long npix, const double* C, int K,
                        const float* kappa_bar, float* maps_bar, int accumulate, double* dots, hipStream_t s);
int pmx_launch_born_flat(pmx_plan_t* p, const float* maps, int S, int ny, int nx, const double* C, int K,
                         const double* sc, float* kappa, int accumulate, hipStream_t s);
int pmx_launch_born_flat_vjp(pmx_plan_t* p, const float* maps, int S, int ny, int nx, const double* C, int K,
                             const double* sc, const float* kappa_bar, float* maps_bar, int accumulate, double* dots,
                             hipStream_t s);

namespace {
thread_local std::string g_err;


#define PMX_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return pmx_set_error(e_ == hipErrorOutOfMemory ? 4 : 3, std::string(#call) + ": " +    \
                                                                  hipGetErrorString(e_));    \
  } while (0)
#define PMX_TRY(expr)       \
  do {                      \
    int rc_ = (expr);       \
    if (rc_) return rc_;    \
  } while (0)

template <class T> int dev_alloc(T** ptr, size_t count, size_t& total) {
  size_t bytes = count * sizeof(T);
  PMX_HIP(hipMalloc((void**)ptr, bytes < 256 ? 256 : bytes));
  total += bytes;
  return 0;
}

std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  return t;
}

// numpy fftfreq(n)[i] * 2 pi (cell units); the Nyquist bin i = n/2 maps to -pi
double wfreq(int i, int n) {
  const int f = (i < (n + 1) / 2) ? i : i - n;
  return 2.0 * M_PI * (double)f / (double)n;
}

void axis_tables(int n, int count, std::vector<float>& g, std::vector<float>& w2, std::vector<float>& dec) {
  g.resize(count);
  w2.resize(count);
  dec.resize(count);
  for (int i = 0; i < count; ++i) {
    const double w = wfreq(i, n);
    const bool nyq = (2 * i == n);
    g[i] = nyq ? 0.f : (float)((8.0 * std::sin(w) - std::sin(2.0 * w)) / 6.0);
    w2[i] = (float)(w * w);
    const double h = 0.5 * w;
    const double s = (h == 0.0) ? 1.0 : std::sin(h) / h;
    dec[i] = (float)(1.0 / (s * s * s * s));
  }
}

template <class T> int upload(T* dst, const std::vector<T>& src) {
  PMX_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

int check_plan(pmx_plan_t* p) {
  if (!p) return pmx_set_error(1, "null plan");
  int dev = -1;
  PMX_HIP(hipGetDevice(&dev));
  if (dev != p->device) return pmx_set_error(1, "plan was created on another device than the current one");
  return 0;
}

int check_particles(pmx_plan_t* p, const void* pos, int npx, int npy, int npz, int mode) {
  if (!pos) return pmx_set_error(1, "null particle array");
  if (npx <= 0 || npy <= 0 || npz <= 0) return pmx_set_error(1, "particle grid must be positive");
  if (mode != 0 && mode != 1) return pmx_set_error(1, "mode must be 0 (relative) or 1 (absolute)");
  if (mode == 0 && (p->nx % npx || p->ny % npy || p->nz % npz))
    return pmx_set_error(1, "relative mode needs the particle grid to divide the mesh");
  return 0;
}

PartArgs part_args(float* x, float* v, int npx, int npy, int npz, int mode, pmx_plan_t* p) {
  PartArgs a;
  a.x = x;
  a.v = v;
  a.npx = npx;
  a.npy = npy;
  a.npz = npz;
  a.mode = mode;
  a.sx = (float)p->nx / npx;
  a.sy = (float)p->ny / npy;
  a.sz = (float)p->nz / npz;
  return a;
}

// The map accumulator is sized by pmx_plan_reserve; entry points only check it.
int ensure_acc(pmx_plan_t* p, size_t count) {
  const size_t bytes = count * sizeof(double);
  if (p->acc_bytes >= bytes) return 0;
  return pmx_set_error(4, "map accumulator holds " + std::to_string(p->acc_bytes / sizeof(double)) +
                              " values, this call needs " + std::to_string(count) +
                              ": call pmx_plan_reserve with acc_count >= " + std::to_string(count) + " first");
}

int grow_acc(pmx_plan_t* p, size_t count) {
  const size_t bytes = count * sizeof(double);
  if (p->acc_bytes >= bytes) return 0;
  if (p->acc) PMX_HIP(hipFree(p->acc));
  p->workspace_bytes -= p->acc_bytes;
  p->acc = nullptr;
  p->acc_bytes = 0;
  PMX_HIP(hipMalloc((void**)&p->acc, bytes));
  p->acc_bytes = bytes;
  p->workspace_bytes += bytes;
  return 0;
}

// Per-kernel timing of pmx_step: when enabled, 8 events per step bracket paint, the five FFT passes
// and the readout on the caller's stream.
void mark(StepMarks* m, int i, hipStream_t s) {
  if (m) (void)hipEventRecord(m->e[i], s);
}

int force_mesh_from(pmx_plan_t* p, const float* src, float scale, unsigned flags, float* out, hipStream_t s,
                    StepMarks* m = nullptr) {
  PMX_TRY(pmx_pass_r2c_z(p, src, p->specA, s));
  mark(m, 2, s);
  PMX_TRY(pmx_pass_c2c_y(p, p->specA, p->specB, s));
  mark(m, 3, s);
  KOp op{};
  op.mode = XP_FORCES;
  op.scale = (float)((double)scale / (double)p->ncell);
  op.flags = flags & PMX_DECONV;
  PMX_TRY(pmx_xpass(p, op, p->specB, p->specC, s));
  mark(m, 4, s);
  PMX_TRY(pmx_pass_c2c_y_inv(p, 3, p->specC, p->specD, s));
  mark(m, 5, s);
  return pmx_pass_c2r_z(p, 3, p->specD, out, s);
}

// The potential: p->rho = G src, G = ifft((force_scale / (12 N k^2)) fft) (optionally
// CIC-compensated), a real symmetric operator, so it is also its own adjoint. `src` may be p->rho.
// A force and a sharpening boost at the same positions as one operator: the Green's function times
// a + b exp(-kl^2 / k^2) exp(-k^4 / ks^4), so its readout is c_kick F + v_scale S.
struct PgdMix {
  float a, b, kl2, ks4;
};

int pot_only(pmx_plan_t* p, const float* src, float force_scale, unsigned flags, hipStream_t s,
             StepMarks* m = nullptr, const PgdMix* mix = nullptr) {
  PMX_TRY(pmx_pass_r2c_z(p, src, p->specA, s));
  mark(m, 2, s);
  PMX_TRY(pmx_pass_c2c_y(p, p->specA, p->specB, s));
  mark(m, 3, s);
  KOp op{};
  op.mode = XP_POT;
  op.scale = (float)((double)(mix ? 1.f : force_scale) / (12.0 * (double)p->ncell));
  op.flags = flags & PMX_DECONV;
  if (mix) {
    op.flags |= 4u;
    op.mix_a = mix->a;
    op.mix_b = mix->b;
    op.pgd_kl2 = mix->kl2;
    op.pgd_ks4 = mix->ks4;
  }
  PMX_TRY(pmx_xpass(p, op, p->specB, p->specC, s));
  mark(m, 4, s);
  PMX_TRY(pmx_pass_c2c_y_inv(p, 1, p->specC, p->specD, s));
  mark(m, 5, s);
  PMX_TRY(pmx_pass_c2r_z(p, 1, p->specD, p->rho, s));
  mark(m, 6, s);
  return 0;
}

// Force mesh through the potential: phi = ifft(delta_hat * force_scale / (12 k^2)) (one inverse
// FFT instead of three), then F = 8 (phi[+1] - phi[-1]) - (phi[+2] - phi[-2]) per axis, which is the
// reference's k-space gradient i (8 sin k - sin 2k) / 6 exactly.
int pot_force_mesh_from(pmx_plan_t* p, const float* src, float force_scale, unsigned flags, float* out, hipStream_t s,
                        StepMarks* m = nullptr) {
  PMX_TRY(pot_only(p, src, force_scale, flags, s, m));
  return pmx_grad4(p, p->rho, out, s);
}

// Force adjoint at the last sorted particle set `a` (whose force mesh is in p->fmesh):
//   xbar (+)= J^T (umul * u),  J = d forces / d x,  forces = R(x) S G P(x)
// = sum_c grad w_c(x) (umul u . M[c] + rhobar[c]),  rhobar = G S^T R(x)^T (umul u).
// R^T is the sorted int64 deposit of the three components (scale from max |u| in p->ps_misc[3]);
// S^T = -sum_d S_d; dot (+)= u . F. With vkick the same gather also applies v += F ckick (the
// reverse step's un-kick).
// M: the force mesh at the particles' positions (the plan's own, or one the caller kept)
// mix: the combined force + sharpening operator of M (then force_scale is not used)
int force_adjoint(pmx_plan_t* p, const PartArgs& a, const float* u, float umul, float force_scale, unsigned flags,
                  const float* M, float* xbar, int accumulate, double* dot, float* vkick, float ckick, hipStream_t s,
                  const PgdMix* mix = nullptr) {
  float* U = reinterpret_cast<float*>(p->specC);  // 3 scalar meshes; consumed by div4 before the FFT reuses it
  PMX_TRY(pmx_launch_set_scale(p->ps_misc + 3, umul, p->ps_scale, s));
  PMX_TRY(pmx_paint_deposit(p, a, u, 3, umul, 0.f, 1, U, 1, 0, s, 3, p->ncell));
  PMX_TRY(pmx_launch_div4(p, U, U + p->ncell, U + 2 * p->ncell, 1, p->rho, s));
  PMX_TRY(pot_only(p, p->rho, force_scale, flags, s, nullptr, mix));
  return pmx_launch_vjp_gather(p, a, 3, u, umul, M, p->rho, nullptr, 1.f, xbar, accumulate, nullptr, dot, vkick, ckick,
                               s);
}

// Events come from a pool created when timing is enabled, so a timed step creates none (event
// creation can stall the host); the pool grows only when more steps are timed than reserved.
bool add_marks(pmx_plan_t* p) {
  StepMarks m;
  for (int i = 0; i < kStepMarks; ++i)
    if (hipEventCreate(&m.e[i]) != hipSuccess) return false;
  p->marks->push_back(m);
  return true;
}

StepMarks* new_marks(pmx_plan_t* p) {
  if (!p->timing || !p->marks) return nullptr;
  if (p->marks_used == p->marks->size() && !add_marks(p)) return nullptr;
  return &(*p->marks)[p->marks_used++];
}

void clear_marks(pmx_plan_t* p) {
  if (!p->marks) return;
  for (auto& m : *p->marks)
    for (int i = 0; i < kStepMarks; ++i) (void)hipEventDestroy(m.e[i]);
  p->marks->clear();
  p->marks_used = 0;
}

}  // namespace

int pmx_set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int pmx_check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return pmx_set_error(3, std::string("kernel launch ") + what + ": " + hipGetErrorString(e));
  return 0;
}

extern "C" {

}  // extern "C"

int pmx_pot_force_mesh(pmx_plan_t* p, const float* src, float force_scale, hipStream_t s) {
  return pot_force_mesh_from(p, src, force_scale, 0u, p->fmesh, s);
}

extern "C" {

int pmx_abi_version(void) { return 1; }
const char* pmx_last_error(void) { return g_err.c_str(); }
int pmx_fft_supported(int n) { return pmx_fft_len_ok(n) ? 1 : 0; }

unsigned pmx_grid_cap = 1u << 20;

int pmx_set_grid_cap(unsigned max_blocks) {
  if (max_blocks == 0) max_blocks = 1u << 20;
  if (max_blocks > (1u << 20)) return pmx_set_error(1, "grid cap above 2^20 workgroups");
  pmx_grid_cap = max_blocks;
  return 0;
}

int pmx_plan_create(pmx_plan* out, int nx, int ny, int nz, unsigned flags) {
  if (!out) return pmx_set_error(1, "null output handle");
  *out = nullptr;
  if (!pmx_fft_len_ok(nx) || !pmx_fft_len_ok(ny) || !pmx_fft_len_ok(nz))
    return pmx_set_error(2, "unsupported mesh size (each axis must be one of 8,16,32,64,96,128,192,256,512,1024)");
  pmx_plan_t* p = new pmx_plan_t();
  std::memset(p, 0, sizeof(*p));
  p->nslots = 1;
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->nzh = nz / 2 + 1;
  p->flags = flags;
  p->ncell = (size_t)nx * ny * nz;
  p->nspec = (size_t)p->nzh * nx * ny;
  p->nzhp = (p->nzh + 7) & ~7;
  p->nspecd = (size_t)p->nzhp * nx * ny;
  auto fail = [&](int rc) {
    pmx_plan_destroy(p);
    return rc;
  };
  hipError_t e = hipGetDevice(&p->device);
  if (e != hipSuccess) {
    delete p;
    return pmx_set_error(3, std::string("hipGetDevice: ") + hipGetErrorString(e));
  }
  size_t tot = 0;
  int rc = 0;
  if ((rc = dev_alloc(&p->tw_x, nx, tot)) || (rc = dev_alloc(&p->tw_y, ny, tot)) || (rc = dev_alloc(&p->tw_z, nz, tot)) ||
      (rc = dev_alloc(&p->tw_zh, nz / 2, tot)) || (rc = dev_alloc(&p->grad_x, nx, tot)) ||
      (rc = dev_alloc(&p->grad_y, ny, tot)) || (rc = dev_alloc(&p->grad_z, p->nzh, tot)) ||
      (rc = dev_alloc(&p->w2_x, nx, tot)) || (rc = dev_alloc(&p->w2_y, ny, tot)) || (rc = dev_alloc(&p->w2_z, p->nzh, tot)) ||
      (rc = dev_alloc(&p->dec_x, nx, tot)) || (rc = dev_alloc(&p->dec_y, ny, tot)) ||
      (rc = dev_alloc(&p->dec_z, p->nzh, tot)) || (rc = dev_alloc(&p->rho, p->ncell, tot)) ||
      (rc = dev_alloc(&p->specA, p->nspec, tot)) || (rc = dev_alloc(&p->specB, p->nspec, tot)) ||
      (rc = dev_alloc(&p->specC, 3 * p->nspec, tot)) || (rc = dev_alloc(&p->specD, 3 * p->nspecd, tot)) ||
      (rc = dev_alloc(&p->fmesh, 3 * p->ncell, tot)))
    return fail(rc);
  if ((rc = dev_alloc(&p->kphys_x, nx, tot)) || (rc = dev_alloc(&p->kphys_y, ny, tot)) ||
      (rc = dev_alloc(&p->kphys_z, p->nzh, tot)))
    return fail(rc);
  int sx, sy, sz, ntiles;
  pmx_paint_tiles(p, sx, sy, sz, ntiles);
  if ((rc = dev_alloc(&p->ps_count, ntiles, tot)) || (rc = dev_alloc(&p->ps_tcount, ntiles, tot)) ||
      (rc = dev_alloc(&p->ps_offset, ntiles, tot)) ||
      (rc = dev_alloc(&p->ps_cursor, ntiles, tot)) || (rc = dev_alloc(&p->ps_misc, 4, tot)) ||
      (rc = dev_alloc(&p->ps_scale, 2, tot)) ||
      (rc = dev_alloc(&p->ps_faces, (size_t)ntiles * pmx_paint_face_cells(p) * 3, tot)) ||
      (rc = dev_alloc(&p->red, (size_t)kRedMax, tot)))
    return fail(rc);
  // the paint's count pass accumulates into zeroed counters; its deposit pass re-zeroes them
  if (hipMemset(p->ps_count, 0, ntiles * sizeof(unsigned)) != hipSuccess ||
      hipMemset(p->ps_misc, 0, 4 * sizeof(unsigned)) != hipSuccess)
    return fail(pmx_set_error(3, "plan: clearing paint counters failed"));
  p->workspace_bytes = tot;
  if (!(flags & PMX_PLAN_NO_RESERVE)) {
    size_t grown = 0;
    rc = pmx_sort_reserve(p, nx, ny, nz, grown);
    p->workspace_bytes += grown;
    if (rc) return fail(rc);
  }
  std::vector<float> g, w2, dec;
  if ((rc = upload(p->tw_x, twiddles(nx))) || (rc = upload(p->tw_y, twiddles(ny))) || (rc = upload(p->tw_z, twiddles(nz))) ||
      (rc = upload(p->tw_zh, twiddles(nz / 2))))
    return fail(rc);
  axis_tables(nx, nx, g, w2, dec);
  if ((rc = upload(p->grad_x, g)) || (rc = upload(p->w2_x, w2)) || (rc = upload(p->dec_x, dec))) return fail(rc);
  axis_tables(ny, ny, g, w2, dec);
  if ((rc = upload(p->grad_y, g)) || (rc = upload(p->w2_y, w2)) || (rc = upload(p->dec_y, dec))) return fail(rc);
  axis_tables(nz, p->nzh, g, w2, dec);
  if ((rc = upload(p->grad_z, g)) || (rc = upload(p->w2_z, w2)) || (rc = upload(p->dec_z, dec))) return fail(rc);
  *out = p;
  return 0;
}

int pmx_plan_reserve(pmx_plan p, int npx, int npy, int npz, size_t acc_count) {
  PMX_TRY(check_plan(p));
  if (npx < 0 || npy < 0 || npz < 0) return pmx_set_error(1, "negative particle grid");
  if (npx > 0 && npy > 0 && npz > 0) {
    if ((size_t)npx * npy * npz * 3 >= ((size_t)1 << 32))
      return pmx_set_error(2, "particle grid too large (3 * count must be below 2^32)");
    size_t grown = 0;
    const int rc = pmx_sort_reserve(p, npx, npy, npz, grown);
    p->workspace_bytes += grown;
    if (rc) return rc;
  }
  if (acc_count > 0) PMX_TRY(grow_acc(p, acc_count));
  return 0;
}

}  // extern "C"

namespace {
// Regions of one slot of a batched plan (byte offsets, each 256-byte aligned): density, spectra A, B,
// C (3), D (3), force mesh, the colouring's k tables (x, y, z), paint-sort counters, offsets, cursors,
// misc, scale, faces, window counts, segments, packs, and the caller's region.
struct BatchLayout {
  size_t o[20];
  size_t stride;
};
BatchLayout batch_layout(int n, size_t extra) {
  pmx_plan_t t;
  std::memset(&t, 0, sizeof(t));
  t.nx = t.ny = t.nz = n;
  int sx, sy, sz, ntiles;
  pmx_paint_tiles(&t, sx, sy, sz, ntiles);
  const size_t ncell = (size_t)n * n * n, nzh = (size_t)n / 2 + 1, nspec = nzh * n * n,
               nspecd = (size_t)((nzh + 7) & ~(size_t)7) * n * n;
  const size_t bytes[20] = {ncell * 4, nspec * 8, nspec * 8, 3 * nspec * 8, 3 * nspecd * 8, 3 * ncell * 4,
                            (size_t)n * 8, (size_t)n * 8, nzh * 8, (size_t)ntiles * 4, (size_t)ntiles * 4,
                            (size_t)ntiles * 4, (size_t)ntiles * 4, 16, 8,
                            (size_t)ntiles * pmx_paint_face_cells(&t) * 3 * 4, pmx_sort_wg_entries(n, n, n) * 4,
                            ncell * 4, ncell * 4, extra};
  BatchLayout L;
  size_t off = 0;
  for (int i = 0; i < 20; ++i) {
    L.o[i] = off;
    off += (bytes[i] + 255) & ~(size_t)255;
  }
  L.stride = off;
  return L;
}
}  // namespace

size_t pmx_batched_slot_bytes(int n, size_t extra) { return batch_layout(n, extra).stride; }

// A batched plan for `nslots` cubic meshes of side n (the request server's size classes): the FFT
// and k-space tables are allocated once and shared; every slot gets, in one arena, the work buffers
// of an ordinary plan (density, spectra, force mesh, k tables for the colouring, paint sort for n^3
// particles) plus `extra` bytes for the caller, each region 256-byte aligned, slots `stride` bytes
// apart. *extra_off is the offset of the caller's region inside a slot. The plan's pointers are slot
// 0's; pmx_slot_view gives an ordinary (nslots = 1) view of one slot.
int pmx_plan_create_batched(pmx_plan_t** out, int n, int nslots, size_t extra, size_t* extra_off) {
  *out = nullptr;
  if (!pmx_fft_len_ok(n)) return pmx_set_error(2, "unsupported mesh size");
  if (nslots < 1) return pmx_set_error(1, "need at least one slot");
  pmx_plan_t* p = new pmx_plan_t();
  std::memset(p, 0, sizeof(*p));
  p->nx = p->ny = p->nz = n;
  p->nzh = n / 2 + 1;
  p->ncell = (size_t)n * n * n;
  p->nspec = (size_t)p->nzh * n * n;
  p->nzhp = (p->nzh + 7) & ~7;
  p->nspecd = (size_t)p->nzhp * n * n;
  auto fail = [&](int rc) {
    pmx_plan_destroy(p);
    return rc;
  };
  hipError_t e = hipGetDevice(&p->device);
  if (e != hipSuccess) {
    delete p;
    return pmx_set_error(3, std::string("hipGetDevice: ") + hipGetErrorString(e));
  }
  size_t tot = 0;
  int rc = 0;
  if ((rc = dev_alloc(&p->tw_x, n, tot)) || (rc = dev_alloc(&p->tw_y, n, tot)) || (rc = dev_alloc(&p->tw_z, n, tot)) ||
      (rc = dev_alloc(&p->tw_zh, n / 2, tot)) || (rc = dev_alloc(&p->grad_x, n, tot)) ||
      (rc = dev_alloc(&p->grad_y, n, tot)) || (rc = dev_alloc(&p->grad_z, p->nzh, tot)) ||
      (rc = dev_alloc(&p->w2_x, n, tot)) || (rc = dev_alloc(&p->w2_y, n, tot)) || (rc = dev_alloc(&p->w2_z, p->nzh, tot)) ||
      (rc = dev_alloc(&p->dec_x, n, tot)) || (rc = dev_alloc(&p->dec_y, n, tot)) ||
      (rc = dev_alloc(&p->dec_z, p->nzh, tot)) || (rc = dev_alloc(&p->red, (size_t)kRedMax, tot)))
    return fail(rc);
  // one slot's regions
  const BatchLayout L = batch_layout(n, extra);
  const size_t o_rho = L.o[0], o_A = L.o[1], o_B = L.o[2], o_C = L.o[3], o_D = L.o[4], o_F = L.o[5], o_kx = L.o[6],
               o_ky = L.o[7], o_kz = L.o[8], o_cnt = L.o[9], o_tcnt = L.o[10], o_off = L.o[11], o_cur = L.o[12],
               o_misc = L.o[13], o_scale = L.o[14], o_faces = L.o[15], o_wg = L.o[16], o_seg = L.o[17],
               o_pack = L.o[18];
  *extra_off = L.o[19];
  const size_t stride = L.stride;
  char* arena = nullptr;
  e = hipMalloc((void**)&arena, stride * (size_t)nslots);
  if (e != hipSuccess) return fail(pmx_set_error(4, std::string("batched plan arena: ") + hipGetErrorString(e)));
  p->arena = arena;
  tot += stride * (size_t)nslots;
  p->rho = reinterpret_cast<float*>(arena + o_rho);
  p->specA = reinterpret_cast<float2*>(arena + o_A);
  p->specB = reinterpret_cast<float2*>(arena + o_B);
  p->specC = reinterpret_cast<float2*>(arena + o_C);
  p->specD = reinterpret_cast<float2*>(arena + o_D);
  p->fmesh = reinterpret_cast<float*>(arena + o_F);
  p->kphys_x = reinterpret_cast<double*>(arena + o_kx);
  p->kphys_y = reinterpret_cast<double*>(arena + o_ky);
  p->kphys_z = reinterpret_cast<double*>(arena + o_kz);
  p->ps_count = reinterpret_cast<unsigned*>(arena + o_cnt);
  p->ps_tcount = reinterpret_cast<unsigned*>(arena + o_tcnt);
  p->ps_offset = reinterpret_cast<unsigned*>(arena + o_off);
  p->ps_cursor = reinterpret_cast<unsigned*>(arena + o_cur);
  p->ps_misc = reinterpret_cast<unsigned*>(arena + o_misc);
  p->ps_scale = reinterpret_cast<float*>(arena + o_scale);
  p->ps_faces = reinterpret_cast<float*>(arena + o_faces);
  p->ps_wgcnt = reinterpret_cast<unsigned*>(arena + o_wg);
  p->ps_seg = reinterpret_cast<unsigned*>(arena + o_seg);
  p->ps_pack = reinterpret_cast<unsigned*>(arena + o_pack);
  p->ps_wg_cap = pmx_sort_wg_entries(n, n, n);
  p->ps_seg_cap = p->ncell;
  p->nslots = nslots;
  p->slots.stride = stride;
  p->workspace_bytes = tot;
  // the count pass accumulates into zeroed counters (its deposit pass re-zeroes them); misc holds the
  // far-particle counters
  if (hipMemset(arena, 0, stride * (size_t)nslots) != hipSuccess)
    return fail(pmx_set_error(3, "batched plan: clearing the arena failed"));
  std::vector<float> g, w2, dec;
  if ((rc = upload(p->tw_x, twiddles(n))) || (rc = upload(p->tw_y, twiddles(n))) || (rc = upload(p->tw_z, twiddles(n))) ||
      (rc = upload(p->tw_zh, twiddles(n / 2))))
    return fail(rc);
  axis_tables(n, n, g, w2, dec);
  if ((rc = upload(p->grad_x, g)) || (rc = upload(p->w2_x, w2)) || (rc = upload(p->dec_x, dec)) ||
      (rc = upload(p->grad_y, g)) || (rc = upload(p->w2_y, w2)) || (rc = upload(p->dec_y, dec)))
    return fail(rc);
  axis_tables(n, p->nzh, g, w2, dec);
  if ((rc = upload(p->grad_z, g)) || (rc = upload(p->w2_z, w2)) || (rc = upload(p->dec_z, dec))) return fail(rc);
  *out = p;
  return 0;
}

// An ordinary plan (nslots = 1) over slot z of a batched plan: the same shared tables, the slot's
// buffers. It owns nothing.
pmx_plan_t pmx_slot_view(const pmx_plan_t* b, int z) {
  pmx_plan_t v = *b;
  const size_t st = b->slots.stride;
  v.nslots = 1;
  v.slots = Slots{};
  v.arena = nullptr;
  v.marks = nullptr;
  v.timing = 0;
  v.rho = at_slot(b->rho, st, z);
  v.specA = at_slot(b->specA, st, z);
  v.specB = at_slot(b->specB, st, z);
  v.specC = at_slot(b->specC, st, z);
  v.specD = at_slot(b->specD, st, z);
  v.fmesh = at_slot(b->fmesh, st, z);
  v.kphys_x = at_slot(b->kphys_x, st, z);
  v.kphys_y = at_slot(b->kphys_y, st, z);
  v.kphys_z = at_slot(b->kphys_z, st, z);
  v.ps_count = at_slot(b->ps_count, st, z);
  v.ps_tcount = at_slot(b->ps_tcount, st, z);
  v.ps_offset = at_slot(b->ps_offset, st, z);
  v.ps_cursor = at_slot(b->ps_cursor, st, z);
  v.ps_misc = at_slot(b->ps_misc, st, z);
  v.ps_scale = at_slot(b->ps_scale, st, z);
  v.ps_faces = at_slot(b->ps_faces, st, z);
  v.ps_wgcnt = at_slot(b->ps_wgcnt, st, z);
  v.ps_seg = at_slot(b->ps_seg, st, z);
  v.ps_pack = at_slot(b->ps_pack, st, z);
  return v;
}

extern "C" {

int pmx_plan_destroy(pmx_plan p) {
  if (!p) return 0;
  if (p->arena) {  // a batched plan: the shared tables and the arena (slot buffers point into it)
    void* bufs[] = {p->tw_x, p->tw_y, p->tw_z, p->tw_zh, p->grad_x, p->grad_y, p->grad_z, p->w2_x, p->w2_y,
                    p->w2_z, p->dec_x, p->dec_y, p->dec_z, p->red, p->arena};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
    delete p;
    return 0;
  }
  void* bufs[] = {p->tw_x, p->tw_y, p->tw_z, p->tw_zh, p->grad_x, p->grad_y, p->grad_z, p->w2_x, p->w2_y, p->w2_z,
                  p->dec_x, p->dec_y, p->dec_z, p->rho, p->specA, p->specB, p->specC, p->specD, p->fmesh, p->acc,
                  p->kphys_x, p->kphys_y, p->kphys_z, p->ps_count, p->ps_tcount, p->ps_offset, p->ps_cursor, p->ps_wgcnt,
                  p->ps_seg, p->ps_pack, p->ps_misc, p->ps_scale, p->ps_faces, p->red};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  clear_marks(p);
  delete p->marks;
  delete p;
  return 0;
}

size_t pmx_plan_workspace_bytes(pmx_plan p) { return p ? p->workspace_bytes : 0; }

int pmx_plan_escapes(pmx_plan p, uint64_t* count) {
  PMX_TRY(check_plan(p));
  if (!count) return pmx_set_error(1, "null output");
  unsigned c = 0;
  PMX_HIP(hipDeviceSynchronize());
  PMX_HIP(hipMemcpy(&c, p->ps_misc + 2, sizeof(unsigned), hipMemcpyDeviceToHost));
  *count = c;
  return 0;
}

int pmx_plan_set_box(pmx_plan p, const double box[3]) {
  PMX_TRY(check_plan(p));
  if (!box || !(box[0] > 0) || !(box[1] > 0) || !(box[2] > 0)) return pmx_set_error(1, "box must be positive");
  // the estimator's wavevectors: (2 pi n / L) * fftfreq(n), fftfreq(n)[i] = f(i) * (1/n)
  auto tab = [&](int n, int count, double L) {
    std::vector<double> t(count);
    const double val = 1.0 / (double)n;
    for (int i = 0; i < count; ++i) {
      const int f = (i < (n + 1) / 2) ? i : i - n;
      t[i] = (2.0 * M_PI * (double)n / L) * ((double)f * val);
    }
    return t;
  };
  PMX_TRY(upload(p->kphys_x, tab(p->nx, p->nx, box[0])));
  PMX_TRY(upload(p->kphys_y, tab(p->ny, p->ny, box[1])));
  PMX_TRY(upload(p->kphys_z, tab(p->nz, p->nzh, box[2])));
  p->box_set = true;
  return 0;
}

int pmx_paint_cic(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, const float* weights, float weight,
                  float* mesh, int accumulate, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!mesh) return pmx_set_error(1, "null mesh");
  hipStream_t s = (hipStream_t)stream;
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  return pmx_launch_paint(p, a, weights, weight, 0.f, mesh, accumulate, (p->flags & PMX_PAINT_ATOMIC) != 0, s);
}

int pmx_readout_cic(pmx_plan p, const float* mesh, int ncomp, const float* pos, int npx, int npy, int npz, int mode,
                    float* out, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!mesh || !out) return pmx_set_error(1, "null mesh or output");
  if (ncomp != 1 && ncomp != 3) return pmx_set_error(1, "ncomp must be 1 or 3");
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  return pmx_launch_readout(p, a, mesh, ncomp, out, 0.f, 0.f, 1, (hipStream_t)stream);
}

int pmx_fft_r2c(pmx_plan p, const float* in, void* spec, void* stream) {
  PMX_TRY(check_plan(p));
  if (!in || !spec) return pmx_set_error(1, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  PMX_TRY(pmx_fft_forward_to_B(p, in, s));
  KOp op{};
  op.mode = XP_FWD;
  return pmx_xpass(p, op, p->specB, (float2*)spec, s);
}

int pmx_fft_c2r(pmx_plan p, const void* spec, float* out, void* stream) {
  PMX_TRY(check_plan(p));
  if (!spec || !out) return pmx_set_error(1, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  KOp op{};
  op.mode = XP_INV;
  PMX_TRY(pmx_xpass(p, op, (const float2*)spec, p->specC, s));
  return pmx_inverse_from_C(p, 1, p->specC, out, s);
}

int pmx_force_mesh(pmx_plan p, const float* density, float scale, unsigned flags, float* fmesh, void* stream) {
  PMX_TRY(check_plan(p));
  if (!density) return pmx_set_error(1, "null density");
  return force_mesh_from(p, density, scale, flags, fmesh ? fmesh : p->fmesh, (hipStream_t)stream);
}

int pmx_forces(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, float scale, unsigned flags,
               float* forces, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!forces) return pmx_set_error(1, "null forces");
  hipStream_t s = (hipStream_t)stream;
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  PMX_TRY(pmx_launch_paint(p, a, nullptr, 1.f, 0.f, p->rho, 0, ((p->flags | flags) & PMX_PAINT_ATOMIC) != 0, s));
  PMX_TRY(pot_force_mesh_from(p, p->rho, scale, flags, p->fmesh, s));
  return pmx_launch_readout(p, a, p->fmesh, 3, forces, 0.f, 0.f, 1, s);
}

int pmx_step(pmx_plan p, float* x, float* v, int npx, int npy, int npz, int mode, float force_scale, float c_drift_pre,
             float c_kick, float c_drift_post, unsigned flags, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, x, npx, npy, npz, mode));
  if (!v) return pmx_set_error(1, "null velocities");
  hipStream_t s = (hipStream_t)stream;
  StepMarks* m = new_marks(p);
  mark(m, 0, s);
  PartArgs a = part_args(x, (flags & PMX_DRIFT_PRE) ? v : nullptr, npx, npy, npz, mode, p);
  PMX_TRY(pmx_launch_paint(p, a, nullptr, 1.f, c_drift_pre, p->rho, 0, ((p->flags | flags) & PMX_PAINT_ATOMIC) != 0, s));
  mark(m, 1, s);
  PMX_TRY(pot_force_mesh_from(p, p->rho, force_scale, flags, p->fmesh, s, m));
  mark(m, 7, s);
  a.v = v;
  const int what = 2 | ((flags & PMX_DRIFT_POST) ? 4 : 0);
  PMX_TRY(pmx_launch_readout(p, a, p->fmesh, 3, nullptr, c_kick, c_drift_post, what, s));
  mark(m, 8, s);
  return 0;
}

int pmx_step_keep(pmx_plan p, float* x, float* v, int npx, int npy, int npz, int mode, float force_scale,
                  float c_drift_pre, float c_kick, float c_drift_post, unsigned flags, float* force_mesh, void* stream) {
  PMX_TRY(check_plan(p));
  if (!force_mesh) return pmx_set_error(1, "null force mesh");
  // the launches capture the buffer pointer when they are queued, so the plan's own mesh can be
  // restored at once
  float* own = p->fmesh;
  p->fmesh = force_mesh;
  const int rc = pmx_step(p, x, v, npx, npy, npz, mode, force_scale, c_drift_pre, c_kick, c_drift_post, flags, stream);
  p->fmesh = own;
  return rc;
}

int pmx_pgd(pmx_plan p, float* x, float* v, int npx, int npy, int npz, int mode, float alpha, float kl, float ks,
            int target, float v_scale, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, x, npx, npy, npz, mode));
  if (target != 0 && target != 1) return pmx_set_error(1, "target must be 0 (positions) or 1 (velocities)");
  if (target == 1 && !v) return pmx_set_error(1, "null velocities");
  if (!(ks > 0.f) || !(kl >= 0.f)) return pmx_set_error(1, "need kl >= 0 and ks > 0");
  hipStream_t s = (hipStream_t)stream;
  PartArgs a = part_args(x, nullptr, npx, npy, npz, mode, p);
  PMX_TRY(pmx_launch_paint(p, a, nullptr, 1.f, 0.f, p->rho, 0, (p->flags & PMX_PAINT_ATOMIC) != 0, s));
  KOp op{};
  op.mode = XP_POT;
  op.scale = (float)((double)alpha / (12.0 * (double)p->ncell));
  op.flags = 2u;
  op.pgd_kl2 = kl * kl;
  op.pgd_ks4 = (ks * ks) * (ks * ks);
  PMX_TRY(pmx_pass_r2c_z(p, p->rho, p->specA, s));
  PMX_TRY(pmx_pass_c2c_y(p, p->specA, p->specB, s));
  PMX_TRY(pmx_xpass(p, op, p->specB, p->specC, s));
  PMX_TRY(pmx_pass_c2c_y_inv(p, 1, p->specC, p->specD, s));
  PMX_TRY(pmx_pass_c2r_z(p, 1, p->specD, p->rho, s));
  PMX_TRY(pmx_grad4(p, p->rho, p->fmesh, s));
  a.v = v;
  return pmx_launch_readout(p, a, p->fmesh, 3, nullptr, target == 0 ? 1.f : v_scale, 0.f, target == 0 ? 8 : 2, s);
}

int pmx_forces_vjp(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, float scale, unsigned flags,
                   const float* forces_bar, float bar_scale, float* pos_bar, int accumulate, double* dot,
                   void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!forces_bar || !pos_bar) return pmx_set_error(1, "null cotangent");
  if (!std::isfinite(bar_scale)) return pmx_set_error(1, "bar_scale must be finite");
  hipStream_t s = (hipStream_t)stream;
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  const size_t np = (size_t)npx * npy * npz;
  PMX_TRY(pmx_paint_sort(p, a, 0.f, nullptr, 1, 1.f, s));
  PMX_TRY(pmx_paint_deposit(p, a, nullptr, 1, 1.f, 1.f, 0, p->rho, 1, 0, s));
  PMX_TRY(pot_force_mesh_from(p, p->rho, scale, flags, p->fmesh, s));
  PMX_TRY(pmx_launch_absmax(forces_bar, 3 * np, p->ps_misc + 3, s));
  return force_adjoint(p, a, forces_bar, bar_scale, scale, flags, p->fmesh, pos_bar, accumulate, dot, nullptr, 0.f,
                       s);
}

int pmx_forces_vjp_saved(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, float scale,
                         unsigned flags, const float* force_mesh, const float* forces_bar, float bar_scale,
                         float* pos_bar, int accumulate, double* dot, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!forces_bar || !pos_bar || !force_mesh) return pmx_set_error(1, "null buffer");
  if (!std::isfinite(bar_scale)) return pmx_set_error(1, "bar_scale must be finite");
  hipStream_t s = (hipStream_t)stream;
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  const size_t np = (size_t)npx * npy * npz;
  PMX_TRY(pmx_paint_sort(p, a, 0.f, nullptr, 1, 1.f, s));  // the adjoint's deposit reuses the sort
  PMX_TRY(pmx_launch_absmax(forces_bar, 3 * np, p->ps_misc + 3, s));
  return force_adjoint(p, a, forces_bar, bar_scale, scale, flags, force_mesh, pos_bar, accumulate, dot, nullptr, 0.f,
                       s);
}

int pmx_axpy(float* out, const float* x, const float* v, float c, size_t n, void* stream) {
  if (!out || !x || !v) return pmx_set_error(1, "null buffer");
  if (n == 0) return 0;
  return pmx_launch_axpy(out, x, v, c, n, (hipStream_t)stream);
}

int pmx_adjoint_drift(pmx_plan p, const float* xbar, float* vbar, const float* v, size_t n, float c, double* dot,
                      void* stream) {
  PMX_TRY(check_plan(p));
  if (!xbar || !vbar || !v || !dot) return pmx_set_error(1, "null buffer");
  if (n == 0) return 0;
  return pmx_launch_adj_prep(p, xbar, vbar, v, n, c, dot, nullptr, (hipStream_t)stream);
}

int pmx_pgd_vjp(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, float alpha, float kl, float ks,
                const float* disp_bar, float* pos_bar, int accumulate, double* dot, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!disp_bar || !pos_bar) return pmx_set_error(1, "null cotangent");
  if (!(ks > 0.f) || !(kl >= 0.f)) return pmx_set_error(1, "need kl >= 0 and ks > 0");
  hipStream_t s = (hipStream_t)stream;
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  const size_t np = (size_t)npx * npy * npz;
  const PgdMix mix{0.f, alpha, kl * kl, (ks * ks) * (ks * ks)};  // the PGD operator alone
  PMX_TRY(pmx_paint_sort(p, a, 0.f, nullptr, 1, 1.f, s));
  PMX_TRY(pmx_paint_deposit(p, a, nullptr, 1, 1.f, 1.f, 0, p->rho, 1, 0, s));
  PMX_TRY(pot_only(p, p->rho, 1.f, 0, s, nullptr, &mix));
  PMX_TRY(pmx_grad4(p, p->rho, p->fmesh, s));
  PMX_TRY(pmx_launch_absmax(disp_bar, 3 * np, p->ps_misc + 3, s));
  return force_adjoint(p, a, disp_bar, 1.f, 1.f, 0, p->fmesh, pos_bar, accumulate, dot, nullptr, 0.f, s, &mix);
}

int pmx_force_mesh_vjp(pmx_plan p, const float* fmesh_bar, float scale, unsigned flags, float* density_bar,
                       void* stream) {
  PMX_TRY(check_plan(p));
  if (!fmesh_bar || !density_bar) return pmx_set_error(1, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  PMX_TRY(pmx_launch_div4(p, fmesh_bar, fmesh_bar + 1, fmesh_bar + 2, 3, p->rho, s));
  PMX_TRY(pot_only(p, p->rho, scale, flags, s));
  PMX_HIP(hipMemcpyAsync(density_bar, p->rho, p->ncell * sizeof(float), hipMemcpyDeviceToDevice, s));
  return 0;
}

int pmx_paint_cic_vjp(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                      float weight, const float* mesh_bar, float* pos_bar, float* weights_bar, int accumulate,
                      void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!mesh_bar || !pos_bar) return pmx_set_error(1, "null cotangent");
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  return pmx_launch_vjp_gather(p, a, 1, nullptr, 0.f, nullptr, mesh_bar, weights, weight, pos_bar, accumulate,
                               weights_bar, nullptr, nullptr, 0.f, (hipStream_t)stream);
}

int pmx_readout_cic_vjp(pmx_plan p, const float* mesh, int ncomp, const float* pos, int npx, int npy, int npz, int mode,
                        const float* out_bar, float* mesh_bar, float* pos_bar, int accumulate, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (ncomp != 1 && ncomp != 3) return pmx_set_error(1, "ncomp must be 1 or 3");
  if (!out_bar || (!mesh_bar && !pos_bar) || (pos_bar && !mesh)) return pmx_set_error(1, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  if (pos_bar)
    PMX_TRY(pmx_launch_vjp_gather(p, a, ncomp, out_bar, 1.f, mesh, nullptr, nullptr, 0.f, pos_bar, accumulate,
                                  nullptr, nullptr, nullptr, 0.f, s));
  if (mesh_bar) {
    PMX_TRY(pmx_paint_sort(p, a, 0.f, out_bar, ncomp, 0.f, s));
    PMX_TRY(pmx_paint_deposit(p, a, out_bar, ncomp, 1.f, 0.f, 1, mesh_bar, ncomp, accumulate, s, ncomp, 1));
  }
  return 0;
}

int pmx_step_reverse_vjp(pmx_plan p, float* x, float* v, int npx, int npy, int npz, int mode, float force_scale,
                         float c_kick, float c_drift, unsigned flags, float* x_bar, float* v_bar, double* coef_bar,
                         void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, x, npx, npy, npz, mode));
  if (!v || !x_bar || !v_bar || !coef_bar) return pmx_set_error(1, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)npx * npy * npz * 3;
  PMX_TRY(pmx_launch_fill_d(coef_bar, 2, s));
  // cotangents through the drift x' = x + v' cd (v' the kicked velocity, still in v)
  PMX_TRY(pmx_launch_adj_prep(p, x_bar, v_bar, v, n, c_drift, coef_bar, p->ps_misc + 3, s));
  // reverse the state: x -= v cd, then (in the adjoint gather) v -= F(x) ck
  PartArgs a = part_args(x, v, npx, npy, npz, mode, p);
  PMX_TRY(pmx_paint_sort(p, a, -c_drift, nullptr, 1, 1.f, s));
  a.v = nullptr;
  PMX_TRY(pmx_paint_deposit(p, a, nullptr, 1, 1.f, 1.f, 0, p->rho, 1, 0, s));
  PMX_TRY(pot_force_mesh_from(p, p->rho, force_scale, flags, p->fmesh, s));
  // x_bar += J(x)^T (ck v_bar); coef_bar[1] = v_bar . F(x); the same gather un-kicks v -= F(x) ck
  return force_adjoint(p, a, v_bar, c_kick, force_scale, flags, p->fmesh, x_bar, 1, coef_bar + 1, v, -c_kick, s);
}

int pmx_step_reverse_vjp_sharpen(pmx_plan p, float* x, float* v, int npx, int npy, int npz, int mode,
                                 float force_scale, float c_kick, float c_drift, unsigned flags, float alpha, float kl,
                                 float ks, float v_scale, float* x_bar, float* v_bar, double* coef_bar, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, x, npx, npy, npz, mode));
  if (!v || !x_bar || !v_bar || !coef_bar) return pmx_set_error(1, "null buffer");
  if (!(ks > 0.f) || !(kl >= 0.f)) return pmx_set_error(1, "need kl >= 0 and ks > 0");
  if (flags & PMX_DECONV) return pmx_set_error(1, "the sharpening boost has no CIC deconvolution to combine with");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)npx * npy * npz * 3;
  const PgdMix mix{c_kick * force_scale, v_scale * alpha, kl * kl, (ks * ks) * (ks * ks)};
  PMX_TRY(pmx_launch_fill_d(coef_bar, 2, s));
  PMX_TRY(pmx_launch_adj_prep(p, x_bar, v_bar, v, n, c_drift, coef_bar, p->ps_misc + 3, s));
  PartArgs a = part_args(x, v, npx, npy, npz, mode, p);
  PMX_TRY(pmx_paint_sort(p, a, -c_drift, nullptr, 1, 1.f, s));
  a.v = nullptr;
  PMX_TRY(pmx_paint_deposit(p, a, nullptr, 1, 1.f, 1.f, 0, p->rho, 1, 0, s));
  // M = the combined operator's mesh at x: its readout is c_kick F(x) + v_scale S(x)
  PMX_TRY(pot_only(p, p->rho, 1.f, flags, s, nullptr, &mix));
  PMX_TRY(pmx_grad4(p, p->rho, p->fmesh, s));
  // x_bar += J_F^T (c_kick v_bar) + J_S^T (v_scale v_bar) in one gather, which also undoes the kick
  // and the boost: v -= c_kick F(x) + v_scale S(x)
  return force_adjoint(p, a, v_bar, 1.f, force_scale, flags, p->fmesh, x_bar, 1, coef_bar + 1, v, -1.f, s, &mix);
}

int pmx_step_reverse_vjp_saved(pmx_plan p, float* x, float* v, int npx, int npy, int npz, int mode,
                               float force_scale, float c_kick, float c_drift, unsigned flags, const float* force_mesh,
                               float* x_bar, float* v_bar, double* coef_bar, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, x, npx, npy, npz, mode));
  if (!v || !x_bar || !v_bar || !coef_bar || !force_mesh) return pmx_set_error(1, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)npx * npy * npz * 3;
  PMX_TRY(pmx_launch_fill_d(coef_bar, 2, s));
  PMX_TRY(pmx_launch_adj_prep(p, x_bar, v_bar, v, n, c_drift, coef_bar, p->ps_misc + 3, s));
  // x -= v cd inside the sort's count pass; the sort serves the adjoint's deposit, and the force
  // mesh at x is the one the forward step kept
  PartArgs a = part_args(x, v, npx, npy, npz, mode, p);
  PMX_TRY(pmx_paint_sort(p, a, -c_drift, nullptr, 1, 1.f, s));
  a.v = nullptr;
  return force_adjoint(p, a, v_bar, c_kick, force_scale, flags, force_mesh, x_bar, 1, coef_bar + 1, v, -c_kick, s);
}

int pmx_plan_timing(pmx_plan p, int enable) {
  PMX_TRY(check_plan(p));
  if (!p->marks) p->marks = new std::vector<StepMarks>();
  if (enable > 0) {
    clear_marks(p);
    const int reserve = enable > 64 ? enable : 64;
    p->marks->reserve(reserve);
    for (int i = 0; i < reserve; ++i)
      if (!add_marks(p)) return pmx_set_error(3, "timing: hipEventCreate failed");
  }
  p->timing = enable > 0 ? 1 : 0;
  return 0;
}

int pmx_plan_timing_read(pmx_plan p, double* ms, int* nsteps) {
  PMX_TRY(check_plan(p));
  if (!ms || !nsteps) return pmx_set_error(1, "null output");
  for (int i = 0; i < kStepMarks - 1; ++i) ms[i] = 0.0;  // kStepMarks - 1 phases
  *nsteps = 0;
  if (!p->marks || p->marks_used == 0) return 0;
  PMX_HIP(hipEventSynchronize((*p->marks)[p->marks_used - 1].e[kStepMarks - 1]));
  for (size_t k = 0; k < p->marks_used; ++k) {
    const StepMarks& mk = (*p->marks)[k];
    for (int i = 0; i < kStepMarks - 1; ++i) {
      float t = 0.f;
      PMX_HIP(hipEventElapsedTime(&t, mk.e[i], mk.e[i + 1]));
      ms[i] += t;
    }
  }
  *nsteps = (int)p->marks_used;
  return 0;
}

int pmx_power(pmx_plan p, const float* mesh, const double* kedges, int nedges, double* pk_sum, double* k_sum,
              double* k_cnt, void* stream) {
  const int pole0 = 0;
  return pmx_power_multipoles(p, mesh, nullptr, kedges, nedges, &pole0, 1, nullptr, pk_sum, nullptr, k_sum, k_cnt,
                              stream);
}

int pmx_power_multipoles(pmx_plan p, const float* mesh, const float* mesh2, const double* kedges, int nedges,
                         const int* poles, int npoles, const double* los, double* p_re, double* p_im,
                         double* k_sum, double* k_cnt, void* stream) {
  PMX_TRY(check_plan(p));
  if (!p->box_set) return pmx_set_error(1, "pmx_plan_set_box must be called before pmx_power");
  if (!mesh || !kedges || !p_re || !k_sum || !k_cnt || nedges < 1 || nedges > 512)
    return pmx_set_error(1, "bad power-spectrum arguments");
  if (!poles || npoles < 1 || npoles > 16) return pmx_set_error(1, "npoles must be in [1, 16]");
  int lmax = 0;
  for (int j = 0; j < npoles; ++j) {
    if (poles[j] < 0 || poles[j] > 16) return pmx_set_error(1, "multipole orders must be in [0, 16]");
    lmax = std::max(lmax, poles[j]);
  }
  if (mesh2 && !p_im) return pmx_set_error(1, "a cross spectrum needs p_im");
  double u[3] = {0.0, 0.0, 0.0};
  if (lmax > 0) {
    if (!los) return pmx_set_error(1, "multipoles need a line of sight");
    const double nrm = std::sqrt(los[0] * los[0] + los[1] * los[1] + los[2] * los[2]);
    if (!(nrm > 0.0) || !std::isfinite(nrm)) return pmx_set_error(1, "line of sight must be a nonzero vector");
    for (int c = 0; c < 3; ++c) u[c] = los[c] / nrm;
  }
  const int per = pmx_xpass_power_poles(p, nedges);
  if (per < 1) return pmx_set_error(1, "too many k bins for this mesh");
  hipStream_t s = (hipStream_t)stream;
  const int nb = nedges + 1;
  PMX_TRY(pmx_launch_fill_d(p_re, (size_t)npoles * nb, s));
  if (mesh2) PMX_TRY(pmx_launch_fill_d(p_im, (size_t)npoles * nb, s));
  PMX_TRY(pmx_launch_fill_d(k_sum, nb, s));
  PMX_TRY(pmx_launch_fill_d(k_cnt, nb, s));
  KOp op{};
  if (mesh2) {  // m2 in full [kz][ky][kx] layout in specC; specA/B are then reused for m1
    PMX_TRY(pmx_fft_forward_to_B(p, mesh2, s));
    op.mode = XP_FWD;
    PMX_TRY(pmx_xpass(p, op, p->specB, p->specC, s));
  }
  PMX_TRY(pmx_fft_forward_to_B(p, mesh, s));
  op.mode = XP_POWER;
  op.kphys_x = p->kphys_x;
  op.kphys_y = p->kphys_y;
  op.kphys_z = p->kphys_z;
  op.kedges = kedges;
  op.nedges = nedges;
  op.k_sum = k_sum;
  op.k_cnt = k_cnt;
  op.spec2 = mesh2 ? p->specC : nullptr;
  for (int c = 0; c < 3; ++c) op.los[c] = u[c];
  for (int j0 = 0; j0 < npoles; j0 += per) {
    op.npoles = std::min(per, npoles - j0);
    for (int j = 0; j < op.npoles; ++j) op.poles[j] = poles[j0 + j];
    op.pk_sum = p_re + (size_t)j0 * nb;
    op.pk_im = mesh2 ? p_im + (size_t)j0 * nb : nullptr;
    op.acc_k = (j0 == 0);
    PMX_TRY(pmx_xpass(p, op, p->specB, nullptr, s));
  }
  return 0;
}

int pmx_plan_set_rbf_width(pmx_plan p, double sigma_rad) {
  PMX_TRY(check_plan(p));
  if (!(sigma_rad >= 0.0) || !std::isfinite(sigma_rad)) return pmx_set_error(1, "rbf width must be finite and >= 0");
  p->rbf_sigma = sigma_rad;
  return 0;
}

int pmx_healpix_ud_grade(const float* in, int nside_in, float* out, int nside_out, float power, int order_in,
                         int order_out, int pess, void* stream) {
  auto pow2 = [](int n) { return n >= 1 && n <= (1 << 13) && (n & (n - 1)) == 0; };
  if (!in || !out || in == out) return pmx_set_error(1, "ud_grade needs distinct input and output maps");
  if (!pow2(nside_in) || !pow2(nside_out)) return pmx_set_error(1, "ud_grade: nside must be a power of two <= 8192");
  if ((order_in != 0 && order_in != 1) || (order_out != 0 && order_out != 1))
    return pmx_set_error(1, "ud_grade: order must be 0 (RING) or 1 (NESTED)");
  const double ratio = power != 0.f ? std::pow((double)nside_out / (double)nside_in, (double)power) : 1.0;
  return pmx_launch_ud_grade(in, nside_in, out, nside_out, ratio, order_in == 0, order_out == 0, pess != 0,
                             (hipStream_t)stream);
}

int pmx_white_noise(pmx_plan p, uint64_t seed, float* out, void* stream) {
  PMX_TRY(check_plan(p));
  if (!out) return pmx_set_error(1, "null output");
  return pmx_launch_white_noise(out, p->ncell, seed, (hipStream_t)stream);
}

int pmx_philox4x32(uint64_t seed, size_t nblocks, uint32_t* out, void* stream) {
  if (!out) return pmx_set_error(1, "null output");
  return pmx_launch_philox_raw(out, nblocks, seed, (hipStream_t)stream);
}

int pmx_color_field(pmx_plan p, const float* field, const double* ktab, const double* ptab, int ntab, float* out,
                    void* stream) {
  PMX_TRY(check_plan(p));
  if (!p->box_set) return pmx_set_error(1, "pmx_plan_set_box must be called before pmx_color_field");
  if (!field || !ktab || !ptab || ntab < 2 || !out) return pmx_set_error(1, "bad colouring arguments");
  hipStream_t s = (hipStream_t)stream;
  PMX_TRY(pmx_fft_forward_to_B(p, field, s));
  KOp op{};
  op.mode = XP_COLOR;
  op.scale = (float)(1.0 / (double)p->ncell);
  op.kphys_x = p->kphys_x;
  op.kphys_y = p->kphys_y;
  op.kphys_z = p->kphys_z;
  op.ktab = ktab;
  op.ptab = ptab;
  op.ntab = ntab;
  op.cnorm = 1.0;  // the caller folds ncell / volume into ptab
  PMX_TRY(pmx_xpass(p, op, p->specB, p->specC, s));
  return pmx_inverse_from_C(p, 1, p->specC, out, s);
}

int pmx_lpt1(pmx_plan p, const float* ic, float growth, float vel_factor, float* dx, float* pv, void* stream) {
  PMX_TRY(check_plan(p));
  return pmx_lpt1_lattice(p, ic, p->nx, p->ny, p->nz, growth, vel_factor, dx, pv, stream);
}

int pmx_lpt1_lattice(pmx_plan p, const float* ic, int npx, int npy, int npz, float growth, float vel_factor, float* dx,
                     float* pv, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, dx, npx, npy, npz, 0));
  if (!ic || !pv) return pmx_set_error(1, "null buffer");
  hipStream_t s = (hipStream_t)stream;
  PMX_TRY(force_mesh_from(p, ic, 1.0f, 0u, p->fmesh, s));
  return pmx_launch_lpt1(p, p->fmesh, dx, pv, npx, npy, npz, growth, vel_factor, s);
}

int pmx_lpt(pmx_plan p, const float* ic, int npx, int npy, int npz, int order, const double* coef, int ntab,
            double r_max, const double box[3], const double observer[3], float* dx, float* pv, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, dx, npx, npy, npz, 0));
  if (!ic || !pv || !coef) return pmx_set_error(1, "null buffer");
  if (order != 1 && order != 2) return pmx_set_error(1, "order must be 1 or 2");
  if (ntab < 1 || ntab > (1 << 20)) return pmx_set_error(1, "ntab must be in [1, 2^20]");
  LptGeom g{};
  g.coef = coef;
  g.ntab = ntab;
  if (ntab > 1) {
    if (!box || !observer || !(r_max > 0.0)) return pmx_set_error(1, "a coefficient table needs box, observer, r_max > 0");
    const int n[3] = {p->nx, p->ny, p->nz};
    for (int c = 0; c < 3; ++c) {
      if (!(box[c] > 0.0)) return pmx_set_error(1, "box must be positive");
      g.cell[c] = box[c] / (double)n[c];
      g.obs[c] = observer[c];
    }
    g.r_max = r_max;
  }
  hipStream_t s = (hipStream_t)stream;
  if (order == 2) {
    // delta2 into rho from the six Hessian components (two passes of three through fmesh)
    PMX_TRY(pmx_fft_forward_to_B(p, ic, s));
    KOp op{};
    op.mode = XP_HESS;
    op.scale = (float)(1.0 / (double)p->ncell);
    for (int off = 0; off < 2; ++off) {
      op.flags = off ? kHessOffDiag : 0u;
      PMX_TRY(pmx_xpass(p, op, p->specB, p->specC, s));
      PMX_TRY(pmx_inverse_from_C(p, 3, p->specC, p->fmesh, s));
      PMX_TRY(pmx_launch_delta2(p->fmesh, p->rho, p->ncell, off, s));
    }
    PMX_TRY(force_mesh_from(p, p->rho, 1.0f, 0u, p->fmesh, s));
    PMX_TRY(pmx_launch_lpt_apply(p, p->fmesh, dx, pv, npx, npy, npz, g, 1, 0, s));
  }
  PMX_TRY(force_mesh_from(p, ic, 1.0f, 0u, p->fmesh, s));
  return pmx_launch_lpt_apply(p, p->fmesh, dx, pv, npx, npy, npz, g, 0, order == 2 ? 1 : 0, s);
}

int pmx_lpt1_vjp(pmx_plan p, const float* dx_bar, const float* p_bar, int npx, int npy, int npz, float growth,
                 float vel_factor, float* ic_bar, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, dx_bar ? (const void*)dx_bar : (const void*)p_bar, npx, npy, npz, 0));
  if (!ic_bar) return pmx_set_error(1, "null output");
  hipStream_t s = (hipStream_t)stream;
  PMX_TRY(pmx_launch_lpt1_vjp(p, dx_bar, p_bar, npx, npy, npz, growth, vel_factor, p->fmesh, s));
  return pmx_force_mesh_vjp(p, p->fmesh, 1.0f, 0u, ic_bar, stream);
}

namespace {
int check_shell(pmx_plan_t* p, const float* pos, int npx, int npy, int npz, int mode, const double* box,
                const double* observer, double rmin, double rmax, int nside, int scheme, int ntiles, const int* rot,
                const double* shift) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!box || !observer) return pmx_set_error(1, "null argument");
  if (nside < 1 || nside > 8192) return pmx_set_error(1, "nside out of range");
  if (scheme < 0 || scheme > 2) return pmx_set_error(2, "scheme must be 0 (ngp), 1 (bilinear) or 2 (rbf_neighbor)");
  if (!(rmax > rmin) || rmin < 0) return pmx_set_error(1, "need 0 <= rmin < rmax");
  if (rot && (ntiles < 1 || ntiles > pmx_sky_max_tiles() || !shift))
    return pmx_set_error(1, "image table: need 1 <= ntiles <= " + std::to_string(pmx_sky_max_tiles()) +
                                " and a shift table");
  return 0;
}
}  // namespace

int pmx_paint_healpix(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                      float weight, const double box[3], const double observer[3], double rmin, double rmax, int nside,
                      int scheme, float* map, int accumulate, void* stream) {
  return pmx_paint_healpix_tiled(p, pos, npx, npy, npz, mode, weights, weight, box, observer, rmin, rmax, nside, scheme,
                                 1, nullptr, nullptr, map, accumulate, stream);
}

int pmx_paint_healpix_tiled(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                            float weight, const double box[3], const double observer[3], double rmin, double rmax,
                            int nside, int scheme, int ntiles, const int* tile_rot, const double* tile_shift,
                            float* map, int accumulate, void* stream) {
  PMX_TRY(check_shell(p, pos, npx, npy, npz, mode, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                      tile_shift));
  if (!map) return pmx_set_error(1, "null map");
  PMX_TRY(ensure_acc(p, (size_t)12 * nside * nside + 4 * (size_t)nside));  // maps + ring table
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  return pmx_launch_healpix(p, a, weights, weight, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                            tile_shift, map, accumulate, (hipStream_t)stream);
}

// ------------------------------------------------------------ HEALPix shell paint without allocation
// The map accumulator (12 nside^2 float64 values) and the ring colatitude table (4 nside) of the
// shell paint live in one file-scope device array per GPU. Host code never passes `scratch` itself
// (its host-side name is not a device address): each device's address comes from
// hipGetSymbolAddress on the entry point's first call on that device.
namespace {
constexpr int kStaticMaxNside = 512;
constexpr size_t kStaticScratch = (size_t)12 * kStaticMaxNside * kStaticMaxNside + 4 * (size_t)kStaticMaxNside;
constexpr int kStaticMaxDevices = 64;
__device__ __align__(256) double scratch[kStaticScratch];
std::mutex g_scratch_mu;
double* g_scratch_ptr[kStaticMaxDevices] = {};

int static_scratch(double** out) {
  int dev = -1;
  PMX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kStaticMaxDevices) return pmx_set_error(2, "device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return pmx_set_error(3, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

int pmx_paint_healpix_noalloc(int nx, int ny, int nz, const float* pos, int npx, int npy, int npz, int mode,
                              const float* weights, float weight, const double box[3], const double observer[3],
                              double rmin, double rmax, int nside, int scheme, int ntiles, const int* tile_rot,
                              const double* tile_shift, float* map, int accumulate, void* stream) {
  if (!pmx_fft_len_ok(nx) || !pmx_fft_len_ok(ny) || !pmx_fft_len_ok(nz))
    return pmx_set_error(2, "unsupported mesh size (each axis must be one of 8,16,32,64,96,128,192,256,512,1024)");
  if (nside < 1 || nside > kStaticMaxNside)
    return pmx_set_error(2, "pmx_paint_healpix_noalloc: nside must be in [1, " + std::to_string(kStaticMaxNside) +
                                "] (its static scratch holds " + std::to_string(kStaticScratch) + " doubles)");
  double* ptr = nullptr;
  PMX_TRY(static_scratch(&ptr));
  // a plan-shaped view of the mesh geometry and the scratch: nothing here owns device memory
  pmx_plan_t v;
  std::memset(&v, 0, sizeof(v));
  v.nslots = 1;
  v.nx = nx;
  v.ny = ny;
  v.nz = nz;
  v.nzh = nz / 2 + 1;
  v.ncell = (size_t)nx * ny * nz;
  PMX_HIP(hipGetDevice(&v.device));
  v.acc = ptr;
  v.acc_bytes = kStaticScratch * sizeof(double);
  PMX_TRY(check_shell(&v, pos, npx, npy, npz, mode, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                      tile_shift));
  if (!map) return pmx_set_error(1, "null map");
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, &v);
  return pmx_launch_healpix(&v, a, weights, weight, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                            tile_shift, map, accumulate, (hipStream_t)stream);
}

int pmx_paint_healpix_lightcone(pmx_plan p, const float* pos, const float* vel, int npx, int npy, int npz, int mode,
                                const float* weights, float weight, const double box[3], const double observer[3],
                                double rmin, double rmax, int nside, int scheme, int ntiles, const int* tile_rot,
                                const double* tile_shift, const double* drift_tables, int ntab, double chi_max,
                                double a_lo, double a_hi, double a_now, double gp_now, double drift_bound, float* map,
                                int accumulate, void* stream) {
  PMX_TRY(check_shell(p, pos, npx, npy, npz, mode, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                      tile_shift));
  if (!map || !vel || !drift_tables) return pmx_set_error(1, "null argument");
  if (ntab < 2 || !(chi_max > 0.0) || !(a_hi > a_lo) || !(a_now > 0.0) || !(drift_bound >= 0.0))
    return pmx_set_error(1, "bad drift tables: need ntab >= 2, chi_max > 0, a_hi > a_lo, a_now > 0");
  PMX_TRY(ensure_acc(p, (size_t)12 * nside * nside + 4 * (size_t)nside));  // maps + ring table
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  const PmxDrift d{vel, drift_tables, ntab, chi_max, a_lo, a_hi, a_now, gp_now, drift_bound};
  return pmx_launch_healpix(p, a, weights, weight, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                            tile_shift, map, accumulate, (hipStream_t)stream, &d);
}

int pmx_paint_healpix_vjp(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                          float weight, const double box[3], const double observer[3], double rmin, double rmax,
                          int nside, int scheme, int ntiles, const int* tile_rot, const double* tile_shift,
                          const float* map_bar, float* pos_bar, int accumulate, void* stream) {
  if (scheme == 2) return pmx_set_error(2, "the rbf_neighbor scheme has no adjoint");
  PMX_TRY(check_shell(p, pos, npx, npy, npz, mode, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                      tile_shift));
  if (!map_bar || !pos_bar) return pmx_set_error(1, "null cotangent");
  PMX_TRY(ensure_acc(p, (size_t)12 * nside * nside + 4 * (size_t)nside));  // the ring table after the maps
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  return pmx_launch_healpix_vjp(p, a, weights, weight, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                                tile_shift, map_bar, pos_bar, accumulate, (hipStream_t)stream);
}

int pmx_paint_healpix_lightcone_vjp(pmx_plan p, const float* pos, const float* vel, int npx, int npy, int npz,
                                    int mode, const float* weights, float weight, const double box[3],
                                    const double observer[3], double rmin, double rmax, int nside, int scheme,
                                    int ntiles, const int* tile_rot, const double* tile_shift,
                                    const double* drift_tables, int ntab, double chi_max, double a_lo, double a_hi,
                                    double a_now, double gp_now, double drift_bound, const float* map_bar,
                                    float* pos_bar, float* vel_bar, int accumulate, int nparam,
                                    const double* param_tables, double* param_dots, void* stream) {
  if (scheme == 2) return pmx_set_error(2, "the rbf_neighbor scheme has no adjoint");
  PMX_TRY(check_shell(p, pos, npx, npy, npz, mode, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                      tile_shift));
  if (!map_bar || !pos_bar || !vel_bar) return pmx_set_error(1, "null cotangent");
  if (!vel || !drift_tables) return pmx_set_error(1, "null argument");
  if (ntab < 2 || !(chi_max > 0.0) || !(a_hi > a_lo) || !(a_now > 0.0) || !(drift_bound >= 0.0))
    return pmx_set_error(1, "bad drift tables: need ntab >= 2, chi_max > 0, a_hi > a_lo, a_now > 0");
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  if (nparam < 0 || nparam > 4 || (nparam > 0 && (!param_tables || !param_dots)))
    return pmx_set_error(1, "need 0 <= nparam <= 4 with parameter tables and dots");
  PMX_TRY(ensure_acc(p, (size_t)12 * nside * nside + 4 * (size_t)nside));  // the ring table after the maps
  PmxDrift d{vel, drift_tables, ntab, chi_max, a_lo, a_hi, a_now, gp_now, drift_bound};
  d.ptab = param_tables;
  d.nparam = nparam;
  d.pdots = param_dots;
  return pmx_launch_healpix_vjp(p, a, weights, weight, box, observer, rmin, rmax, nside, scheme, ntiles, tile_rot,
                                tile_shift, map_bar, pos_bar, accumulate, (hipStream_t)stream, &d, vel_bar);
}

namespace {
int check_born(pmx_plan_t* p, const float* maps, int nshell, int64_t npix, const double* coef, int nsrc) {
  PMX_TRY(check_plan(p));
  if (!maps || !coef) return pmx_set_error(1, "null buffer");
  if (nshell < 1 || nshell > pmx_born_max_shells())
    return pmx_set_error(1, "need 1 <= nshell <= " + std::to_string(pmx_born_max_shells()));
  if (nsrc < 1 || nsrc > pmx_born_max_sources())
    return pmx_set_error(1, "need 1 <= nsrc <= " + std::to_string(pmx_born_max_sources()));
  if (npix < 1) return pmx_set_error(1, "npix must be positive");
  return 0;
}
}  // namespace

int pmx_born(pmx_plan p, const float* maps, int nshell, int64_t npix, const double* coef, int nsrc, float* kappa,
             int accumulate, void* stream) {
  PMX_TRY(check_born(p, maps, nshell, npix, coef, nsrc));
  if (!kappa) return pmx_set_error(1, "null output");
  return pmx_launch_born(p, maps, nshell, (long)npix, coef, nsrc, kappa, accumulate, (hipStream_t)stream);
}

int pmx_born_vjp(pmx_plan p, const float* maps, int nshell, int64_t npix, const double* coef, int nsrc,
                 const float* kappa_bar, float* maps_bar, int accumulate, double* dots, void* stream) {
  PMX_TRY(check_born(p, maps, nshell, npix, coef, nsrc));
  if (!kappa_bar || !maps_bar) return pmx_set_error(1, "null cotangent");
  return pmx_launch_born_vjp(p, maps, nshell, (long)npix, coef, nsrc, kappa_bar, maps_bar, accumulate, dots,
                             (hipStream_t)stream);
}

int pmx_born_flat(pmx_plan p, const float* maps, int nshell, int ny, int nx, const double* coef, int nsrc,
                  const double* scales, float* kappa, int accumulate, void* stream) {
  PMX_TRY(check_born(p, maps, nshell, (int64_t)ny * nx, coef, nsrc));
  if (ny < 1 || nx < 1) return pmx_set_error(1, "need ny, nx >= 1");
  if (!kappa || !scales) return pmx_set_error(1, "null buffer");
  return pmx_launch_born_flat(p, maps, nshell, ny, nx, coef, nsrc, scales, kappa, accumulate, (hipStream_t)stream);
}

int pmx_born_flat_vjp(pmx_plan p, const float* maps, int nshell, int ny, int nx, const double* coef, int nsrc,
                      const double* scales, const float* kappa_bar, float* maps_bar, int accumulate, double* dots,
                      void* stream) {
  PMX_TRY(check_born(p, maps, nshell, (int64_t)ny * nx, coef, nsrc));
  if (ny < 1 || nx < 1) return pmx_set_error(1, "need ny, nx >= 1");
  if (!kappa_bar || !maps_bar || !scales) return pmx_set_error(1, "null buffer");
  return pmx_launch_born_flat_vjp(p, maps, nshell, ny, nx, coef, nsrc, scales, kappa_bar, maps_bar, accumulate, dots,
                                  (hipStream_t)stream);
}

int pmx_paint_flat_vjp(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                       float weight, int fnx, int fny, float zlo, float zhi, double scale, const float* map_bar,
                       float* pos_bar, int accumulate, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!map_bar || !pos_bar || fnx < 1 || fny < 1) return pmx_set_error(1, "bad flat-sky arguments");
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  return pmx_launch_flat_vjp(p, a, weights, weight, fnx, fny, zlo, zhi, map_bar, (float)scale, pos_bar, accumulate,
                             (hipStream_t)stream);
}

int pmx_image_positions(pmx_plan p, const float* pos, const float* vel, int npx, int npy, int npz, int mode,
                        const double box[3], const double observer[3], int rot, const double shift[3],
                        const double* drift_tables, int ntab, double chi_max, double a_lo, double a_hi, double a_now,
                        double gp_now, float* out, void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!out || !box || !observer || !shift) return pmx_set_error(1, "null argument");
  if (rot < 0 || rot >= 48) return pmx_set_error(1, "rotation code must be in [0, 48)");
  for (int c = 0; c < 3; ++c)
    if (!(box[c] > 0.0)) return pmx_set_error(1, "box sizes must be positive");
  if ((size_t)npx * npy * npz * 3 >= ((size_t)1 << 32)) return pmx_set_error(1, "too many particles");
  if (!vel) return pmx_launch_image_positions(p, part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p),
                                              box, observer, rot, shift, out, (hipStream_t)stream, nullptr);
  if (!drift_tables || ntab < 2 || !(chi_max > 0.0) || !(a_hi > a_lo) || !(a_now > 0.0))
    return pmx_set_error(1, "bad drift tables: need ntab >= 2, chi_max > 0, a_hi > a_lo, a_now > 0");
  const PmxDrift d{vel, drift_tables, ntab, chi_max, a_lo, a_hi, a_now, gp_now, 0.0};
  return pmx_launch_image_positions(p, part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p), box,
                                    observer, rot, shift, out, (hipStream_t)stream, &d);
}

int pmx_image_positions_vjp(pmx_plan p, const float* pos, const float* vel, int npx, int npy, int npz, int mode,
                            const double box[3], const double observer[3], int rot, const double shift[3],
                            const double* drift_tables, int ntab, double chi_max, double a_lo, double a_hi,
                            double a_now, double gp_now, const float* out_bar, float* pos_bar, float* vel_bar,
                            int accumulate, int nparam, const double* param_tables, double* param_dots,
                            void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!out_bar || !pos_bar || !box || !observer || !shift) return pmx_set_error(1, "null argument");
  if (rot < 0 || rot >= 48) return pmx_set_error(1, "rotation code must be in [0, 48)");
  for (int c = 0; c < 3; ++c)
    if (!(box[c] > 0.0)) return pmx_set_error(1, "box sizes must be positive");
  if ((size_t)npx * npy * npz * 3 >= ((size_t)1 << 32)) return pmx_set_error(1, "too many particles");
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  if (!vel)
    return pmx_launch_image_positions_vjp(p, a, box, observer, rot, shift, out_bar, pos_bar, nullptr, accumulate,
                                          (hipStream_t)stream, nullptr);
  if (!vel_bar) return pmx_set_error(1, "a drifted image needs vel_bar");
  if (!drift_tables || ntab < 2 || !(chi_max > 0.0) || !(a_hi > a_lo) || !(a_now > 0.0))
    return pmx_set_error(1, "bad drift tables: need ntab >= 2, chi_max > 0, a_hi > a_lo, a_now > 0");
  if (nparam < 0 || nparam > 4 || (nparam > 0 && (!param_tables || !param_dots)))
    return pmx_set_error(1, "need 0 <= nparam <= 4 with parameter tables and dots");
  PmxDrift d{vel, drift_tables, ntab, chi_max, a_lo, a_hi, a_now, gp_now, 0.0};
  d.ptab = param_tables;
  d.nparam = nparam;
  d.pdots = param_dots;
  return pmx_launch_image_positions_vjp(p, a, box, observer, rot, shift, out_bar, pos_bar, vel_bar, accumulate,
                                        (hipStream_t)stream, &d);
}

int pmx_paint_flat(pmx_plan p, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                   float weight, int fnx, int fny, float zlo, float zhi, double scale, float* map, int accumulate,
                   void* stream) {
  PMX_TRY(check_plan(p));
  PMX_TRY(check_particles(p, pos, npx, npy, npz, mode));
  if (!map || fnx < 1 || fny < 1) return pmx_set_error(1, "bad flat-sky arguments");
  const size_t n = (size_t)fnx * fny;
  PMX_TRY(ensure_acc(p, n));
  hipStream_t s = (hipStream_t)stream;
  float* acc = reinterpret_cast<float*>(p->acc);
  PMX_TRY(pmx_launch_fill(acc, n, 0.f, s));
  PartArgs a = part_args(const_cast<float*>(pos), nullptr, npx, npy, npz, mode, p);
  PMX_TRY(pmx_launch_flat(p, a, weights, weight, fnx, fny, zlo, zhi, acc, s));
  return pmx_launch_finalize_map(acc, map, (long)n, scale, accumulate, s);
}

}  // extern "C"
