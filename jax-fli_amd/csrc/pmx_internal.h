// Internal declarations shared by the pmx translation units.
//
// Layout conventions (all row-major, FP32 unless noted):
//   real mesh      R[x][y][z]                (nx, ny, nz)
//   particles      P[i][j][k][3]             (npx, npy, npz, 3)  -- the reference's (X, Y, Z, 3)
//   force mesh     F[x][y][z][3]             interleaved so one readout corner is 12 contiguous bytes
//   spectra        half spectrum along z, nzh = nz/2 + 1 complex bins, stored in pass-specific
//                  transposed orders (see the pass kernels in pmx_fft.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <string>
#include <vector>

// per-step timing marks of pmx_step (paint | r2c_z | c2c_y | xpass | c2c_y_inv | c2r_z | grad4 | readout)
constexpr int kStepMarks = 9;
struct StepMarks {
  hipEvent_t e[kStepMarks];
};

// Slot batching. A batched plan holds nslots copies of its work buffers (and of the particles it
// steps), each slot's `stride` bytes after the previous one's, and shares its FFT and k-space
// tables. Its launches add the slot to the grid (blockIdx.z, times any z extent the kernel already
// uses) and every kernel moves its slot-owned pointers by slot * stride. Device arrays, per slot:
// live (nonzero: the slot runs this launch; null: every slot runs) and coef (4 floats: unused, kick,
// drift after the kick, potential scale; null: the launch arguments hold the coefficients).
struct Slots {
  size_t stride = 0;
  const int* live = nullptr;
  const float* coef = nullptr;
};
template <class T>
__host__ __device__ __forceinline__ T* at_slot(T* p, size_t stride, unsigned z) {
  return p ? reinterpret_cast<T*>(reinterpret_cast<char*>(const_cast<typename std::remove_const<T>::type*>(p)) +
                                  (size_t)z * stride)
           : p;
}
__device__ __forceinline__ bool slot_live(const Slots& s, unsigned z) { return !s.live || s.live[z] != 0; }

struct pmx_plan_t {
  int nx, ny, nz, nzh;
  int nslots;       // slots of a batched plan (1: an ordinary plan)
  Slots slots;      // the batched plan's slot stride and per-slot control arrays
  void* arena;      // a batched plan's slot buffers (one allocation); null for an ordinary plan
  int device;
  unsigned flags;
  size_t ncell;   // nx*ny*nz
  size_t nspec;   // nzh*nx*ny (complex)
  int nzhp;       // row length of the D layout: nzh rounded up to 8 complex (64-byte aligned rows)
  size_t nspecd;  // nzhp*nx*ny (complex), one component of D
  // twiddle tables exp(-2 pi i k / n), one per line length that the passes use
  float2 *tw_x, *tw_y, *tw_z, *tw_zh;
  // per-axis spectral tables, index = FFT bin along that axis (z: 0..nz/2)
  float *grad_x, *grad_y, *grad_z;  // FastPM 4-point gradient (8 sin w - sin 2w)/6, w = 2 pi fftfreq
  float *w2_x, *w2_y, *w2_z;        // w^2
  float *dec_x, *dec_y, *dec_z;     // CIC window compensation sinc(w/2)^-4
  // work buffers owned by the plan
  float* rho;       // density mesh
  float2* specA;    // [kz][x][y]
  float2* specB;    // [kz][ky][x]
  float2* specC;    // 3 x [kz][x][ky]
  float2* specD;    // 3 x [x][y][kz]
  float* fmesh;     // force mesh [x][y][z][3]
  double *kphys_x, *kphys_y, *kphys_z;  // estimator wavevectors (h/Mpc) per bin, set by pmx_plan_set_box
  bool box_set;
  double rbf_sigma = 0.0;  // rbf_neighbor shell paints: Gaussian width in radians (0: the pixel resolution as FWHM)
  // sorted paint (pmx_particles.hip): per paint tile particle counts (zero between paints), the
  // last sort's counts, segment offsets, reservation cursors; per particle sub-block window counts; segments of particle
  // indices; misc = {max |weight| bits, far entries, far entries of the last paint}; fixed-point
  // scale {2^k, 2^-k}
  unsigned *ps_count, *ps_tcount, *ps_offset, *ps_cursor, *ps_wgcnt, *ps_seg, *ps_misc;
  unsigned* ps_pack;  // per particle: window slot << 16 | rank within the sub-block's slot (count -> scatter)
  float* ps_scale;
  float* ps_faces;  // per tile: its upper face cells (pmx_paint_face_cells floats), 3 blocks for vectors
  double* red;      // per-workgroup partial sums of the adjoint dot products (kRedMax doubles)
  size_t ps_wg_cap, ps_seg_cap;
  double* acc;      // double accumulators (healpix maps, flat maps); grown on demand
  size_t acc_bytes;
  size_t workspace_bytes;
  int timing;
  std::vector<StepMarks>* marks;
  size_t marks_used;
};

constexpr int kRedMax = 65536;  // partial-sum slots: the largest particle-kernel grid

// batched plans of the request server (pmx_api.hip)
int pmx_plan_create_batched(pmx_plan_t** out, int n, int nslots, size_t extra, size_t* extra_off);
size_t pmx_batched_slot_bytes(int n, size_t extra);  // the slot stride pmx_plan_create_batched would use
pmx_plan_t pmx_slot_view(const pmx_plan_t* b, int z);
int pmx_pot_force_mesh(pmx_plan_t* p, const float* src, float force_scale, hipStream_t s);

// error reporting (pmx_api.hip)
int pmx_set_error(int code, const std::string& msg);
int pmx_check_launch(const char* what);

// supported line lengths for the in-LDS Stockham FFT
bool pmx_fft_len_ok(int n);

// k-space operators of the x pass: forward only, inverse only, 3 force components (Green x FastPM
// gradient), colouring by sqrt(P(k)), power-spectrum binning, potential delta/k^2
enum { XP_FWD = 0, XP_INV = 1, XP_FORCES = 2, XP_COLOR = 3, XP_POWER = 4, XP_POT = 5, XP_HESS = 6 };
// XP_HESS: three components of the Hessian of the potential, H_ij = ifft(g_i g_j fft(src) / k^2):
// the diagonal (00, 11, 22), or with this flag the off-diagonal (01, 02, 12)
constexpr unsigned kHessOffDiag = 8u;
// LPT term at the lattice (pmx_particles.hip): (D, c) = columns (2 term, 2 term + 1) of coef; ntab == 1:
// constants, else a table over the particle's comoving distance r on the uniform grid [0, r_max]
struct LptGeom {
  const double* coef;
  int ntab;
  double r_max, cell[3], obs[3];
};

// FFT passes (pmx_fft.hip). All asynchronous on `stream`. Return 0 or an error code.
struct KOp {
  int mode;               // see enum in pmx_fft.hip
  float scale;            // multiplies the spectrum (already divided by the cell count where relevant)
  unsigned flags;         // bit0: CIC deconvolution; bit1 (XP_POT): PGD range filter; bit2 (XP_POT):
                          // mix_a + mix_b * PGD range filter
  float pgd_kl2, pgd_ks4; // PGD filter exp(-kl^2 / k^2) exp(-k^4 / ks^4), k in cell units
  float mix_a, mix_b;     // force and sharpening weights of one combined potential
  // colouring: amplitude = sqrt(P(|k|) * norm), P linearly interpolated in (ktab, ptab)
  const double* kphys_x;  // |k| components in h/Mpc per bin index, as float64
  const double* kphys_y;
  const double* kphys_z;
  const double* ktab;
  const double* ptab;
  int ntab;
  double cnorm;
  // power spectrum binning: per pole, sums of Re and Im of m1 conj(m2) (2l+1) P_l(mu) over the
  // nedges+1 bins of |k|; the auto spectrum has m2 = m1 and no Im sums
  const double* kedges;
  int nedges;
  double* pk_sum;         // npoles x (nedges+1) real sums
  double* pk_im;          // npoles x (nedges+1) imaginary sums, cross spectra only (else null)
  double* k_sum;          // nedges+1 (accumulated only when acc_k)
  double* k_cnt;
  const float2* spec2;    // second mesh's full spectrum [kz][ky][kx] (cross spectrum) or null
  int npoles;             // 1..kMaxPolesLaunch
  int poles[8];
  double los[3];          // unit line of sight; mu = k.los / |k| (0 at k = 0)
  int acc_k;
};
constexpr int kMaxPolesLaunch = 8;

int pmx_fft_forward_to_B(pmx_plan_t* p, const float* in, hipStream_t s);                 // R -> specB
int pmx_xpass(pmx_plan_t* p, const KOp& op, const float2* inB, float2* out, hipStream_t s);  // specB -> out
// poles one XP_POWER launch can bin for nedges edges on this plan (0: too many edges)
int pmx_xpass_power_poles(const pmx_plan_t* p, int nedges);
int pmx_inverse_from_C(pmx_plan_t* p, int ncomp, const float2* inC, float* out, hipStream_t s);  // specC -> real
int pmx_pass_r2c_z(pmx_plan_t* p, const float* in, float2* outA, hipStream_t s);
int pmx_pass_c2c_y(pmx_plan_t* p, const float2* inA, float2* outB, hipStream_t s);
int pmx_pass_c2c_y_inv(pmx_plan_t* p, int ncomp, const float2* inC, float2* outD, hipStream_t s);
int pmx_pass_c2r_z(pmx_plan_t* p, int ncomp, const float2* inD, float* out, hipStream_t s);
// force mesh (x,y,z,3) from the potential by the FastPM 4-point stencil
int pmx_grad4(pmx_plan_t* p, const float* phi, float* F, hipStream_t s);

// particle kernels (pmx_particles.hip)
struct PartArgs {
  float* x;            // positions/displacements (npx,npy,npz,3)
  float* v;            // velocities (may be null)
  int npx, npy, npz;
  int mode;            // 0 relative (lattice + dx), 1 absolute
  float sx, sy, sz;    // lattice spacing in cells (mesh / particle grid)
};
// drift to the lightcone of the HEALPix shell paint (pmx_sky.hip)
struct PmxDrift {
  const float* v;
  const double* tables;  // a(chi) on [0, chi_max]; Gp(a), dGfa(a) on [a_lo, a_hi]; ntab points each
  int ntab;
  double chi_max, a_lo, a_hi, a_now, gp_now, fbound;
  // adjoint only: per-parameter tables of df/dtheta on the a(chi) grid, and the dots they feed
  const double* ptab = nullptr;
  int nparam = 0;
  double* pdots = nullptr;
};
// paint tiles of a plan (log2 of the tile dims, tile count)
void pmx_paint_tiles(const pmx_plan_t* p, int& sx, int& sy, int& sz, int& ntiles);
int pmx_paint_face_cells(const pmx_plan_t* p);
int pmx_launch_paint(pmx_plan_t* p, const PartArgs& a, const float* w, float w0, float cd_pre,
                     float* mesh, int accumulate, int atomic, hipStream_t s);
int pmx_launch_fill(float* a, size_t n, float v, hipStream_t s);
// the sorted paint in two halves: sort (count, scan, scatter) then deposit(s) of the sorted particles
int pmx_paint_sort(pmx_plan_t* p, const PartArgs& a, float cd_pre, const float* w, int wstride, float w0,
                   hipStream_t s);
size_t pmx_sort_wg_entries(int npx, int npy, int npz);
// size the sort buffers for a particle grid (synchronous, grow only; adds the new bytes to grown)
int pmx_sort_reserve(pmx_plan_t* p, int npx, int npy, int npz, size_t& grown);
// ncomp > 1: component c carries w[p * wstride + c] and goes to mesh + c * mcomp
int pmx_paint_deposit(pmx_plan_t* p, const PartArgs& a, const float* w, int wstride, float wmul, float w0,
                      int is_signed, float* mesh, int mstride, int accumulate, hipStream_t s, int ncomp = 1,
                      size_t mcomp = 0);
// adjoint building blocks (pmx_particles.hip)
int pmx_launch_absmax(const float* a, size_t n, unsigned* bits, hipStream_t s);
int pmx_launch_set_scale(unsigned* bits, float mul, float* scale, hipStream_t s);
int pmx_launch_div4(pmx_plan_t* p, const float* U0, const float* U1, const float* U2, int ustride, float* out,
                    hipStream_t s);
int pmx_launch_vjp_gather(pmx_plan_t* p, const PartArgs& a, int ncomp, const float* u, float umul, const float* M,
                          const float* rhob, const float* w, float w0, float* xbar, int accumulate, float* wbar,
                          double* dot, float* vkick, float ckick, hipStream_t s);

int pmx_launch_axpy(float* out, const float* x, const float* v, float c, size_t n, hipStream_t s);
int pmx_launch_adj_prep(pmx_plan_t* p, const float* xbar, float* vbar, const float* v, size_t n, float cd,
                        double* dot, unsigned* bits, hipStream_t s);
int pmx_launch_fill_d(double* a, size_t n, hipStream_t s);
int pmx_launch_readout(pmx_plan_t* p, const PartArgs& a, const float* mesh, int ncomp, float* out,
                       float ck, float cd_post, int what, hipStream_t s);

