/* pmx: MI355X-native particle-mesh engine behind a plain C ABI.
 *
 * Every entry point takes raw device pointers and a hipStream_t (passed as void*), is asynchronous
 * on that stream unless stated otherwise, returns 0 on success or a nonzero code (1 invalid
 * argument, 2 unsupported size or mode, 3 HIP runtime error, 4 out of device memory), and leaves
 * a message for pmx_last_error() on failure.
 *
 * Layouts follow the reference's arrays, row-major float32:
 *   particle field  (npx, npy, npz, 3)  displacements from the lattice (mode 0, GRID_RELATIVE) or
 *                                       absolute grid coordinates (mode 1, GRID_ABSOLUTE)
 *   density mesh    (nx, ny, nz)
 *   force mesh      (nx, ny, nz, 3)
 *   HEALPix map     (12 nside^2,)  RING ordering
 * In mode 0 the particle grid must divide the mesh (mesh/np integer per axis).
 * Supported mesh sizes per axis: 8, 16, 32, 64, 96, 128, 192, 256, 512, 1024.
 * A plan owns its work buffers: calls sharing one plan must be issued on one stream.
 */
#ifndef PMX_H
#define PMX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmx_plan_t* pmx_plan;

#define PMX_DECONV 1u      /* divide the force spectrum by the CIC window of paint and readout */
#define PMX_DRIFT_PRE 2u   /* pmx_step: x += v * c_drift_pre before painting */
#define PMX_DRIFT_POST 4u  /* pmx_step: x += v * c_drift_post after the kick */
#define PMX_PAINT_ATOMIC 8u /* paint with the one-pass global-atomic scatter instead of the sorted
                               owner-computes tiles (plan flag or per-call flag) */
#define PMX_PLAN_NO_RESERVE 16u /* pmx_plan_create: reserve no paint-sort buffers (pmx_plan_reserve sizes them) */

int pmx_abi_version(void);
const char* pmx_last_error(void);
int pmx_fft_supported(int n);
/* Workgroup cap of the grid-stride kernels (fills, noise, LPT, image transforms, map finalisation,
 * Born): 0 restores the default 2^20; smaller values make every such kernel loop over its range
 * (a check that large inputs are covered). Process-wide; not a reference interface. */
int pmx_set_grid_cap(unsigned max_blocks);

/* Plan for one mesh shape; allocates its workspace (synchronous): FFT tables and spectra, density
 * and force meshes, and, unless flags has PMX_PLAN_NO_RESERVE, the paint-sort buffers for one
 * particle per cell (an (nx, ny, nz) particle grid). */
int pmx_plan_create(pmx_plan* plan, int nx, int ny, int nz, unsigned flags);
int pmx_plan_destroy(pmx_plan plan);
/* Bytes of device memory the plan holds (everything pmx_plan_create and pmx_plan_reserve allocated). */
size_t pmx_plan_workspace_bytes(pmx_plan plan);
/* Grow the plan's buffers so that later calls allocate nothing (synchronous; never shrinks):
 *   npx, npy, npz > 0: paint-sort buffers for that particle grid (and any grid with no more particles
 *                      and particle sub-blocks of 8 x 8 x 64);
 *   acc_count > 0:     the float64 map accumulator of the HEALPix and flat-sky paints, which need
 *                      12 nside^2 + 4 nside values (HEALPix) or fnx * fny (flat).
 * Entry points never allocate: a paint whose grid or map exceeds the reservation fails with code 4
 * and names the size to reserve. No reference counterpart (XLA allocates per call). */
int pmx_plan_reserve(pmx_plan plan, int npx, int npy, int npz, size_t acc_count);
/* Entries of the last sorted paint on this plan that fell outside their particle sub-block's
 * 5x5x8-tile window and were counted and placed one by one with global atomics (particles displaced
 * by more than about two tiles). Synchronises the device. */
int pmx_plan_escapes(pmx_plan plan, uint64_t* count);
/* Per-kernel timing of pmx_step on this plan. enable > 0 clears previous records, creates the HIP
 * events for max(enable, 64) steps up front (timed steps create none) and brackets every following
 * pmx_step with them on its stream; enable = 0 stops; pmx_plan_timing_read synchronises on the last
 * event and returns the summed milliseconds of the 8 phases (paint, r2c_z, c2c_y, k-space x pass,
 * inverse c2c_y, c2r_z, gradient stencil, readout+kick+drift) and the number of steps recorded. */
int pmx_plan_timing(pmx_plan plan, int enable);
int pmx_plan_timing_read(pmx_plan plan, double* ms8, int* nsteps);
/* Physical box (Mpc/h) for the k tables of pmx_power and pmx_color_field (synchronous upload). */
int pmx_plan_set_box(pmx_plan plan, const double box[3]);

/* CIC paint. Replaces cic_paint_dx / cic_paint as called by _single_paint
 * (src/jax_fli/_src/fields/_painting.py:66-82; ParticleField.paint, fields/particles.py:120-178).
 * weights: per-particle device array or NULL for the scalar `weight`. accumulate=0 zeroes mesh. */
int pmx_paint_cic(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                  float weight, float* mesh, int accumulate, void* stream);

/* CIC readout of an ncomp (1 or 3) component mesh (ncomp interleaved last). Replaces
 * cic_read_dx / cic_read in ParticleField.read_out (fields/particles.py:183-229). out: (np, ncomp). */
int pmx_readout_cic(pmx_plan plan, const float* mesh, int ncomp, const float* pos, int npx, int npy, int npz,
                    int mode, float* out, void* stream);

/* Unnormalised 3D real FFT pair (fft3d / ifft3d of the reference, initial.py:175,182 and inside
 * pm_forces). Spectrum layout: complex64 [nz/2+1][ny][nx] (kz slowest, kx fastest). */
int pmx_fft_r2c(pmx_plan plan, const float* in, void* spec, void* stream);
int pmx_fft_c2r(pmx_plan plan, const void* spec, float* out, void* stream);

/* Force mesh (nx,ny,nz,3) of a density mesh: FFT, Green's function -1/k^2, FastPM 4-point gradient,
 * optional CIC compensation (PMX_DECONV), inverse FFT; times `scale`. The k-space part of
 * jaxpm.pm.pm_forces as called by _forces (src/jax_fli/pm/solvers.py:19-32). fmesh may be NULL
 * (result kept in the plan). */
int pmx_force_mesh(pmx_plan plan, const float* density, float scale, unsigned flags, float* fmesh, void* stream);

/* Particle forces pm_forces(pos) * scale: paint, potential (one inverse FFT), 4-point gradient
 * stencil, readout. Replaces _forces
 * (src/jax_fli/pm/solvers.py:19-32) with scale = 1.5 * Omega_m. forces: (npx,npy,npz,3). */
int pmx_forces(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, float scale, unsigned flags,
               float* forces, void* stream);

/* Fused PM step on (x, v) in place (force as in pmx_forces):
 *   [PMX_DRIFT_PRE]  x += v * c_drift_pre
 *   v += (pm_forces(x) * force_scale) * c_kick
 *   [PMX_DRIFT_POST] x += v * c_drift_post
 * The drift/kick/drift bodies of EfficientDriftDoubleKick.step (solvers.py:151-202, with the
 * drift of the next step fused after the kick), ReversibleDoubleKickDrift.step (solvers.py:258-325)
 * and its exact inverse (in place of ReversibleDoubleKickDrift.reverse, solvers.py:339-411).
 * Coefficients are the host-computed scalar factors of those methods. */
int pmx_step(pmx_plan plan, float* x, float* v, int npx, int npy, int npz, int mode, float force_scale,
             float c_drift_pre, float c_kick, float c_drift_post, unsigned flags, void* stream);

/* pmx_step, with the step's force mesh written to force_mesh (nx*ny*nz*3 floats, caller-owned,
 * interleaved (x, y, z, 3)) instead of the plan's work buffer, where it stays for
 * pmx_step_reverse_vjp_saved. Same bodies as pmx_step (solvers.py:151-202, 258-325). */
int pmx_step_keep(pmx_plan plan, float* x, float* v, int npx, int npy, int npz, int mode, float force_scale,
                  float c_drift_pre, float c_kick, float c_drift_post, unsigned flags, float* force_mesh, void* stream);

/* Binned power spectrum sums of a real mesh (the estimator of _power,
 * src/jax_fli/_src/power/_compute.py:68-119): for bin b = #(kedges <= |k|), b in [0, nedges],
 * pk_sum[b] = sum |FFT(mesh)|^2, k_sum[b] = sum |k|, k_cnt[b] = mode count, over the full c2c grid.
 * Outputs are float64 device arrays of nedges+1 entries (overwritten), 1 <= nedges <= 512.
 * Needs pmx_plan_set_box. */
int pmx_power(pmx_plan plan, const float* mesh, const double* kedges, int nedges, double* pk_sum, double* k_sum,
              double* k_cnt, void* stream);

/* Cross spectra and Legendre multipoles of the same estimator (_power with mesh2 and multipoles,
 * src/jax_fli/_src/power/_compute.py:5-20, 47-63, 92-112). For each of the npoles poles l = poles[j]
 * (host array, 0 <= l <= 16, 1 <= npoles <= 16) and bin b:
 *   p_re[j*(nedges+1) + b] = sum Re(m1 conj(m2)) (2l+1) P_l(mu),  p_im[...] = the same sum of Im,
 * m1 = FFT(mesh), m2 = FFT(mesh2) (m2 = m1 when mesh2 is NULL: then p_im is not written and may be
 * NULL), mu = k.los / |k| (0 at k = 0) with los (host double[3]) normalised here; los may be NULL
 * when every pole is 0. k_sum, k_cnt as for pmx_power. The caller forms |sum| = sqrt(re^2 + im^2) for
 * cross spectra, as the reference does. float64 device outputs, overwritten; 1 <= nedges <= 512.
 * Uses the plan's spectrum buffers: calls on one plan must not overlap. Needs pmx_plan_set_box. */
int pmx_power_multipoles(pmx_plan plan, const float* mesh, const float* mesh2, const double* kedges, int nedges,
                         const int* poles, int npoles, const double* los, double* p_re, double* p_im,
                         double* k_sum, double* k_cnt, void* stream);

/* Gaussian width (radians) of the rbf_neighbor shell paint (scheme 2 of the HEALPix paints): each particle
 * is spread over its pixel and the pixels of the eight directions one pixel resolution away, weighted by
 * exp(-d^2 / (2 sigma^2)) of the angular distance d to each pixel centre and normalised to sum 1 (the
 * "rbf_neighbor" option of PaintingOptions, fields/painting.py:82; the upstream kernel in
 * jaxpm.spherical is not in the reference, so this definition is this library's own). 0 (the default):
 * the pixel resolution sqrt(4 pi / npix) read as a FWHM. The scheme has no adjoint. */
int pmx_plan_set_rbf_width(pmx_plan plan, double sigma_rad);

/* HEALPix resolution change with healpy.ud_grade's semantics (as applied after painting at paint_nside by
 * jaxpm.spherical.paint_particles_spherical, called at src/jax_fli/_src/fields/_painting.py:329-347):
 * in NESTED order a degraded pixel averages its finite, not-UNSEEN (-1.6375e30) children (pess: UNSEEN
 * unless all are good, else UNSEEN when none is) and an upgraded pixel copies its parent; a non-zero power
 * scales by (nside_out / nside_in)^power. order_in / order_out: 0 RING, 1 NESTED. nside: powers of two
 * up to 8192. Device float32 maps of 12 nside^2 values; in and out must not overlap. */
int pmx_healpix_ud_grade(const float* in, int nside_in, float* out, int nside_out, float power, int order_in,
                         int order_out, int pess, void* stream);

/* Counter-based white noise N(0,1) (Philox4x32-10 + Box-Muller) filling (nx,ny,nz) -- stands in for
 * normal_field(key) (initial.py:93). pmx_philox4x32 writes raw 4x32 blocks for known-answer tests. */
int pmx_white_noise(pmx_plan plan, uint64_t seed, float* out, void* stream);
int pmx_philox4x32(uint64_t seed, size_t nblocks, uint32_t* out, void* stream);

/* Colour a real field by sqrt(P(|k|) * ncell / volume): interpolate_initial_conditions
 * (initial.py:124-199). ktab/ptab: device float64 tables (linear interpolation, linear
 * extrapolation). Needs pmx_plan_set_box. */
int pmx_color_field(pmx_plan plan, const float* field, const double* ktab, const double* ptab, int ntab, float* out,
                    void* stream);

/* First-order LPT at the lattice: F = pm_forces(lattice, delta=ic) (unit scale),
 * dx = growth * F, p = vel_factor * dx. jaxpm.pm.lpt(order=1) as called by lpt (pm/lpt.py:119-127). */
int pmx_lpt1(pmx_plan plan, const float* ic, float growth, float vel_factor, float* dx, float* p, void* stream);

/* pmx_lpt1 for a particle lattice coarser than the mesh (npx, npy, npz dividing it): particle
 * (i, j, k) sits on cell (i nx/npx, j ny/npy, k nz/npz) and reads the force there. This is
 * jaxpm.pm.lpt with initial_particles set to that sub-lattice (pm/lpt.py:86-127), the set-up of
 * 512^3 particles on a 1024^3 mesh. */
int pmx_lpt1_lattice(pmx_plan plan, const float* ic, int npx, int npy, int npz, float growth, float vel_factor,
                     float* dx, float* p, void* stream);

/* LPT of order 1 or 2 at the particle lattice (jaxpm.pm.lpt as called by lpt, src/jax_fli/pm/lpt.py:119-127;
 * upstream JaxPM: Jenkins 2010 eqs. 2-3):
 *   dx = D1 F(ic) [+ D2 F(delta2)],  pv = c1 (D1 F(ic)) [+ c2 (D2 F(delta2))],
 * F(.) the FastPM force of a density at the lattice, delta2 = sum_{i<j} (H_ii H_jj - H_ij^2),
 * H_ij = ifft(g_i g_j fft(ic) / k^2) with g the FastPM gradient kernel. coef (float64 device array) holds
 * the columns D1, c1, D2, c2 of ntab rows each: ntab == 1 constants (snapshot LPT: D1 = D(a), c1 = a^2 f E,
 * D2 = 3/7 D2(a)/D2(1), c2 = a^2 f2 E); ntab >= 2 a table over the comoving distance r of each lattice
 * particle, r = |lattice cell * box / mesh - observer| (Mpc/h), on the uniform grid [0, r_max], linearly
 * interpolated (lightcone LPT, per-particle scale factors a(r)). order 1 reads columns D1, c1 only.
 * Uses the plan's density and force buffers. */
int pmx_lpt(pmx_plan plan, const float* ic, int npx, int npy, int npz, int order, const double* coef, int ntab,
            double r_max, const double box[3], const double observer[3], float* dx, float* pv, void* stream);

/* Cotangent of the initial field for pmx_lpt1 / pmx_lpt1_lattice: ic_bar = the transpose of
 * F(ic) applied to growth * (dx_bar + vel_factor * p_bar) placed on the lattice cells. dx_bar or
 * p_bar may be null (zero). Replaces jax.vjp of jaxpm.pm.lpt(order=1). */
int pmx_lpt1_vjp(pmx_plan plan, const float* dx_bar, const float* p_bar, int npx, int npy, int npz, float growth,
                 float vel_factor, float* ic_bar, void* stream);

/* PGD / sharpening displacement S = alpha * read(x, F_pgd), F_pgd the PM force of paint(x) with the
 * Green's function multiplied by pgd_range(k) = exp(-kl^2 / k^2) exp(-k^4 / ks^4) (0 at k = 0, k in
 * cell units, the FastPM 4-point gradient): PGDKernel / SharpeningKernel._compute_boost
 * (src/jax_fli/pm/correction.py:131-333). target 0: x += S (PGDKernel.apply); target 1:
 * v += S * v_scale (SharpeningKernel: v_scale = +1/full_drift_factor to boost, -1/full_drift_factor
 * to rewind). */
int pmx_pgd(pmx_plan plan, float* x, float* v, int npx, int npy, int npz, int mode, float alpha, float kl, float ks,
            int target, float v_scale, void* stream);

/* Cotangent of the PGD displacement S = alpha * read(x, F_pgd) of pmx_pgd: pos_bar (+)= J_S(pos)^T
 * disp_bar; dot (optional, float64, +=) disp_bar . S. The jax.vjp of PGDKernel._compute_boost /
 * SharpeningKernel (src/jax_fli/pm/correction.py:169-200, 301-333) for the checkpointed adjoint. */
int pmx_pgd_vjp(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, float alpha, float kl, float ks,
                const float* disp_bar, float* pos_bar, int accumulate, double* dot, void* stream);

/* HEALPix shell paint: particles with rmin < |x - observer| <= rmax (Mpc/h, periodic box) painted
 * to RING pixels, scheme 0 = ngp, 1 = bilinear (HEALPix ring interpolation weights), 2 = rbf_neighbor
 * (pmx_plan_set_rbf_width), map = weighted count / (pixel solid angle * (rmax^3 - rmin^3) / 3). paint_particles_spherical as called
 * by _single_paint_spherical (src/jax_fli/_src/fields/_painting.py:249-348). */
int pmx_paint_healpix(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                      float weight, const double box[3], const double observer[3], double rmin, double rmax,
                      int nside, int scheme, float* map, int accumulate, void* stream);

/* pmx_paint_healpix over a table of `ntiles` images of the periodic box (at most 4096): image t
 * places the observer-centred position X at R_t X + tile_shift[3t..3t+2] (Mpc/h), R_t the cube
 * symmetry tile_rot[t] of the 48-entry table of get_all_47_symmetries (pm/interp.py:31-43; index
 * 6 * sign_index + permutation_index, 0 = identity). Device arrays tile_rot (int32) and tile_shift
 * (float64). The images are summed into one map: OnionTiler._paint_tiled_27 (pm/interp.py:425-504)
 * with 27 images, and plain periodic replication with identity rotations. tile_rot == NULL paints
 * the box alone (= pmx_paint_healpix). */
int pmx_paint_healpix_tiled(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode,
                            const float* weights, float weight, const double box[3], const double observer[3],
                            double rmin, double rmax, int nside, int scheme, int ntiles, const int* tile_rot,
                            const double* tile_shift, float* map, int accumulate, void* stream);

/* pmx_paint_healpix_tiled without a plan and without allocating: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). (nx, ny, nz) is the mesh the positions refer to (the plan's mesh
 * in pmx_paint_healpix_tiled); everything else as there. The accumulator and ring table live in a
 * static device array of 12 * 512^2 + 4 * 512 float64 values (25.2 MB) that every GPU the library
 * runs on holds until the process ends: nside must be at most 512 (code 2 otherwise). The array is
 * shared by all calls on one GPU and cleared by each call, so calls that overlap on one GPU (other
 * streams, other threads) race: serialise them. */
int pmx_paint_healpix_noalloc(int nx, int ny, int nz, const float* pos, int npx, int npy, int npz, int mode,
                              const float* weights, float weight, const double box[3], const double observer[3],
                              double rmin, double rmax, int nside, int scheme, int ntiles, const int* tile_rot,
                              const double* tile_shift, float* map, int accumulate, void* stream);

/* pmx_paint_healpix_tiled with the drift to the lightcone of DriftInterp / OnionTiler
 * (pm/interp.py:110-121, 212-504): every image position X_t (observer-centred, Mpc/h) first moves by
 * f (R_t vel) with f = (Gp(a_t) - gp_now) / dGfa(sqrt(a_now a_t)), a_t = a(chi = |X_t|), and the shell
 * test and the pixel use the moved position. vel: the velocity array (npx, npy, npz, 3) as given
 * (added to Mpc/h positions unchanged, as _drift_to_time does). drift_tables (device float64, 3 x ntab):
 * a(chi) on the uniform grid [0, chi_max], then Gp(a) and dGfa(a) on [a_lo, a_hi], linearly
 * interpolated and clamped; gp_now = Gp(a_now); drift_bound >= max |f| over the tables (it widens the
 * per-workgroup image culling). tile_rot == NULL: the box alone. */
int pmx_paint_healpix_lightcone(pmx_plan plan, const float* pos, const float* vel, int npx, int npy, int npz,
                                int mode, const float* weights, float weight, const double box[3],
                                const double observer[3], double rmin, double rmax, int nside, int scheme,
                                int ntiles, const int* tile_rot, const double* tile_shift,
                                const double* drift_tables, int ntab, double chi_max, double a_lo, double a_hi,
                                double a_now, double gp_now, double drift_bound, float* map, int accumulate,
                                void* stream);

/* Cotangents of pmx_paint_healpix_lightcone: pos_bar and vel_bar (+)= d<map_bar, map>/d(pos, vel), the
 * bilinear weights differentiated as in pmx_paint_healpix_vjp at the drifted image Z = Y + f(|Y|) R v
 * and chained through the drift: dL/dY = G + f'(|Y|) (R v . G) Y / |Y|, dL/dv = f R^T G, f' the
 * derivative of the tables' linear interpolation (the jax.vjp of DriftInterp / OnionTiler's drifted
 * paint, src/jax_fli/pm/interp.py:110-121, 258-319, 476-493). Shell membership is a step and
 * contributes nothing; ngp maps have zero cotangents. With nparam (<= 4) tables of df/dtheta_q on
 * the a(chi) grid (param_tables, nparam x ntab), param_dots[q] += sum over particle images of
 * (R v . G) df/dtheta_q(|Y|) times the weight: the cotangent of theta_q through the drift factor. */
int pmx_paint_healpix_lightcone_vjp(pmx_plan plan, const float* pos, const float* vel, int npx, int npy, int npz,
                                    int mode, const float* weights, float weight, const double box[3],
                                    const double observer[3], double rmin, double rmax, int nside, int scheme,
                                    int ntiles, const int* tile_rot, const double* tile_shift,
                                    const double* drift_tables, int ntab, double chi_max, double a_lo, double a_hi,
                                    double a_now, double gp_now, double drift_bound, const float* map_bar,
                                    float* pos_bar, float* vel_bar, int accumulate, int nparam,
                                    const double* param_tables, double* param_dots, void* stream);

/* One image of the particles as absolute grid positions (npx*npy*npz*3 floats in out): X = the
 * particle's unwrapped position in Mpc/h (lattice + displacement in relative mode, not wrapped into
 * the box) minus observer, Y = R X + shift with R the cube symmetry `rot` (0..47 in the order of
 * get_all_47_symmetries before it drops entry 0, the identity), and with vel (may be null: no drift)
 * Y += f(|Y|) R v, f the lightcone drift factor from the tables of pmx_paint_healpix_lightcone; out =
 * (Y + observer) / cell. The rotate / shift / drift that TelephotoInterp applies before its paint
 * (_paint_telephoto and _paint_single, src/jax_fli/pm/interp.py:590-660; _drift_to_time, 110-121). */
int pmx_image_positions(pmx_plan plan, const float* pos, const float* vel, int npx, int npy, int npz, int mode,
                        const double box[3], const double observer[3], int rot, const double shift[3],
                        const double* drift_tables, int ntab, double chi_max, double a_lo, double a_hi, double a_now,
                        double gp_now, float* out, void* stream);

/* Cotangents of pmx_image_positions for out_bar (npx*npy*npz*3 floats, of its grid positions):
 * pos_bar (+)= cell R^T Ybar with Ybar = out_bar / cell + f'(|Y|) (R v . out_bar / cell) Y / |Y|, and with
 * vel (the drift) vel_bar (+)= f R^T out_bar / cell; with nparam (<= 4) tables of df/dtheta_q on the
 * a(chi) grid, param_dots[q] += sum (R v . out_bar / cell) df/dtheta_q(|Y|). The jax.vjp of
 * TelephotoInterp's and DriftInterp's image transform and drift (src/jax_fli/pm/interp.py:110-121,
 * 277-289, 590-660). */
int pmx_image_positions_vjp(pmx_plan plan, const float* pos, const float* vel, int npx, int npy, int npz, int mode,
                            const double box[3], const double observer[3], int rot, const double shift[3],
                            const double* drift_tables, int ntab, double chi_max, double a_lo, double a_hi,
                            double a_now, double gp_now, const float* out_bar, float* pos_bar, float* vel_bar,
                            int accumulate, int nparam, const double* param_tables, double* param_dots,
                            void* stream);

/* Flat-sky slab paint: absolute z in (zlo, zhi] (grid units), (x, y) rescaled to (fnx, fny) and
 * wrapped, 2D CIC, times `scale` (1 / pixel volume). _single_paint_2d (_painting.py:117-202). */
int pmx_paint_flat(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                   float weight, int fnx, int fny, float zlo, float zhi, double scale, float* map, int accumulate,
                   void* stream);

/* Position cotangent of pmx_paint_flat for the cotangent map_bar of its (scaled) map: pos_bar (+)=
 * scale * w * d/dpos sum_c w_c(pos) map_bar[c] over the 2D CIC corners (x and y chained through the
 * flat-grid scaling); the slab selection is a step, so z gets nothing. Replaces jax.vjp of
 * _single_paint_2d (src/jax_fli/_src/fields/_painting.py:160-202). */
int pmx_paint_flat_vjp(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                       float weight, int fnx, int fny, float zlo, float zhi, double scale, const float* map_bar,
                       float* pos_bar, int accumulate, void* stream);

/* ---- adjoints (vector-Jacobian products) --------------------------------------------------------
 * Hand-written transposes of the entry points above, for reverse-mode gradients of the reference's
 * `adjoint="reverse"` / `"checkpointed"` integration (src/jax_fli/pm/integrate.py:147-193, 288-538).
 * "accumulate" adds into the cotangent output instead of overwriting it. Dot outputs are float64
 * device scalars, accumulated with atomics (zero them first). */

/* Position (and weight) cotangent of pmx_paint_cic: pos_bar = d<mesh_bar, paint(pos)>/d pos,
 * weights_bar (optional, per particle) = d<mesh_bar, paint>/d w. The CIC weight derivatives follow
 * the reference's 1 - |f - {0,1}| (derivative -sign). Replaces jax.vjp of cic_paint_dx / cic_paint. */
int pmx_paint_cic_vjp(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                      float weight, const float* mesh_bar, float* pos_bar, float* weights_bar, int accumulate,
                      void* stream);

/* Cotangents of pmx_readout_cic: mesh_bar (nx,ny,nz,ncomp; optional) = scatter of out_bar by the
 * CIC weights (the readout's transpose, sorted deposit), pos_bar (optional) = d<out_bar, read>/d pos.
 * Replaces jax.vjp of cic_read_dx / cic_read (fields/particles.py:183-229). */
int pmx_readout_cic_vjp(pmx_plan plan, const float* mesh, int ncomp, const float* pos, int npx, int npy, int npz,
                        int mode, const float* out_bar, float* mesh_bar, float* pos_bar, int accumulate,
                        void* stream);

/* Position cotangent of pmx_paint_healpix_tiled (tile_rot NULL: of pmx_paint_healpix):
 * pos_bar (+)= d<map_bar, map(pos)>/d pos through the bilinear HEALPix weights of every image of a
 * particle inside the shell (the ngp scheme is piecewise constant: zero). Shell membership is a
 * step and contributes nothing, as under jax.vjp. Replaces jax.vjp of paint_particles_spherical
 * (called by _single_paint_spherical, _painting.py:249-348). */
int pmx_paint_healpix_vjp(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, const float* weights,
                          float weight, const double box[3], const double observer[3], double rmin, double rmax,
                          int nside, int scheme, int ntiles, const int* tile_rot, const double* tile_shift,
                          const float* map_bar, float* pos_bar, int accumulate, void* stream);

/* Density cotangent of pmx_force_mesh: density_bar = G^T S^T fmesh_bar (S the 4-point gradient,
 * antisymmetric; G the Green's operator, symmetric). Replaces jax.vjp of the k-space part of
 * pm_forces (used by the LPT adjoint, pm/lpt.py:119). */
int pmx_force_mesh_vjp(pmx_plan plan, const float* fmesh_bar, float scale, unsigned flags, float* density_bar,
                       void* stream);

/* Position cotangent of pmx_forces: pos_bar (+)= J(pos)^T (bar_scale forces_bar), J the Jacobian of
 * pm_forces(pos) * scale (readout-weight and paint-weight terms); dot (optional) += forces_bar . F
 * (forces_bar unscaled). Recomputes the force mesh at pos. Replaces jax.vjp of _forces
 * (src/jax_fli/pm/solvers.py:19-32); bar_scale folds a kick coefficient into the pull-back. */
int pmx_forces_vjp(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, float scale, unsigned flags,
                   const float* forces_bar, float bar_scale, float* pos_bar, int accumulate, double* dot,
                   void* stream);

/* pmx_forces_vjp with the force mesh at pos given (force_mesh, nx*ny*nz*3 floats, as kept by
 * pmx_step_keep for the step that kicked with these positions) instead of recomputed: the sort of pos,
 * the adjoint deposit, the potential pass and the gather only. For the checkpointed adjoint
 * (integrate.py:147-193 with adjoint="checkpointed"). */
int pmx_forces_vjp_saved(pmx_plan plan, const float* pos, int npx, int npy, int npz, int mode, float scale,
                         unsigned flags, const float* force_mesh, const float* forces_bar, float bar_scale,
                         float* pos_bar, int accumulate, double* dot, void* stream);

/* out = x + c v over n floats (out may alias x): the drift x_a = x + c_pre v that the checkpointed
 * adjoint recomputes from a kept state before pulling a kick back (the solvers' drift,
 * src/jax_fli/pm/solvers.py:165-172, 305-312). */
int pmx_axpy(float* out, const float* x, const float* v, float c, size_t n, void* stream);

/* Transpose of a drift x <- x + c v for the cotangents (xbar, vbar) of its output: vbar += c xbar,
 * and dot (float64 device scalar) += <xbar, v>, the cotangent of c (per-workgroup partials summed in a
 * fixed order). The per-op drift pull-back of the checkpointed adjoint (jax.vjp of the solvers'
 * drifts, src/jax_fli/pm/integrate.py:147-193 with adjoint="checkpointed"). */
int pmx_adjoint_drift(pmx_plan plan, const float* xbar, float* vbar, const float* v, size_t n, float c, double* dot,
                      void* stream);

/* One reverse-adjoint step of the reversible KKD solver, in place. The forward step was
 *   v <- v + F(x) * c_kick ;  x <- x + v * c_drift      (F = pm_forces(x) * force_scale)
 * (ReversibleDoubleKickDrift.step, solvers.py:258-325). Given the state after it and the cotangents
 * (x_bar, v_bar) of that state, this restores the state before it (the exact inverse, in place of
 * ReversibleDoubleKickDrift.reverse, solvers.py:339-411) and replaces the cotangents by those of
 * the state before it, reusing the reverse pass's force mesh for the Jacobian:
 *   v_bar <- v_bar + c_drift x_bar ;  x_bar <- x_bar + J(x)^T (c_kick v_bar)
 * coef_bar (2 doubles, overwritten) = {dL/dc_drift, dL/dc_kick} for parameter gradients.
 * Replaces the per-step reverse + jax.vjp(step) of _integrate_bwd (integrate.py:304-358). */
int pmx_step_reverse_vjp(pmx_plan plan, float* x, float* v, int npx, int npy, int npz, int mode, float force_scale,
                         float c_kick, float c_drift, unsigned flags, float* x_bar, float* v_bar, double* coef_bar,
                         void* stream);

/* pmx_step_reverse_vjp with the force mesh of that step taken from force_mesh (nx*ny*nz*3 floats,
 * interleaved as the plan's own), kept by pmx_step_keep during the forward pass, instead of being
 * recomputed at the restored positions: the same inverse and cotangent update without the paint,
 * the five FFT passes and the gradient stencil of the force. The restored x differs from the
 * forward's by the float32 rounding of x - v cd, as in pmx_step_reverse_vjp. Replaces the per-step
 * reverse + jax.vjp(step) of _integrate_bwd (src/jax_fli/pm/integrate.py:304-358) when the HBM
 * holds one force mesh per step. */
int pmx_step_reverse_vjp_saved(pmx_plan plan, float* x, float* v, int npx, int npy, int npz, int mode,
                               float force_scale, float c_kick, float c_drift, unsigned flags, const float* force_mesh,
                               float* x_bar, float* v_bar, double* coef_bar, void* stream);

/* pmx_step_reverse_vjp for a ReversibleDoubleKickDrift step with the sharpening boost
 * (SharpeningKernel, src/jax_fli/pm/correction.py:231-333; solvers.py:258-325, 339-411):
 *   v <- v + F(x) c_kick + S(x) v_scale ;  x <- x + v c_drift     (S as in pmx_pgd, target 1)
 * restored and pulled back in place: v_bar <- v_bar + c_drift x_bar ;
 * x_bar <- x_bar + J_F(x)^T (c_kick v_bar) + J_S(x)^T (v_scale v_bar), with the force and the boost
 * as one operator (Green's function times c_kick force_scale + v_scale alpha exp(-kl^2/k^2)
 * exp(-k^4/ks^4)), so the step costs what pmx_step_reverse_vjp does. coef_bar = {dL/dc_drift,
 * <v_bar, c_kick F + v_scale S>}. */
int pmx_step_reverse_vjp_sharpen(pmx_plan plan, float* x, float* v, int npx, int npy, int npz, int mode,
                                 float force_scale, float c_kick, float c_drift, unsigned flags, float alpha, float kl,
                                 float ks, float v_scale, float* x_bar, float* v_bar, double* coef_bar, void* stream);

/* ---- Born convergence ---------------------------------------------------------------------------
 * kappa[k][p] = sum_s coef[k][s] * (maps[s][p] / mean_s - 1) over `nshell` HEALPix (or any
 * same-length) density maps, mean_s the map's pixel mean (a zero mean is replaced by FLT_EPSILON).
 * coef (device float64, nsrc x nshell, at most 16 x 256) holds
 * 3/2 Om (H0/c)^2 d_r r_s / a_s times the source-weighted lensing efficiency clip(1 - r_s/chi, 0, 1000):
 * _born_core_impl and _born_spherical (src/jax_fli/_src/lensing/_born.py:29-127) with the n(z)
 * quadrature folded into coef on the host. Uses the plan only for scratch. */
int pmx_born(pmx_plan plan, const float* maps, int nshell, int64_t npix, const double* coef, int nsrc, float* kappa,
             int accumulate, void* stream);

/* Cotangent of pmx_born: maps_bar (+)= d<kappa_bar, kappa>/d maps, including the dependence of
 * the shell means. dots (optional, device float64, nsrc*nshell + nsrc + nshell): D[k*nshell+s] =
 * sum_p kappa_bar[k][p] maps[s][p], then E[k] = sum_p kappa_bar[k][p], then the means m[s] as
 * used, from which the host forms d<kappa_bar, kappa>/d coef[k][s] = D / m_s - E. */
int pmx_born_vjp(pmx_plan plan, const float* maps, int nshell, int64_t npix, const double* coef, int nsrc,
                 const float* kappa_bar, float* maps_bar, int accumulate, double* dots, void* stream);

/* Flat-sky Born convergence (_born_core_impl's flat branch, src/jax_fli/_src/lensing/_born.py:53-98,
 * called by _born_flat, :131-190): planes maps (nshell, ny, nx); each plane is resampled at output
 * pixel (i, j) by periodic bilinear interpolation (map_coordinates order 1, mode "wrap") at row
 * coordinate j * scales[2 s] - 1/2 and column coordinate i * scales[2 s + 1] - 1/2 (the reference's
 * angular grid: x angles index rows, y angles columns, times chi_s / pixel size), then
 *   kappa[k] (+)= sum_s coef[k, s] (R_s / mean(maps[s]) - 1)        (nsrc, ny, nx)
 * with the plane means of the unresampled maps (0 -> FLT_EPSILON) and coef as for pmx_born. */
int pmx_born_flat(pmx_plan plan, const float* maps, int nshell, int ny, int nx, const double* coef, int nsrc,
                  const double* scales, float* kappa, int accumulate, void* stream);

/* Cotangent of pmx_born_flat: maps_bar (+)= d<kappa_bar, kappa>/d maps (the resampling's transpose by
 * float atomics, plus the means' term). dots (optional, nsrc*nshell + nsrc + 2*nshell doubles) =
 * {D[k, s] = <kappa_bar_k, R_s>, E[k] = sum kappa_bar_k, the means as used, P[s] = sum_p g_s(p)
 * (dR_s/du j scales[2s] + dR_s/dw i scales[2s+1]) with g_s = sum_k coef[k, s] kappa_bar_k / mean_s}:
 * dL/dcoef[k, s] = D[k, s] / mean_s - E[k] and dL/dchi_s through the sampling positions = P[s] / chi_s. */
int pmx_born_flat_vjp(pmx_plan plan, const float* maps, int nshell, int ny, int nx, const double* coef, int nsrc,
                      const double* scales, const float* kappa_bar, float* maps_bar, int accumulate, double* dots,
                      void* stream);

/* ------------------------------------------------------------------------------------------------
 * Forward-simulation server: continuous batching of seeded forward simulations of mixed mesh sizes
 * and cosmologies on one GPU. In place of the reference's one-simulation-per-call loop
 * (run_simulations, src/jax_fli/scripts/fli_simulate.py:512-585, each call lpt + nbody with
 * EfficientDriftDoubleKick, pm/nbody.py:28-171). A request: white noise from `seed` (Philox, as
 * pmx_white_noise), EH98 P(k) of its flat LCDM cosmology (Omega_k = 0, w0 = -1, wa = 0) normalised to
 * sigma8 colouring it in a cubic box, first-order LPT at a0, the KDK loop with dt0 to a1; results: the
 * final GRID_RELATIVE displacements (mesh^3 x 3 floats) and the painted density (mesh^3 floats), written
 * into the caller's device buffers. One size class per mesh size, each with `slots` concurrent requests
 * (0: as many as hbm_budget / nclass allows, at most 64); a request is admitted when a slot of its size
 * is free, in arrival order per size. Steps of all live slots of a class run as one batched launch per
 * phase, replayed from a captured HIP graph. Host preparation runs on prep_threads threads (0: 4). */
typedef struct pmx_server_t* pmx_server;
typedef struct {
  int mesh;        /* n: n^3 particles on an n^3 mesh, one of the server's sizes */
  uint64_t seed;   /* white-noise seed */
  double Omega_c, Omega_b, h, n_s, sigma8;
  double box;      /* Mpc/h */
  double a0, a1, dt0;
} pmx_request_t;
typedef struct {
  uint64_t submitted, completed, failed, running, queued;
  int peak_running;                        /* most slots held at once since the server was created */
  double p50_ms, p99_ms, mean_ms, max_ms;  /* submit -> completion latency of completed requests */
  double wall_s;                           /* first submission to last completion */
  double throughput_req_s;                 /* completed / wall_s */
  size_t device_bytes;                     /* the server's device memory (slot arenas and tables) */
} pmx_server_stats_t;
#define PMX_REQ_QUEUED 0
#define PMX_REQ_RUNNING 1
#define PMX_REQ_DONE 2
#define PMX_REQ_FAILED (-1)
/* Allocates every class's slots and captures its step graph (synchronous); starts the server threads. */
int pmx_server_create(pmx_server* server, const int* meshes, const int* slots, int nclass, size_t hbm_budget,
                      int prep_threads);
/* Slots of the class for `mesh` and the device bytes of one slot. */
int pmx_server_class(pmx_server server, int mesh, int* slots, size_t* slot_bytes);
/* Queue a request (returns at once). displacement / density: device buffers the result is written
 * into; valid once the request is PMX_REQ_DONE. */
int pmx_server_submit(pmx_server server, const pmx_request_t* request, float* displacement, float* density,
                      uint64_t* id);
/* State of a request (PMX_REQ_*), without waiting. */
int pmx_server_poll(pmx_server server, uint64_t id, int* state);
/* Block until the request is done (timeout_s <= 0: no limit; code 5 on timeout). */
int pmx_server_wait(pmx_server server, uint64_t id, double timeout_s);
/* Submission, admission and completion times of a request (seconds since the server was created). */
int pmx_server_request_times(pmx_server server, uint64_t id, double times[3]);
/* Counts and latency percentiles over the requests with id >= since_id. */
int pmx_server_stats(pmx_server server, uint64_t since_id, pmx_server_stats_t* stats);
/* Finishes the admitted requests, stops the threads and frees everything (synchronous). */
int pmx_server_destroy(pmx_server server);

#ifdef __cplusplus
}
#endif
#endif
