/*
 * spk.h -- C ABI of the MI355X (gfx950) SPOOKY solver library (libspk.so).
 *
 * Plain C types only: handles, device pointers as void*, sizes, doubles.
 * Every call is asynchronous on the context's HIP stream unless it says it
 * synchronizes.  Every call returns SPK_OK (0) or a negative status; the text
 * of the last error on the calling thread is in spk_last_error().  Unlike the
 * reference's CUDA_RT_CALL/CUFFT_CALL (libs/cufft_utils.h:53-84), which print
 * and continue, errors here are returned and nothing continues past them.
 *
 * Layouts are the reference's (common.hpp:12-14, parameters.cu:650):
 *   spectral field : complex [NX][NY][NZ/2+1], kz fastest, interleaved re/im
 *   real field     : [NX][NY][2*(NZ/2+1)] (2 padding values per row)
 *   field set      : field n at offset n*NTC complex, NTC = NX*NY*(NZ/2+1)
 *   field order    : vx,vy,vz,bx,by,bz,th as present (parameters.cu:568-640)
 * The solver state is FP64 complex.  FFTs and real-space products run in the
 * solver's "bulk" precision (SPK_F64, or SPK_F32 for the FP32-bulk mode).
 */
#ifndef SPK_H
#define SPK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPK_ABI_VERSION 1

enum spk_status {
    SPK_OK = 0,
    SPK_ERR_ARG = -1,         /* bad argument or unsupported shape */
    SPK_ERR_HIP = -2,         /* HIP runtime error */
    SPK_ERR_ALLOC = -3,       /* device allocation failed */
    SPK_ERR_STATE = -4        /* call not valid for this solver configuration */
};

enum spk_precision { SPK_F64 = 0, SPK_F32 = 1 };

int spk_abi_version(void);
const char *spk_last_error(void);

/* ---- context: device + stream (replaces the default-stream globals) ---- */
typedef struct spk_ctx_s *spk_ctx;
/* stream: a hipStream_t, or NULL for the null stream */
int spk_ctx_create(int device, void *stream, spk_ctx *out);
int spk_ctx_destroy(spk_ctx ctx);
int spk_ctx_sync(spk_ctx ctx); /* synchronizes the context stream */
int spk_malloc(spk_ctx ctx, size_t bytes, void **out);
int spk_free(spk_ctx ctx, void *ptr);
int spk_memcpy_h2d(spk_ctx ctx, void *dst, const void *src, size_t bytes); /* synchronizes */
int spk_memcpy_d2h(spk_ctx ctx, void *dst, const void *src, size_t bytes); /* synchronizes */

/* ---- 3D real<->complex FFT ----
 * Replaces init_plan / finish_cufft / r2c_fft / c2r_fft
 * (src/cuda/cufft_routines.hpp:10-17, cufft_routines.cu:14-101).
 * Power-of-two sizes, 2 <= NX,NY <= 2048, 2 <= NZ <= 1024.
 * batch transforms fields at a stride of NTC complex (2*NTC reals). */
typedef struct spk_fft_plan_s *spk_fft_plan;
int spk_fft_plan_create(spk_ctx ctx, int nx, int ny, int nz, int precision, int batch,
                        spk_fft_plan *out);
int spk_fft_plan_destroy(spk_fft_plan plan);
/* forward, unnormalized; in: padded real, out: complex; in == out allowed
 * (replaces r2c_fft, cufft_routines.cu:14-20) */
int spk_fft_r2c(spk_fft_plan plan, const void *in, void *out);
/* inverse with the 1/(NX NY NZ) scaling folded in; in == out allowed; when
 * in != out the input is left untouched (replaces scaleKernel + c2r_fft,
 * cufft_routines.cu:23-40).  Imaginary parts on the kz=0 and kz=NZ/2 planes
 * are ignored; the two padding reals of every row are written as 0. */
int spk_fft_c2r(spk_fft_plan plan, const void *in, void *out);

/* One pass of a plan, for timing and for callers that fuse their own passes:
 * axis 0 = x, 1 = y: complex lines, dir -1 forward / +1 inverse, unscaled;
 * axis 2 = z: dir -1 real rows -> half-complex rows, +1 the reverse (padding
 * written as 0).  in == out allowed.  No counterpart call in the reference,
 * whose cufftXtExec runs all three axes at once (cufft_routines.cu:17,35). */
int spk_fft_pass(spk_fft_plan plan, int axis, int dir, const void *in, void *out);

/* ---- run parameters (Parameters, parameters.cu:77-330) ---- */
typedef struct spk_params {
    int nx, ny, nz;
    int incompressible, heat_equation, boussinesq, stratification, strat_direction;
    int mhd, anisotropic_diffusion, supertimestepping, shearing, rotating;
    int with2d, antialiasing;
    int sts_mask;   /* bit f set: field f is an sts_variable, evolved by RKL2 */
    int safety_sts;
    double lx, ly, lz;
    double nu, nu_m, nu_th, reynolds_ani, N2, OmegaT2, omega, shear;
    double cfl, cfl_par, cfl_sts, safety_source;
    double t_initial, t_final, toutput_flow;
} spk_params;

size_t spk_params_size(void);

/* ---- solver: device state of Fields + TimeStepping + RKLegendre ---- */
typedef struct spk_solver_s *spk_solver;
int spk_solver_create(spk_ctx ctx, const spk_params *p, int bulk_precision, spk_solver *out);
int spk_solver_destroy(spk_solver s);
/* vars[10] = VX,VY,VZ,BX,BY,BZ,TH,VEL,MAG,NUM_FIELDS; *ntc = NTC */
int spk_solver_info(spk_solver s, int *vars, size_t *ntc);
/* device pointers: state (Fields::d_all_fields), dU (d_all_dfields), real buffer
 * (d_all_buffer_r, padded real [nf][2 NTC]) */
int spk_solver_buffers(spk_solver s, void **fields, void **dfields, void **realbuf);
/* time state out8 = t, dt, tremap, dt_hyp, dt_par, dt_sts, t_lastsnap, step (synchronizes) */
int spk_solver_get_time(spk_solver s, double *out8);
int spk_solver_set_time(spk_solver s, double t, int step, double t_lastsnap);

/* Fields::CleanFieldDivergence (src/fields/symmetries.cu:114-130) */
int spk_clean_divergence(spk_solver s);
/* Supervisor::Complex2RealFields (src/supervisor/supervisor.cu:261-272): state -> real buffer */
int spk_complex2real(spk_solver s);
/* TimeStepping::compute_dt (src/timestepping/compute_dt.cu:15-309); dt stays on the device */
int spk_compute_dt(spk_solver s);
/* TimeStepping::compute_dfield (src/timestepping/compute_dfield.cu:11-90) into dU */
int spk_compute_dfield(spk_solver s);
/* TimeStepping::RungeKutta3 (src/timestepping/hydro_mhd_advance.cu:82-189); uses the current dt */
int spk_rk3(spk_solver s);
/* RKLegendre::compute_cycle_RKL2 (src/rkl/rkl.cu:437-591); synchronizes to size the cycle */
int spk_rkl2_cycle(spk_solver s);
/* TimeStepping::HydroMHDAdvance (src/timestepping/hydro_mhd_advance.cu:18-75) */
int spk_hydro_mhd_advance(spk_solver s);
/* nsteps iterations of Supervisor::executeMainLoop's body without file output
 * (supervisor.cu:169-190): CleanFieldDivergence + HydroMHDAdvance, with the
 * snapshot clock of CheckSnapshot (check_output.cu:62-94) kept on the device */
int spk_main_loop(spk_solver s, int nsteps);
/* As spk_main_loop, continuing the previous spk_main_loop or spk_main_loop_resume
 * call on s: the caller has not changed the state since (only spk_solver_time and
 * other reads in between, no writes through spk_solver_buffers' pointers).  On
 * the fused back end the first step then starts from the inverse x pass of the
 * state that the last step left (whose stage 3 also made the projection),
 * instead of a fresh CleanFieldDivergence and inverse x pass: a caller that
 * steps one at a time between outputs, as Supervisor::executeMainLoop
 * (supervisor.cu:169-213) does, gets the throughput of one long call.  Any
 * other entry point taking s in between makes the next call start fresh.  No
 * counterpart in the reference. */
int spk_main_loop_resume(spk_solver s, int nsteps);

/* The fused back end of an RK3 stage runs the forward x pass, the spectral
 * epilogue, the RK3 update, the next step's CleanFieldDivergence and the next
 * stage's inverse x pass in one kernel; spk_main_loop can use it where the
 * configuration allows (no shearing or in-RHS conduction; 32 <= NX <= 1024).
 * mode 0: never; 1 (default): when the tile autotune run at spk_solver_create
 * found it faster than the unfused kernels on this grid; 2: always where
 * allowed.  All give the same results to rounding. */
int spk_solver_set_fused(spk_solver s, int mode);
/* Points per thread of the fused kernel's x line: 4 or 8, or 0 for the tile
 * autotune's choice (the default).  All give the same results to rounding.
 * No counterpart in the reference (its x passes are cuFFT's). */
int spk_solver_set_fused_points(spk_solver s, int points);
/* Quasi-2D grids (NZ = 2, NY = 64..512): the inverse y pass, the real-space
 * products and the forward y pass of an RHS evaluation run as one kernel per x
 * plane, and so does the flux of a fused RKL2 stage.  on = 1 (default) uses
 * them where the grid and RHS program allow, 0 never.  Both give the same
 * results to rounding.  No counterpart in the reference (its y and z
 * transforms are cuFFT's). */
int spk_solver_set_plane(spk_solver s, int on);
/* Whether spk_main_loop runs the plane kernel in its RK3 stages (*rk3) and in
 * its RKL2 stages (*rkl2). */
int spk_solver_plane(spk_solver s, int *rk3, int *rkl2);
/* RKL2 cycles planned on the device (on = 1): the stage count and coefficients
 * of RKLegendre::compute_cycle_RKL2 (rkl.cu:466-585) are computed by a one-thread
 * kernel from the step's dt and dt_sts, and the cycle launches the stage count
 * of the sts cap (dt = safety_sts dt_sts, compute_dt.cu:274-282) with the stages
 * beyond the step's count returning at once.  No host synchronization per cycle,
 * so steps can be captured in a graph.  Needs super-time-stepping of an
 * incompressible run and at most 64 stages at the cap.  Same results as the
 * host-planned cycle (on = 0, the default). */
int spk_solver_set_rkl_device(spk_solver s, int on);
/* Large-scale forcing (no counterpart in the reference): a fixed spectral
 * field F on nmodes modes (index[m]: position in a field, 0 <= index < NTC;
 * values[m][3][2]: the complex F of vx, vy, vz there); after every step of
 * spk_main_loop the state gets U += dt F on those modes.  nmodes = 0 removes it.
 * Incompressible runs only; synchronizes. */
int spk_solver_set_forcing(spk_solver s, int nmodes, const long long *index, const double *values);
/* The last device plan: its stage count, whether the step's count exceeded the
 * plan (overflow = 1: the cycle was cut short), and the stages launched per cycle. */
int spk_solver_rkl_plan(spk_solver s, int *stages, int *overflow, int *smax);
/* Which back end spk_main_loop will run: *fused = 1 for the fused kernel;
 * *points = points per thread of its x line (the autotuned tile). */
int spk_solver_pipeline(spk_solver s, int *fused, int *points);

/* Per-kernel timing of the RK3 pipeline, each kernel class launched reps times
 * exactly as inside a step, timed with HIP events on the context stream (the
 * reference times every FFT with a cudaEvent pair, cufft_routines.cu:47-53).
 * out_ms[k] = average ms per launch for k = 0 inverse x pass (from the state),
 * 1 inverse y pass, 2 fused z kernel, 3 forward y pass, 4 forward x pass,
 * 5 spectral epilogue (RK3 stage 2), 6 CleanFieldDivergence, 7 dt, 8 the fused
 * forward x pass + epilogue + next inverse x pass (RK3 stage 2); -1 where a
 * class does not run for this configuration.  *count = 9.  The state is left
 * unchanged; the RK3 scratch is overwritten.  Synchronizes. */
int spk_kernel_timings(spk_solver s, int reps, double *out_ms, int cap, int *count);

/* ---- measurement (src/io/spooky_output_functions.cu); results synchronize ---- */
int spk_energy(spk_solver s, int field, double *out);            /* computeEnergy :23-55 */
int spk_enstrophy(spk_solver s, int vec_first_field, double *out); /* computeEnstrophy :57-86 */
int spk_correlation(spk_solver s, int fa, int fb, double *out);  /* twoFieldCorrelation :188-235 (real buffer) */
/* Spectrum1d (cuda_kernels_generic.cu:441-471) + computeSpectrum1d
 * (spooky_output_spectrum.cu:424-444); out[nbins]; bins beyond nbins are dropped */
int spk_spectrum(spk_solver s, int fa, int fb, double *out, int cap, int *nbins);
/* One quantity of the timevar file by the reference's name, as WriteTimevarOutput
 * computes it (spooky_output_timevar.cu:48-204, spooky_output_functions.cu:23-395):
 * ev em et Kx Ky Kz Mx My Mz vxvy vxvz vyvz bxby bxbz bybz thvx thvz w2 j2 hc hv
 * hm thetaB bz2 phiB pv dissgradT fluxbbgradT.  Quantities from real fields read
 * the real buffer, which must hold the current state (spk_complex2real), as the
 * reference fills d_all_buffer_r before writing a row.  SPK_ERR_ARG for a name
 * that is not a quantity of this configuration ("t" is the host's clock).
 * Differences: computeAnisoDissipation accumulates into scratch it does not
 * clear (DivergenceMask flag 1); here the divergence is written, not added. */
int spk_timevar(spk_solver s, const char *name, double *out);
/* Shell bins of the spectra: nbins and deltak of Wavevector (wavevector.cu:98-110);
 * bin i is centred on i * deltak, as WriteSpectrumOutputHeader writes
 * (spooky_output_spectrum.cu:448-480). */
int spk_spectrum_bins(spk_solver s, int *nbins, double *deltak);
/* One spectrum of the spectrum file by the reference's name, as
 * WriteSpectrumOutput computes it (spooky_output_spectrum.cu:22-400): Kx Ky Kz
 * Mx My Mz Eth vxvy vyvz vxvz bxby bybz bxbz thvx thvy thvz, and the transfer
 * spectra advec_{x,y,z} emfpower_{x,y,z} lorentz_{x,y,z} th_advec
 * helicity_{x,y,z} th_aniso_diss th_aniso_inj.  Transfer spectra read the real
 * buffer, which must hold the current state (spk_complex2real).  Mirrored
 * quirks: bxbz pairs B_x with B_y (:153-158); th_aniso_diss adds the
 * anisotropic dissipation to the unmasked transform of B_y B_z, which the
 * reference's scratch array holds at that point.  SPK_ERR_ARG for a name that
 * is not a spectrum of this configuration (the reference writes no line). */
int spk_spectrum_named(spk_solver s, const char *name, double *out, int cap, int *nbins);
/* One quantity of the user timevar file (WriteUserTimevarOutput,
 * src/user/user_output_timevar.cu:18-76): kpartheta (computekpartheta,
 * user_output_functions.cu:31-58; needs anisotropic_diffusion; reads the real
 * buffer), uservar1 and uservar2 (0, as shipped).  SPK_ERR_ARG otherwise. */
int spk_user_timevar(spk_solver s, const char *name, double *out);
/* The real fields a snapshot holds (CheckSnapshot, check_output.cu:70-84) into
 * the real buffer: Complex2RealFields, or with shearing UnshearOutput
 * (unshear_output.cu:21-54), which first multiplies a copy of every field by
 * exp(-i ky Lx S fmod(t, 2 Ly / (S Lx)) / 2).  Synchronizes to read t. */
int spk_snapshot_real(spk_solver s);

/* ---- the same operations on caller-owned arrays ----
 * The reference's methods take the field arrays as arguments and its host
 * classes allocate every device array (Supervisor::spookyGpuAlloc,
 * supervisor.hpp:74-78).  These entries run the solver's kernels on arrays the
 * caller owns (allocated with spk_malloc or any hipMalloc on the context's
 * device): complex_fields = d_all_fields, [nf][NX][NY][NZ/2+1] complex FP64;
 * real_buffer = d_all_buffer_r, [nf][NX][NY][2(NZ/2+1)] FP64.  The solver handle
 * supplies the parameters, the grid and the library's transform buffers.
 * Complex arrays must be 16-byte aligned.  A real_buffer documented as not read
 * may be NULL. */
/* TimeStepping::compute_dt(complexFields, realBuffer) (timestepping.hpp:39,
 * compute_dt.cu:15-309): dt, dt_hyp, dt_par, dt_sts on the device; when
 * real_buffer is not NULL it receives the real fields, as the reference fills it */
int spk_compute_dt_ptr(spk_solver s, const void *complex_fields, double *real_buffer);
/* TimeStepping::compute_dfield(complex_Fields, real_Buffer, complex_dFields)
 * (timestepping.hpp:40, compute_dfield.cu:11-90); real_buffer is not read */
int spk_compute_dfield_ptr(spk_solver s, const void *complex_fields, double *real_buffer, void *complex_dfields);
/* TimeStepping::RungeKutta3(complex_Fields, real_Buffer) (timestepping.hpp:42,
 * hydro_mhd_advance.cu:82-189) with the dt of the last compute_dt; scratch =
 * d_all_scrtimestep (timestepping.cu:36), nf fields, or NULL for the solver's
 * own; real_buffer: scratch of the shearing-box remap only.  RK3 stage 1 reuses
 * the products of a compute_dt on the same complex_fields, as the reference
 * reuses the real buffer compute_dt filled. */
int spk_rk3_ptr(spk_solver s, void *complex_fields, double *real_buffer, void *scratch);
/* TimeStepping::HydroMHDAdvance (hydro_mhd_advance.cu:18-75) on caller arrays */
int spk_hydro_mhd_advance_ptr(spk_solver s, void *complex_fields, double *real_buffer, void *scratch);
/* Fields::CleanFieldDivergence (symmetries.cu:114-130) on caller fields */
int spk_clean_divergence_ptr(spk_solver s, void *complex_fields);
/* CleanDivergence of one 3-component vector field (cuda_kernels_generic.cu:288-313) */
int spk_project(spk_solver s, void *complex_vec);
/* Supervisor::Complex2RealFields(ComplexField_in, RealField_out, num_fields)
 * (supervisor.hpp:72, supervisor.cu:261-272); the input is not modified */
int spk_complex2real_ptr(spk_solver s, const void *complex_in, double *real_out, int num_fields);
/* RKLegendre::compute_cycle(complex_Fields, real_Buffer) (rkl.hpp:42, RKL2 cycle
 * rkl.cu:437-591); synchronizes to size the cycle; real_buffer is not read */
int spk_rkl2_cycle_ptr(spk_solver s, void *complex_fields, double *real_buffer);
/* One stage of that cycle (rkl.cu:496-585): ParabolicTermsSTS of the sts
 * variables (physics.cu:233-330, with AnisotropicHeatFlux for theta) and the
 * rkl2_stage update (cuda_kernels_generic.cu:160-172).  coef = mu_j, nu_j,
 * mu~_j, gamma~_j of the recurrence (rkl.cu:466-585; stage 1 uses mu~_1 only);
 * uc0, uc1, du0: the RKLegendre arrays d_all_Uc0, d_all_Uc1, d_all_dU0
 * (rkl.cu:125-153), num_sts_vars fields each, caller-owned. */
int spk_rkl2_stage(spk_solver s, int stage, const double *coef, double dt, void *complex_fields, void *uc0,
                   void *uc1, void *du0);
/* Physics::AnisotropicHeatFlux(complex_Fields, real_Buffer, complex_dTheta)
 * (physics.hpp:27, anisotropic_conduction.cu:10-62): complex_dtheta += i mask
 * k . FFT[chi b (b . grad theta + OmegaT2 b_s) / B^2]; real_buffer is not read */
int spk_aniso_flux(spk_solver s, const void *complex_fields, double *real_buffer, void *complex_dtheta);
/* max |x| over the physical points of one padded real field into a device
 * double: cublasIdamax + the read-back of the reference (compute_dt.cu:90-106)
 * without the host round trip; asynchronous */
int spk_maxabs(spk_solver s, const double *real_field, double *out_dev);
/* computeEnergy(data_type *vcomplex) (spooky_output_functions.cu:23-55); synchronizes */
int spk_energy_ptr(spk_solver s, const void *complex_field, double *out);
/* No counterpart in the reference.  spk_energy_ptr for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): computeEnergy of one spectral field,
 * [NX][NY][NZ/2+1] complex FP64 (16-byte aligned), into one device double
 * out_dev (8-byte aligned).  Needs no solver; allocates nothing; asynchronous
 * on the context stream, no synchronization.  Its reduction partials live in a
 * static device array of SPK_NOALLOC_PARTS doubles (8 KiB) that each GPU holds
 * until the process ends; a call whose partials would not fit is rejected
 * (the reduction strides over the field, so no grid needs more).  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU race. */
#define SPK_NOALLOC_PARTS 1024
int spk_energy_noalloc(spk_ctx ctx, int nx, int ny, int nz, const void *complex_field, double *out_dev);
/* computeEnstrophy(data_type *complex_vecField) (spooky_output_functions.cu:57-86) */
int spk_enstrophy_ptr(spk_solver s, const void *complex_vec, double *out);
/* twoFieldCorrelation(scalar_type *v1, scalar_type *v2) (spooky_output_functions.cu:188-235) */
int spk_correlation_ptr(spk_solver s, const double *real_a, const double *real_b, double *out);
/* computeSpectrum1d(v1, v2, ...) (inputoutput.hpp:30, spooky_output_spectrum.cu:424-444) */
int spk_spectrum_ptr(spk_solver s, const void *complex_a, const void *complex_b, double *out, int cap, int *nbins);

/* ---- batched ensemble: many independent boxes advanced together ----
 * The reference runs one box per process (Supervisor, supervisor.cu:20-213);
 * an ensemble holds nboxes solvers (box i: params[i], any grid and physics the
 * solver supports except shearing boxes, all in precision bulk) on device
 * `device`.  Boxes that differ only in nu, nu_m, nu_th and reynolds_ani form a
 * shape group, and a group is advanced as one batch: each kernel of the step
 * is launched once for all of its boxes (the box in blockIdx.z), so 48 boxes
 * at 64^3 are one launch per kernel.  Batches run on their own HIP streams.
 * `lanes` (0: 4) is the stream count of the box-by-box mode (split 0, box i on
 * lane i % lanes).  Boxes with super-time-stepping plan their RKL2 cycles on
 * the device, each with its own stage count. */
typedef struct spk_ensemble_s *spk_ensemble;
int spk_ensemble_create(int device, int nboxes, const spk_params *params, int bulk, int lanes, spk_ensemble *out);
int spk_ensemble_destroy(spk_ensemble e);
int spk_ensemble_size(spk_ensemble e, int *nboxes, int *lanes);
/* The solver of box i, for initial conditions, time state and diagnostics
 * (calls on it are ordered after the ensemble's queued steps). */
int spk_ensemble_box(spk_ensemble e, int i, spk_solver *out);
/* Every box advances nsteps (even) steps of the main loop (spk_main_loop) with
 * its own dt.  With the graph on (default) two steps of every box are captured
 * once into one HIP graph and each pair of steps is one graph launch; the graph
 * is captured again when a box was stepped outside it.  Asynchronous. */
int spk_ensemble_run(spk_ensemble e, int nsteps);
int spk_ensemble_sync(spk_ensemble e);
/* on = 0: the same steps launched box by box into the lanes, without a graph */
int spk_ensemble_set_graph(spk_ensemble e, int on);
/* box-by-box launches over `lanes` streams (split 0) */
int spk_ensemble_set_lanes(spk_ensemble e, int lanes);
/* split >= 1: each shape group advanced in `split` batches of consecutive boxes
 * (the default is 1: one batch per group); 0: box by box over the lanes */
int spk_ensemble_set_split(spk_ensemble e, int split);
int spk_ensemble_layout(spk_ensemble e, int *ngroups, int *split, int *nbatches);
/* Autotune of the launch layout: batched, over splits 1, 2, 4; box by box,
 * over 1, 2, 4, 8, 16 lanes.  Each candidate times `probe` (even) steps of
 * every box after a warm-up pair; every box's state and clock are saved before
 * and restored after each candidate.  Keeps the fastest; *best = its split
 * (batched) or lane count. */
int spk_ensemble_tune(spk_ensemble e, int probe, int *best, double *best_ms);

/* ---- learned subgrid-stress closure (no counterpart in the reference) ----
 * A 4-layer 3D convolution network on a periodic n^3 grid (n a multiple of 32):
 * the filtered velocity (3 channels) -> 32 -> 32 -> 32 channels with 3^3 taps
 * and GELU -> the 6 subgrid stresses tau_xx, tau_xy, tau_xz, tau_yy, tau_yz,
 * tau_zz.  BF16 weights and activations on the matrix cores, FP32
 * accumulation.  Parameters (FP32, spk_closure_num_params of them), layer by
 * layer l = 1..4: W_l[27][32][cin_l] then b_l[32], cin = 16, 32, 32, 32; tap
 * t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) multiplies the input at (x + dx,
 * y + dy, z + dz); input channels 3..15 of layer 1 and output channels 6..31
 * of layer 4 are unused.  Velocities and stresses are device arrays in the
 * padded real layout, FP64 [3 or 6][n][n][n + 2]; the network sees
 * u * in_scale and predicts tau / out_scale[c]. */
typedef struct spk_closure_s *spk_closure;
size_t spk_closure_num_params(void);
int spk_closure_create(spk_ctx ctx, int n, spk_closure *out); /* parameters start at 0 */
int spk_closure_destroy(spk_closure c);
int spk_closure_set_params(spk_closure c, const float *host_params); /* also resets Adam; synchronizes */
int spk_closure_get_params(spk_closure c, float *host_params);       /* synchronizes */
int spk_closure_set_scales(spk_closure c, double in_scale, const double *out_scale6);
int spk_closure_forward(spk_closure c, const void *u, void *tau);
/* gradient of the loss mean_{c < 6, points} (tau_pred / s_c - tau_true / s_c)^2
 * with respect to every parameter (readable with spk_closure_get_grads); loss
 * (optional) synchronizes */
int spk_closure_grads(spk_closure c, const void *u, const void *tau_true, double *loss);
int spk_closure_get_grads(spk_closure c, float *host_grads); /* synchronizes */
/* spk_closure_grads then one Adam step (beta 0.9, 0.999, eps 1e-8) at rate lr */
int spk_closure_train_step(spk_closure c, const void *u, const void *tau_true, double lr, double *loss);
/* The closure inline in a run on its grid (incompressible, same stream): after a
 * step, U_i += dt P(-i mask k_j tau_ij(u)), tau from the network on the current
 * real velocity, P the projection on k-perpendicular.  Overwrites the real buffer. */
int spk_closure_apply(spk_closure c, spk_solver s);
/* Training data from a fine run: with F the sharp cut to |k_i| <= kc,
 * kc = (2/3)(n/2 - 1) (the dealiasing cut of the coarse grid), the coarse
 * velocity ubar = F u and the subgrid stress tau_ij = F(u_i u_j) - F(ubar_i ubar_j)
 * (products on the fine grid), on the coarse grid n^3.  U_hi: device spectral FP64
 * velocity [3][N][N][N/2+1] (N a multiple of n); ubar_lo, tau_lo: device padded
 * real FP64 [3][n][n][n + 2] and [6][n][n][n + 2].  Allocates fine-grid scratch
 * (7 real fields of N^3) for the call; synchronizes. */
int spk_sgs_stress(spk_ctx ctx, const void *U_hi, int N, void *ubar_lo, void *tau_lo, int n);

#ifdef __cplusplus
}
#endif
#endif /* SPK_H */
