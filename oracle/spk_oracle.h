/*
 * spk_oracle.h -- FP64 CPU restatement of SPOOKY's pseudospectral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * solver: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (libspk) never links or calls it.
 *
 * Every routine restates one piece of the reference (LorenzoLMP/spooky-git,
 * src/...) and cites the file:line it follows.  Layouts are the reference's:
 *   spectral field  : complex [NX][NY][NZ/2+1], kz fastest, interleaved re/im
 *   real field      : double  [NX][NY][2*(NZ/2+1)] (2 padding reals per row)
 *   field n of a set: offset n*NTC complex (n*2*NTC reals)
 * Field order follows Parameters::popVariablesGrid (parameters.cu:568-654).
 */
#ifndef SPK_ORACLE_H
#define SPK_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters.  Mirrors the Parameters members read in parameters.cu:77-330
 * that the hot path uses.  Layout is shared, by construction, with spk_params
 * in include/spk.h (both are plain ints/doubles in the same order). */
typedef struct orc_params {
    int nx, ny, nz;
    int incompressible, heat_equation, boussinesq, stratification, strat_direction;
    int mhd, anisotropic_diffusion, supertimestepping, shearing, rotating;
    int with2d, antialiasing;
    int sts_mask;        /* bit f set: field f is an sts_variable (rkl2) */
    int safety_sts;
    double lx, ly, lz;
    double nu, nu_m, nu_th, reynolds_ani, N2, OmegaT2, omega, shear;
    double cfl, cfl_par, cfl_sts, safety_source;
    double t_initial, t_final, toutput_flow;
} orc_params;

size_t orc_params_size(void);
int orc_num_threads(void); /* OpenMP threads the oracle runs on */

/* ---- transforms (cufft_routines.cu:14-40) ---- */
/* forward R2C, unnormalized; in: padded real, out: complex; in may equal out */
void orc_fft_r2c(int nx, int ny, int nz, const double *in, double *out);
/* inverse C2R scaled by 1/(nx*ny*nz) (scaleKernel, cufft_routines.cu:30);
 * in may equal out; imaginary parts of the kz=0 and kz=nz/2 planes are ignored */
void orc_fft_c2r(int nx, int ny, int nz, const double *in, double *out);
/* 1D complex DFT along a contiguous line (sign -1 forward, +1 inverse) */
void orc_fft1d(double *line, int n, int sign);

/* ---- wavevector (wavevector.cu:22-112) ---- */
/* kvec: 3*NTC doubles, mask: NTC doubles; scal[6] = kxmax,kymax,kzmax,kmax,deltak,nbins */
void orc_wavevector(const orc_params *p, double *kvec, double *mask, double *scal);

/* ---- simulation (Fields + Physics + TimeStepping + RKLegendre) ---- */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_create(const orc_params *p);
void orc_sim_destroy(orc_sim *s);
int orc_sim_nfields(const orc_sim *s);
size_t orc_sim_ntc(const orc_sim *s);
/* field indices: out[0..9] = VX,VY,VZ,BX,BY,BZ,TH,VEL,MAG,NUM_FIELDS */
void orc_sim_vars(const orc_sim *s, int *out);
double *orc_sim_fields(orc_sim *s);    /* complex [nf][ntc] */
double *orc_sim_dfields(orc_sim *s);   /* complex [nf][ntc] */
double *orc_sim_realbuf(orc_sim *s);   /* real padded [nf][2 ntc] */
double *orc_sim_kvec(orc_sim *s);      /* current (possibly sheared) kvec */
/* time state: t, dt, tremap, dt_hyp, dt_par, dt_sts, t_lastsnap, step */
void orc_sim_get_time(const orc_sim *s, double *out8);
void orc_sim_set_time(orc_sim *s, double t, int step, double t_lastsnap);

void orc_clean_divergence(orc_sim *s);   /* symmetries.cu:114-130 */
void orc_complex2real(orc_sim *s);       /* supervisor.cu:261-272 */
void orc_compute_dt(orc_sim *s);         /* compute_dt.cu:15-309 */
void orc_compute_dfield(orc_sim *s, int stage_step); /* compute_dfield.cu:11-90 */
void orc_rk3(orc_sim *s);                /* hydro_mhd_advance.cu:82-189 */
void orc_rkl2_cycle(orc_sim *s);         /* rkl.cu:437-591 */
void orc_hydro_mhd_advance(orc_sim *s);   /* hydro_mhd_advance.cu:18-75 */
/* n iterations of Supervisor::executeMainLoop's body (supervisor.cu:183-190)
 * without file output; t_lastsnap follows CheckSnapshot (check_output.cu:62-94) */
void orc_main_loop(orc_sim *s, int nsteps);
void orc_shift_time(orc_sim *s);         /* shearing_remapping.cu:15-24 */
void orc_shear_wavevector(orc_sim *s, double tremap); /* wavevector.cu:162-172 */
void orc_remap_all_fields(orc_sim *s);   /* shearing_remapping.cu:26-59 */

/* RKL2 stage count and per-stage coefficients (rkl.cu:466-585).
 * coef[4*(j-1)+{0,1,2,3}] = mu_j, nu_j, mu_tilde_j, gamma_j for stage j=1..s
 * (stage 1 stores mu_tilde_1 only).  Returns s.  coef may be NULL. */
int orc_rkl2_coeffs(double dt, double dt_sts, double *coef, int cap);

/* ---- diagnostics (spooky_output_functions.cu) ---- */
double orc_energy(orc_sim *s, int field);              /* :23-55 */
double orc_enstrophy(orc_sim *s, int vec_first_field); /* :57-86 */
double orc_correlation(orc_sim *s, int fa, int fb);    /* :188-235, on realbuf */
/* shell spectrum Re(v1 v2*)/N^2 (cuda_kernels_generic.cu:441-471); returns nbins+1 */
int orc_spectrum(orc_sim *s, int fa, int fb, double *out, int cap);

#ifdef __cplusplus
}
#endif
#endif
