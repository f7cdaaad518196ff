/*
 * spk_oracle.c -- FP64 CPU restatement of SPOOKY's hot path (TEST INFRASTRUCTURE).
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the bench
 * load this file's library.  The product never does.  Each function cites the
 * reference routine (LorenzoLMP/spooky-git, src/...) whose arithmetic it
 * restates; operation order inside each elementwise expression follows the
 * reference kernel so that FP64 results agree to a few ulp.
 *
 * The 3D transform is an independent radix-2 Cooley-Tukey implementation (the
 * reference calls closed cuFFT); it is pinned by numpy rfftn/irfftn golden
 * vectors, a brute-force DFT and closed forms (tests/test_oracle_fft.py).
 */
#include "spk_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_L 3.141592653589793238462643383279502884L

size_t orc_params_size(void) { return sizeof(orc_params); }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* 1D FFT: iterative radix-2, twiddles from long-double sin/cos               */
/* ------------------------------------------------------------------------- */

#define ORC_MAXLOG 16
static double *g_tw[ORC_MAXLOG + 1]; /* g_tw[l][2k], [2k+1] = cos, sin(2 pi k / 2^l) */

static int ilog2i(int n) {
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

static void ensure_tw(int n) {
    int l = ilog2i(n);
    if (l > ORC_MAXLOG || g_tw[l]) return;
#pragma omp critical(orc_tw)
    {
        if (!g_tw[l]) {
            double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
            for (int k = 0; k < n; k++) {
                long double a = 2.0L * PI_L * (long double)k / (long double)n;
                t[2 * k] = (double)cosl(a);
                t[2 * k + 1] = (double)sinl(a);
            }
            g_tw[l] = t;
        }
    }
}

void orc_fft1d(double *a, int n, int sign) {
    if (n <= 1) return;
    ensure_tw(n);
    const double *tw = g_tw[ilog2i(n)];
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = a[2 * i], ti = a[2 * i + 1];
            a[2 * i] = a[2 * j];
            a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = tr;
            a[2 * j + 1] = ti;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                double wr = tw[2 * k * step], wi = sign * tw[2 * k * step + 1];
                double *u = a + 2 * (i + k), *v = a + 2 * (i + k + half);
                double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - vr;
                v[1] = u[1] - vi;
                u[0] += vr;
                u[1] += vi;
            }
        }
    }
}

/* strided complex lines of one 3D array: apply orc_fft1d along axis ax (0=x,1=y) */
static void fft_axis(double *c, int nx, int ny, int nzh, int ax, int sign) {
    int n = ax == 0 ? nx : ny;
    ensure_tw(n);
    long nlines = ax == 0 ? (long)ny * nzh : (long)nx * nzh;
    size_t stride = ax == 0 ? (size_t)ny * nzh : (size_t)nzh;
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
        for (long L = 0; L < nlines; L++) {
            size_t base;
            if (ax == 0) {
                base = (size_t)L; /* (j, k) flattened */
            } else {
                long i = L / nzh, k = L % nzh;
                base = (size_t)i * ny * nzh + (size_t)k;
            }
            for (int m = 0; m < n; m++) {
                line[2 * m] = c[2 * (base + m * stride)];
                line[2 * m + 1] = c[2 * (base + m * stride) + 1];
            }
            orc_fft1d(line, n, sign);
            for (int m = 0; m < n; m++) {
                c[2 * (base + m * stride)] = line[2 * m];
                c[2 * (base + m * stride) + 1] = line[2 * m + 1];
            }
        }
        free(line);
    }
}

/* cufft_routines.cu:14-20 (r2c_fft): cufftXtExec D2Z, forward, unnormalized */
void orc_fft_r2c(int nx, int ny, int nz, const double *in, double *out) {
    int nzh = nz / 2 + 1;
    size_t ntc = (size_t)nx * ny * nzh;
    double *c = (double *)malloc(sizeof(double) * 2 * ntc);
    ensure_tw(nz);
    long rows = (long)nx * ny;
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * (size_t)nz);
#pragma omp for schedule(static)
        for (long r = 0; r < rows; r++) {
            const double *src = in + (size_t)r * 2 * nzh;
            for (int m = 0; m < nz; m++) {
                line[2 * m] = src[m];
                line[2 * m + 1] = 0.0;
            }
            orc_fft1d(line, nz, -1);
            memcpy(c + (size_t)r * 2 * nzh, line, sizeof(double) * 2 * nzh);
        }
        free(line);
    }
    fft_axis(c, nx, ny, nzh, 1, -1);
    fft_axis(c, nx, ny, nzh, 0, -1);
    memcpy(out, c, sizeof(double) * 2 * ntc);
    free(c);
}

/* cufft_routines.cu:23-40 (c2r_fft): scaleKernel by 1/(NX NY NZ), then Z2D */
void orc_fft_c2r(int nx, int ny, int nz, const double *in, double *out) {
    int nzh = nz / 2 + 1;
    size_t ntc = (size_t)nx * ny * nzh;
    double scale = 1.0 / ((double)nx * ny * nz);
    double *c = (double *)malloc(sizeof(double) * 2 * ntc);
    for (size_t i = 0; i < 2 * ntc; i++) c[i] = in[i] * scale;
    fft_axis(c, nx, ny, nzh, 0, +1);
    fft_axis(c, nx, ny, nzh, 1, +1);
    ensure_tw(nz);
    long rows = (long)nx * ny;
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * (size_t)nz);
#pragma omp for schedule(static)
        for (long r = 0; r < rows; r++) {
            const double *src = c + (size_t)r * 2 * nzh;
            /* Hermitian completion of the row; only the real part is kept, so the
             * imaginary parts of kz=0 and kz=nz/2 drop out */
            for (int k = 0; k < nzh; k++) {
                line[2 * k] = src[2 * k];
                line[2 * k + 1] = src[2 * k + 1];
            }
            for (int k = 1; k < nz - nzh + 1; k++) {
                line[2 * (nz - k)] = src[2 * k];
                line[2 * (nz - k) + 1] = -src[2 * k + 1];
            }
            orc_fft1d(line, nz, +1);
            double *dst = out + (size_t)r * 2 * nzh;
            for (int m = 0; m < nz; m++) dst[m] = line[2 * m];
            for (int m = nz; m < 2 * nzh; m++) dst[m] = 0.0;
        }
        free(line);
    }
    free(c);
}

/* ------------------------------------------------------------------------- */
/* wavevector.cu:22-112                                                       */
/* ------------------------------------------------------------------------- */
void orc_wavevector(const orc_params *p, double *kvec, double *mask, double *scal) {
    int NX = p->nx, NY = p->ny, NZ = p->nz, nzh = NZ / 2 + 1;
    size_t ntc = (size_t)NX * NY * nzh;
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
            for (int k = 0; k < nzh; k++) {
                size_t idx = (size_t)k + (size_t)nzh * (j + (size_t)i * NY);
                kvec[idx] = (2.0 * M_PI) / p->lx * (fmod((double)i + ((double)NX / 2), NX) - (double)NX / 2);
                kvec[ntc + idx] = (2.0 * M_PI) / p->ly * (fmod((double)j + ((double)NY / 2), NY) - (double)NY / 2);
                kvec[2 * ntc + idx] = p->with2d ? 0.0 : (2.0 * M_PI) / p->lz * (double)k;
            }
    double kxmax = 2.0 * M_PI / p->lx * ((NX / 2) - 1);
    double kymax = 2.0 * M_PI / p->ly * ((NY / 2) - 1);
    double kzmax = 2.0 * M_PI / p->lz * ((NZ / 2) - 1);
    if (p->with2d) kzmax = 0.0;
    for (size_t idx = 0; idx < ntc; idx++) mask[idx] = 1.0;
    if (p->antialiasing) {
        for (size_t idx = 0; idx < ntc; idx++) {
            mask[idx] = 1.0;
            if (fabs(kvec[idx]) > 2.0 / 3.0 * kxmax) mask[idx] = 0.0;
            if (fabs(kvec[ntc + idx]) > 2.0 / 3.0 * kymax) mask[idx] = 0.0;
            if (fabs(kvec[2 * ntc + idx]) > 2.0 / 3.0 * kzmax) mask[idx] = 0.0;
        }
        kxmax = (2.0 / 3.0) * kxmax;
        kymax = (2.0 / 3.0) * kymax;
        kzmax = (2.0 / 3.0) * kzmax;
    }
    double kmax = pow(kxmax * kxmax + kymax * kymax + kzmax * kzmax, 0.5);
    double lmin = p->lx;
    if (p->ly < lmin) lmin = p->ly;
    if (p->lz < lmin) lmin = p->lz;
    double deltak = 2.0 * M_PI / lmin;
    int nbins = (int)(kmax / deltak + 0.5);
    if (scal) {
        scal[0] = kxmax;
        scal[1] = kymax;
        scal[2] = kzmax;
        scal[3] = kmax;
        scal[4] = deltak;
        scal[5] = (double)nbins;
    }
}

/* ------------------------------------------------------------------------- */
/* simulation state                                                           */
/* ------------------------------------------------------------------------- */
#define NTMP 6
struct orc_sim {
    orc_params p;
    int VX, VY, VZ, BX, BY, BZ, TH, VEL, MAG, NF;
    size_t ntc, ntot;
    double *kvec, *kvec0, *mask;
    double kxmax, kymax, kzmax, kmax, deltak;
    int nbins;
    double *fields, *dfields, *scr, *realbuf, *tmp;
    int nsts, sts_index[8], sts_pos[8];
    double *dU, *dU0, *Uc0, *Uc1;
    double t, dt, dt_hyp, dt_par, dt_sts, tremap, t_lastsnap;
    int step, stage_step, loop_started;
};

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

orc_sim *orc_sim_create(const orc_params *p) {
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->p = *p;
    /* parameters.cu:568-640 */
    s->NF = 0;
    if (p->heat_equation) {
        s->TH = 0;
        s->NF = 1;
    } else if (p->incompressible) {
        s->VX = 0; s->VY = 1; s->VZ = 2; s->VEL = 0; s->NF = 3;
        if (p->mhd) {
            s->BX = 3; s->BY = 4; s->BZ = 5; s->MAG = 3; s->NF += 3;
            if (p->boussinesq) { s->TH = 6; s->NF += 1; }
        } else if (p->boussinesq) {
            s->TH = 3; s->NF += 1;
        }
    }
    int nzh = p->nz / 2 + 1;
    s->ntc = (size_t)p->nx * p->ny * nzh;
    s->ntot = (size_t)p->nx * p->ny * p->nz;
    s->kvec = zalloc(3 * s->ntc);
    s->kvec0 = zalloc(3 * s->ntc);
    s->mask = zalloc(s->ntc);
    double scal[6];
    orc_wavevector(p, s->kvec0, s->mask, scal);
    memcpy(s->kvec, s->kvec0, sizeof(double) * 3 * s->ntc);
    s->kxmax = scal[0]; s->kymax = scal[1]; s->kzmax = scal[2];
    s->kmax = scal[3]; s->deltak = scal[4]; s->nbins = (int)scal[5];
    size_t fsz = 2 * s->ntc * (size_t)s->NF;
    s->fields = zalloc(fsz);
    s->dfields = zalloc(fsz);
    s->scr = zalloc(fsz);
    s->realbuf = zalloc(fsz);
    s->tmp = zalloc(2 * s->ntc * NTMP);
    /* rkl.cu:40-57: sts_variables_index / sts_variables_pos */
    s->nsts = 0;
    for (int f = 0; f < s->NF; f++) {
        s->sts_pos[f] = -1;
        if (p->supertimestepping && (p->sts_mask >> f & 1)) {
            s->sts_index[s->nsts] = f;
            s->sts_pos[f] = s->nsts++;
        }
    }
    size_t ssz = 2 * s->ntc * (size_t)(s->nsts ? s->nsts : 1);
    s->dU = zalloc(ssz); s->dU0 = zalloc(ssz); s->Uc0 = zalloc(ssz); s->Uc1 = zalloc(ssz);
    s->t = 0.0; /* timestepping.cu:27 */
    return s;
}

void orc_sim_destroy(orc_sim *s) {
    if (!s) return;
    free(s->kvec); free(s->kvec0); free(s->mask);
    free(s->fields); free(s->dfields); free(s->scr); free(s->realbuf); free(s->tmp);
    free(s->dU); free(s->dU0); free(s->Uc0); free(s->Uc1);
    free(s);
}

int orc_sim_nfields(const orc_sim *s) { return s->NF; }
size_t orc_sim_ntc(const orc_sim *s) { return s->ntc; }
void orc_sim_vars(const orc_sim *s, int *o) {
    o[0] = s->VX; o[1] = s->VY; o[2] = s->VZ; o[3] = s->BX; o[4] = s->BY;
    o[5] = s->BZ; o[6] = s->TH; o[7] = s->VEL; o[8] = s->MAG; o[9] = s->NF;
}
double *orc_sim_fields(orc_sim *s) { return s->fields; }
double *orc_sim_dfields(orc_sim *s) { return s->dfields; }
double *orc_sim_realbuf(orc_sim *s) { return s->realbuf; }
double *orc_sim_kvec(orc_sim *s) { return s->kvec; }
void orc_sim_get_time(const orc_sim *s, double *o) {
    o[0] = s->t; o[1] = s->dt; o[2] = s->tremap; o[3] = s->dt_hyp;
    o[4] = s->dt_par; o[5] = s->dt_sts; o[6] = s->t_lastsnap; o[7] = s->step;
}
void orc_sim_set_time(orc_sim *s, double t, int step, double tls) {
    s->t = t; s->step = step; s->t_lastsnap = tls;
}

/* complex helpers: pointer to element i of complex array a */
#define RE(a, i) (a)[2 * (i)]
#define IM(a, i) (a)[2 * (i) + 1]
#define CF(s, f) ((s)->fields + 2 * (s)->ntc * (size_t)(f))
#define DF(s, f) ((s)->dfields + 2 * (s)->ntc * (size_t)(f))
#define RB(s, f) ((s)->realbuf + 2 * (s)->ntc * (size_t)(f))
#define TMPC(s, n) ((s)->tmp + 2 * (s)->ntc * (size_t)(n))
#define KX(s) ((s)->kvec)
#define KY(s) ((s)->kvec + (s)->ntc)
#define KZ(s) ((s)->kvec + 2 * (s)->ntc)

/* ------------------------------------------------------------------------- */
/* elementwise kernels (cuda_kernels.cu, cuda_kernels_generic.cu)             */
/* ------------------------------------------------------------------------- */

/* cuda_kernels_generic.cu:288-313 CleanDivergence (in place) */
static void k_clean_divergence(orc_sim *s, double *X) {
    size_t N = s->ntc;
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s);
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        /* q0 = I * (kx X0 + ky X1 + kz X2) */
        double sr = kx[i] * RE(X, i) + ky[i] * RE(X, N + i) + kz[i] * RE(X, 2 * N + i);
        double si = kx[i] * IM(X, i) + ky[i] * IM(X, N + i) + kz[i] * IM(X, 2 * N + i);
        double q0r = -si, q0i = sr;
        double k2 = kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i];
        double ik2 = k2 > 0.0 ? 1.0 / k2 : 1.0;
        const double *kk[3] = {kx, ky, kz};
        for (int c = 0; c < 3; c++) {
            /* Z = X + I k_c q0 ik2 */
            double tr = kk[c][i] * q0r, ti = kk[c][i] * q0i; /* k q0 */
            double ur = -ti, ui = tr;                         /* I k q0 */
            RE(X, c * N + i) = RE(X, c * N + i) + ur * ik2;
            IM(X, c * N + i) = IM(X, c * N + i) + ui * ik2;
        }
    }
}

/* symmetries.cu:114-130 */
void orc_clean_divergence(orc_sim *s) {
    if (s->p.incompressible) k_clean_divergence(s, CF(s, s->VEL));
    if (s->p.mhd) k_clean_divergence(s, CF(s, s->MAG));
}

/* supervisor.cu:261-272: copy complex fields to the real buffer, C2R each in place */
static void complex2real_n(orc_sim *s, const double *cin, double *rout, int n) {
    const orc_params *p = &s->p;
    for (int f = 0; f < n; f++)
        orc_fft_c2r(p->nx, p->ny, p->nz, cin + 2 * s->ntc * (size_t)f, rout + 2 * s->ntc * (size_t)f);
}
void orc_complex2real(orc_sim *s) { complex2real_n(s, s->fields, s->realbuf, s->NF); }

static void r2c_inplace(orc_sim *s, double *buf) {
    orc_fft_r2c(s->p.nx, s->p.ny, s->p.nz, buf, buf);
}

/* max |f| over the physical points of a padded real field (compute_dt.cu:90-106
 * uses cublasIdamax over 2*NTC including the padding; padding is excluded here) */
static double maxabs_real(orc_sim *s, const double *r) {
    int nz = s->p.nz, nzp = 2 * (nz / 2 + 1);
    long rows = (long)s->p.nx * s->p.ny;
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long row = 0; row < rows; row++)
        for (int k = 0; k < nz; k++) {
            double a = fabs(r[(size_t)row * nzp + k]);
            if (a > m) m = a;
        }
    return m;
}

/* compute_dt.cu:15-309 */
void orc_compute_dt(orc_sim *s) {
    const orc_params *p = &s->p;
    s->dt_hyp = 0.0; s->dt_par = 0.0; s->dt_sts = 0.0;
    double gamma_v = 0.0, gamma_th = 0.0, gamma_par = 0.0, gamma_b = 0.0;
    double gamma_tmp = 0.0, gamma_sts = 0.0;
    double kxmax = s->kxmax, kymax = s->kymax, kzmax = s->kzmax;
    double tremap = s->tremap;
#define STSPOS(f) (s->sts_pos[f])
    if (p->heat_equation) {
        gamma_tmp = ((kxmax) * (kxmax) + kymax * kymax + kzmax * kzmax) * p->nu_th;
        if (STSPOS(s->TH) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        s->dt_par = p->cfl * p->cfl_par / gamma_par;
        s->dt_sts = p->cfl_par / gamma_sts;
        s->dt = s->dt_par;
    }
    if (p->incompressible) {
        orc_complex2real(s);
        double maxfx = maxabs_real(s, RB(s, s->VX));
        double maxfy = maxabs_real(s, RB(s, s->VY));
        double maxfz = maxabs_real(s, RB(s, s->VZ));
        gamma_v = ((kxmax + fabs(tremap) * kymax) * maxfx + kymax * maxfy + kzmax * maxfz);
        if (p->rotating) gamma_v += fabs(p->omega) / p->safety_source;
        if (p->shearing) gamma_v += fabs(p->shear) / p->safety_source;
        gamma_tmp = (kxmax + fabs(tremap) * kymax) * (kxmax + fabs(tremap) * kymax) * p->nu;
        if (STSPOS(s->VX) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        gamma_tmp = kymax * kymax * p->nu;
        if (STSPOS(s->VY) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        gamma_tmp = kzmax * kzmax * p->nu;
        if (STSPOS(s->VZ) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        if (p->boussinesq) {
            gamma_th += pow(fabs(p->N2), 0.5) / p->safety_source;
            if (p->anisotropic_diffusion) {
                gamma_th += pow(fabs(p->OmegaT2), 0.5) / p->safety_source;
                gamma_tmp = ((kxmax + fabs(tremap) * kymax) * (kxmax + fabs(tremap) * kymax) + kymax * kymax + kzmax * kzmax) * (1. / p->reynolds_ani);
            } else {
                gamma_tmp = ((kxmax + fabs(tremap) * kymax) * (kxmax + fabs(tremap) * kymax) + kymax * kymax + kzmax * kzmax) * p->nu_th;
            }
            if (STSPOS(s->TH) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        }
        if (p->mhd) {
            double maxbx = maxabs_real(s, RB(s, s->BX));
            double maxby = maxabs_real(s, RB(s, s->BY));
            double maxbz = maxabs_real(s, RB(s, s->BZ));
            gamma_b = (kxmax + fabs(tremap) * kymax) * maxbx + kymax * maxby + kzmax * maxbz;
            gamma_tmp = (kxmax + fabs(tremap) * kymax) * (kxmax + fabs(tremap) * kymax) * p->nu_m;
            if (STSPOS(s->BX) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
            gamma_tmp = kymax * kymax * p->nu_m;
            if (STSPOS(s->BY) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
            gamma_tmp = kzmax * kzmax * p->nu_m;
            if (STSPOS(s->BZ) < 0) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        }
        s->dt_hyp = p->cfl / (gamma_v + gamma_th + gamma_b);
        s->dt_par = p->cfl_par / gamma_par;
        s->dt_sts = p->cfl_sts / gamma_sts;
        s->dt = (s->dt_hyp < s->dt_par) ? s->dt_hyp : s->dt_par;
        if (p->supertimestepping) {
            if (s->dt > s->dt_sts * p->safety_sts) s->dt = s->dt_sts * p->safety_sts;
        }
    }
#undef STSPOS
    if (s->t + s->dt > p->t_final) {
        s->dt = p->t_final - s->t;
    } else if (s->t + s->dt - s->t_lastsnap > p->toutput_flow) {
        s->dt = p->toutput_flow - s->t + s->t_lastsnap;
    }
    if (p->supertimestepping) {
        if (s->dt_sts > s->dt) s->dt_sts = s->dt;
    }
}

/* hydro.cu:19-75 BasdevantHydro: TracelessShearMatrix[MHD] (cuda_kernels.cu:39-55,
 * 216-234) on the padded real buffer, 5 R2C, NonLinHydroAdv (cuda_kernels.cu:82-98) */
static void basdevant(orc_sim *s, double *dVel) {
    size_t N2 = 2 * s->ntc, N = s->ntc;
    const double *V = RB(s, s->VEL);
    const double *B = s->p.mhd ? RB(s, s->MAG) : NULL;
    double *S = s->tmp; /* real view, 5 arrays of 2N */
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N2; li++) {
        size_t i = (size_t)li;
        if (B) {
            S[i] = (2.0 * V[i] * V[i] - V[N2 + i] * V[N2 + i] - V[2 * N2 + i] * V[2 * N2 + i]
                    - 2.0 * B[i] * B[i] + B[N2 + i] * B[N2 + i] + B[2 * N2 + i] * B[2 * N2 + i]) / 3.0;
            S[N2 + i] = V[i] * V[N2 + i] - B[i] * B[N2 + i];
            S[2 * N2 + i] = V[i] * V[2 * N2 + i] - B[i] * B[2 * N2 + i];
            S[3 * N2 + i] = (-V[i] * V[i] + 2.0 * V[N2 + i] * V[N2 + i] - V[2 * N2 + i] * V[2 * N2 + i]
                             + B[i] * B[i] - 2.0 * B[N2 + i] * B[N2 + i] + B[2 * N2 + i] * B[2 * N2 + i]) / 3.0;
            S[4 * N2 + i] = V[N2 + i] * V[2 * N2 + i] - B[N2 + i] * B[2 * N2 + i];
        } else {
            S[i] = (2.0 * V[i] * V[i] - V[N2 + i] * V[N2 + i] - V[2 * N2 + i] * V[2 * N2 + i]) / 3.0;
            S[N2 + i] = V[i] * V[N2 + i];
            S[2 * N2 + i] = V[i] * V[2 * N2 + i];
            S[3 * N2 + i] = (-V[i] * V[i] + 2.0 * V[N2 + i] * V[N2 + i] - V[2 * N2 + i] * V[2 * N2 + i]) / 3.0;
            S[4 * N2 + i] = V[N2 + i] * V[2 * N2 + i];
        }
    }
    for (int n = 0; n < 5; n++) r2c_inplace(s, TMPC(s, n));
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s), *mask = s->mask;
    const double *Sc = s->tmp;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        double m = mask[i];
        /* dU_x = -I m (kx Sxx + ky Sxy + kz Sxz) */
        double ar = kx[i] * RE(Sc, i) + ky[i] * RE(Sc, N + i) + kz[i] * RE(Sc, 2 * N + i);
        double ai = kx[i] * IM(Sc, i) + ky[i] * IM(Sc, N + i) + kz[i] * IM(Sc, 2 * N + i);
        RE(dVel, i) = m * ai; IM(dVel, i) = -m * ar;
        ar = kx[i] * RE(Sc, N + i) + ky[i] * RE(Sc, 3 * N + i) + kz[i] * RE(Sc, 4 * N + i);
        ai = kx[i] * IM(Sc, N + i) + ky[i] * IM(Sc, 3 * N + i) + kz[i] * IM(Sc, 4 * N + i);
        RE(dVel, N + i) = m * ai; IM(dVel, N + i) = -m * ar;
        ar = kx[i] * RE(Sc, 2 * N + i) + ky[i] * RE(Sc, 4 * N + i) - kz[i] * (RE(Sc, i) + RE(Sc, 3 * N + i));
        ai = kx[i] * IM(Sc, 2 * N + i) + ky[i] * IM(Sc, 4 * N + i) - kz[i] * (IM(Sc, i) + IM(Sc, 3 * N + i));
        RE(dVel, 2 * N + i) = m * ai; IM(dVel, 2 * N + i) = -m * ar;
    }
}

/* mhd.cu:20-62 CurlEMF: MagneticEmf (cuda_kernels.cu:240-252), 3 R2C,
 * MagneticShear (cuda_kernels.cu:257-272) */
static void curl_emf(orc_sim *s, double *dMag) {
    size_t N2 = 2 * s->ntc, N = s->ntc;
    const double *V = RB(s, s->VEL), *B = RB(s, s->MAG);
    double *E = s->tmp;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N2; li++) {
        size_t i = (size_t)li;
        E[i] = -B[N2 + i] * V[2 * N2 + i] + B[2 * N2 + i] * V[N2 + i];
        E[N2 + i] = B[i] * V[2 * N2 + i] - B[2 * N2 + i] * V[i];
        E[2 * N2 + i] = -B[i] * V[N2 + i] + B[N2 + i] * V[i];
    }
    for (int n = 0; n < 3; n++) r2c_inplace(s, TMPC(s, n));
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s), *mask = s->mask;
    const double *Ec = s->tmp;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        double m = mask[i];
        /* dB_x = I m (ky Ez - kz Ey) */
        double ar = ky[i] * RE(Ec, 2 * N + i) - kz[i] * RE(Ec, N + i);
        double ai = ky[i] * IM(Ec, 2 * N + i) - kz[i] * IM(Ec, N + i);
        RE(dMag, i) = -m * ai; IM(dMag, i) = m * ar;
        ar = kz[i] * RE(Ec, i) - kx[i] * RE(Ec, 2 * N + i);
        ai = kz[i] * IM(Ec, i) - kx[i] * IM(Ec, 2 * N + i);
        RE(dMag, N + i) = -m * ai; IM(dMag, N + i) = m * ar;
        ar = kx[i] * RE(Ec, N + i) - ky[i] * RE(Ec, i);
        ai = kx[i] * IM(Ec, N + i) - ky[i] * IM(Ec, i);
        RE(dMag, 2 * N + i) = -m * ai; IM(dMag, 2 * N + i) = m * ar;
    }
}

/* boussinesq.cu:10-42 AdvectTemperature: EnergyFluxVector (cuda_kernels.cu:277-289),
 * 3 R2C, NonLinBoussinesqAdv (cuda_kernels.cu:293-304) */
static void advect_temperature(orc_sim *s, double *dTh) {
    size_t N2 = 2 * s->ntc, N = s->ntc;
    const double *V = RB(s, s->VEL), *Th = RB(s, s->TH);
    double *E = s->tmp;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N2; li++) {
        size_t i = (size_t)li;
        E[i] = V[i] * Th[i];
        E[N2 + i] = V[N2 + i] * Th[i];
        E[2 * N2 + i] = V[2 * N2 + i] * Th[i];
    }
    for (int n = 0; n < 3; n++) r2c_inplace(s, TMPC(s, n));
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s), *mask = s->mask;
    const double *Ec = s->tmp;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        double m = mask[i];
        double ar = kx[i] * RE(Ec, i) + ky[i] * RE(Ec, N + i) + kz[i] * RE(Ec, 2 * N + i);
        double ai = kx[i] * IM(Ec, i) + ky[i] * IM(Ec, N + i) + kz[i] * IM(Ec, 2 * N + i);
        RE(dTh, i) = m * ai; IM(dTh, i) = -m * ar;
    }
}

/* anisotropic_conduction.cu:10-62 AnisotropicHeatFlux (adds into dTh) */
static void anisotropic_heat_flux(orc_sim *s, double *dTh) {
    size_t N2 = 2 * s->ntc, N = s->ntc;
    const orc_params *p = &s->p;
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s), *mask = s->mask;
    const double *Th = CF(s, s->TH);
    double *G = s->tmp;
    /* Gradient (cuda_kernels_generic.cu:249-260) */
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        RE(G, i) = -kx[i] * IM(Th, i); IM(G, i) = kx[i] * RE(Th, i);
        RE(G, N + i) = -ky[i] * IM(Th, i); IM(G, N + i) = ky[i] * RE(Th, i);
        RE(G, 2 * N + i) = -kz[i] * IM(Th, i); IM(G, 2 * N + i) = kz[i] * RE(Th, i);
    }
    for (int n = 0; n < 3; n++) orc_fft_c2r(p->nx, p->ny, p->nz, TMPC(s, n), TMPC(s, n));
    const double *B = RB(s, s->MAG);
    double *bg = TMPC(s, 3); /* real view */
    double *Z = s->tmp;
    int sd = p->strat_direction;
    double chi = 1. / p->reynolds_ani, OmegaT2 = p->OmegaT2;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N2; li++) {
        size_t i = (size_t)li;
        /* ComputeBGradTheta (cuda_kernels.cu:360-367) */
        bg[i] = B[i] * G[i] + B[N2 + i] * G[N2 + i] + B[2 * N2 + i] * G[2 * N2 + i];
    }
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N2; li++) {
        size_t i = (size_t)li;
        /* ComputeAnisotropicHeatFlux (cuda_kernels.cu:371-383) */
        double B2 = B[i] * B[i] + B[N2 + i] * B[N2 + i] + B[2 * N2 + i] * B[2 * N2 + i];
        double bs = B[sd * N2 + i];
        Z[i] = chi * B[i] * (bg[i] + OmegaT2 * bs) / B2;
        Z[N2 + i] = chi * B[N2 + i] * (bg[i] + OmegaT2 * bs) / B2;
        Z[2 * N2 + i] = chi * B[2 * N2 + i] * (bg[i] + OmegaT2 * bs) / B2;
    }
    for (int n = 0; n < 3; n++) r2c_inplace(s, TMPC(s, n));
    const double *Q = s->tmp;
    /* DivergenceMask flag 1 (cuda_kernels_generic.cu:315-334) */
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        double m = mask[i];
        double ar = kx[i] * RE(Q, i) + ky[i] * RE(Q, N + i) + kz[i] * RE(Q, 2 * N + i);
        double ai = kx[i] * IM(Q, i) + ky[i] * IM(Q, N + i) + kz[i] * IM(Q, 2 * N + i);
        RE(dTh, i) += -m * ai; IM(dTh, i) += m * ar;
    }
}

/* nablaOpScalar (cuda_kernels_generic.cu:198-214): Z (=|+=) -a k^2 X */
static void nabla_scalar(orc_sim *s, const double *X, double *Z, double a, int flag) {
    size_t N = s->ntc;
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s);
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        double f = -a * (kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i]);
        if (flag == 0) {
            RE(Z, i) = f * RE(X, i); IM(Z, i) = f * IM(X, i);
        } else {
            RE(Z, i) += f * RE(X, i); IM(Z, i) += f * IM(X, i);
        }
    }
}

/* physics.cu:30-103 HyperbolicTerms */
static void hyperbolic_terms(orc_sim *s) {
    if (s->p.incompressible) basdevant(s, DF(s, s->VEL));
    if (s->p.mhd) curl_emf(s, DF(s, s->MAG));
    if (s->p.boussinesq) advect_temperature(s, DF(s, s->TH));
}

/* physics.cu:105-142 SourceTerms; boussinesq.cu:45-69, shear_rotation.cu:11-63 */
static void source_terms(orc_sim *s) {
    const orc_params *p = &s->p;
    size_t N = s->ntc;
    if (p->stratification) {
        /* BoussinesqStrat (cuda_kernels.cu:310-325) */
        double *dV = DF(s, s->VEL), *dT = DF(s, s->TH);
        const double *V = CF(s, s->VEL), *T = CF(s, s->TH);
        int sd = p->strat_direction;
#pragma omp parallel for schedule(static)
        for (long li = 0; li < (long)N; li++) {
            size_t i = (size_t)li;
            RE(dV, sd * N + i) += -RE(T, i); IM(dV, sd * N + i) += -IM(T, i);
            RE(dT, i) += p->N2 * RE(V, sd * N + i); IM(dT, i) += p->N2 * IM(V, sd * N + i);
        }
    }
    if (p->shearing) {
        /* ShearingFlow (cuda_kernels.cu:332-340) */
        double *dvy = DF(s, s->VY);
        const double *vx = CF(s, s->VX);
        for (size_t i = 0; i < N; i++) { RE(dvy, i) += p->shear * RE(vx, i); IM(dvy, i) += p->shear * IM(vx, i); }
        if (p->mhd) {
            double *dby = DF(s, s->BY);
            const double *bx = CF(s, s->BX);
            for (size_t i = 0; i < N; i++) { RE(dby, i) += -p->shear * RE(bx, i); IM(dby, i) += -p->shear * IM(bx, i); }
        }
    }
    if (p->rotating) {
        /* BackgroundRotation is gated on shearing (shear_rotation.cu:47);
         * CoriolisForce (cuda_kernels.cu:344-355) */
        if (p->shearing) {
            const double *vx = CF(s, s->VX), *vy = CF(s, s->VY);
            double *dvx = DF(s, s->VX), *dvy = DF(s, s->VY);
            for (size_t i = 0; i < N; i++) {
                RE(dvx, i) += 2.0 * p->omega * RE(vy, i); IM(dvx, i) += 2.0 * p->omega * IM(vy, i);
                RE(dvy, i) -= 2.0 * p->omega * RE(vx, i); IM(dvy, i) -= 2.0 * p->omega * IM(vx, i);
            }
        }
    }
}

/* physics.cu:145-230 ParabolicTerms */
static void parabolic_terms(orc_sim *s) {
    const orc_params *p = &s->p;
    if (p->incompressible)
        for (int c = 0; c < 3; c++)
            if (s->sts_pos[s->VEL + c] < 0) nabla_scalar(s, CF(s, s->VEL + c), DF(s, s->VEL + c), p->nu, 1);
    if (p->mhd)
        for (int c = 0; c < 3; c++)
            if (s->sts_pos[s->MAG + c] < 0) nabla_scalar(s, CF(s, s->MAG + c), DF(s, s->MAG + c), p->nu_m, 1);
    if (p->boussinesq || p->heat_equation) {
        if (s->sts_pos[s->TH] < 0) {
            if (p->anisotropic_diffusion) anisotropic_heat_flux(s, DF(s, s->TH));
            else nabla_scalar(s, CF(s, s->TH), DF(s, s->TH), p->nu_th, p->heat_equation ? 0 : 1);
        }
    }
}

/* GradPseudoPressure[Shearing] (cuda_kernels.cu:128-197) */
static void grad_pseudo_pressure(orc_sim *s) {
    size_t N = s->ntc;
    double *dV = DF(s, s->VEL);
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s);
    const double *vx = CF(s, s->VX);
    int sh = s->p.shearing;
    double shear = s->p.shear;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        double sr = kx[i] * RE(dV, i) + ky[i] * RE(dV, N + i) + kz[i] * RE(dV, 2 * N + i);
        double si = kx[i] * IM(dV, i) + ky[i] * IM(dV, N + i) + kz[i] * IM(dV, 2 * N + i);
        if (sh) {
            sr = shear * ky[i] * RE(vx, i) + kx[i] * RE(dV, i) + ky[i] * RE(dV, N + i) + kz[i] * RE(dV, 2 * N + i);
            si = shear * ky[i] * IM(vx, i) + kx[i] * IM(dV, i) + ky[i] * IM(dV, N + i) + kz[i] * IM(dV, 2 * N + i);
        }
        double dr = -si, di = sr; /* divDeltaField = I * (...) */
        double k2 = kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i];
        double ik2 = k2 > 0.0 ? 1.0 / k2 : 1.0;
        const double *kk[3] = {kx, ky, kz};
        for (int c = 0; c < 3; c++) {
            /* += I k_c ik2 div */
            double a = kk[c][i] * ik2;
            RE(dV, c * N + i) += -a * di;
            IM(dV, c * N + i) += a * dr;
        }
    }
}

/* compute_dfield.cu:11-90 */
void orc_compute_dfield(orc_sim *s, int stage_step) {
    if (s->p.incompressible) {
        if (stage_step > 0) orc_complex2real(s);
        hyperbolic_terms(s);
        source_terms(s);
        parabolic_terms(s);
        grad_pseudo_pressure(s);
    } else if (s->p.heat_equation) {
        parabolic_terms(s);
    }
}

/* axpyComplex (cuda_kernels_generic.cu:137-144): Z = a X + b Y */
static void axpy(size_t n, const double *X, const double *Y, double *Z, double a, double b) {
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)n; li++) {
        size_t i = (size_t)li;
        double zr = a * X[2 * i] + b * Y[2 * i];
        double zi = a * X[2 * i + 1] + b * Y[2 * i + 1];
        Z[2 * i] = zr;
        Z[2 * i + 1] = zi;
    }
}

/* wavevector.cu:162-172 + ShearWavevector (cuda_kernels_generic.cu:362-384) */
void orc_shear_wavevector(orc_sim *s, double tremap) {
    const orc_params *p = &s->p;
    double kxmin = (2.0 * M_PI) / p->lx, ts = tremap * p->shear;
    int nzh = p->nz / 2 + 1;
    for (int i = 0; i < p->nx; i++)
        for (int j = 0; j < p->ny; j++)
            for (int k = 0; k < nzh; k++) {
                size_t idx = (size_t)k + (size_t)nzh * (j + (size_t)i * p->ny);
                double kx0 = kxmin * (fmod((double)i + ((double)p->nx / 2), (double)p->nx) - (double)p->nx / 2);
                s->kvec[idx] = kx0 + ts * s->kvec[s->ntc + idx];
            }
}

/* shearing_remapping.cu:15-24 */
void orc_shift_time(orc_sim *s) {
    const orc_params *p = &s->p;
    s->tremap = fmod(s->t + p->ly / (2.0 * p->shear * p->lx), p->ly / (p->shear * p->lx)) - p->ly / (2.0 * p->shear * p->lx);
}

/* shearing_remapping.cu:26-59, RemapComplexVec + MaskVector (cuda_kernels_generic.cu:386-423) */
void orc_remap_all_fields(orc_sim *s) {
    const orc_params *p = &s->p;
    int NX = p->nx, NY = p->ny, nzh = p->nz / 2 + 1;
    size_t N = s->ntc;
    double *re = zalloc(2 * N);
    for (int f = 0; f < s->NF; f++) {
        double *v = CF(s, f);
        memset(re, 0, sizeof(double) * 2 * N);
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NY; j++)
                for (int k = 0; k < nzh; k++) {
                    size_t idx = (size_t)k + (size_t)nzh * (j + (size_t)i * NY);
                    int nx = (int)fmod((double)i + (NX / 2), (double)NX) - NX / 2;
                    int ny = (int)fmod((double)j + (NY / 2), (double)NY) - NY / 2;
                    int nxt = nx + ny;
                    if ((nxt > -NX / 2) && (nxt < NX / 2)) {
                        if (nxt < 0) nxt += NX;
                        size_t d = (size_t)k + (size_t)nzh * j + (size_t)nzh * NY * nxt;
                        re[2 * d] = v[2 * idx];
                        re[2 * d + 1] = v[2 * idx + 1];
                    }
                }
        for (size_t i = 0; i < N; i++) {
            v[2 * i] = re[2 * i] * s->mask[i];
            v[2 * i + 1] = re[2 * i + 1] * s->mask[i];
        }
    }
    free(re);
}

static const double gammaRK[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double xiRK[2] = {-17.0 / 60.0, -5.0 / 12.0};

/* hydro_mhd_advance.cu:82-189 RungeKutta3 */
void orc_rk3(orc_sim *s) {
    const orc_params *p = &s->p;
    size_t n = s->ntc * (size_t)s->NF;
    double dt = s->dt;
    s->stage_step = 0;
    orc_compute_dfield(s, s->stage_step);
    axpy(n, s->fields, s->dfields, s->fields, 1.0, gammaRK[0] * dt);
    axpy(n, s->fields, s->dfields, s->scr, 1.0, xiRK[0] * dt);
    s->stage_step++;
    if (p->shearing) orc_shear_wavevector(s, s->tremap + gammaRK[0] * dt);
    orc_compute_dfield(s, s->stage_step);
    axpy(n, s->scr, s->dfields, s->fields, 1.0, gammaRK[1] * dt);
    axpy(n, s->fields, s->dfields, s->scr, 1.0, xiRK[1] * dt);
    s->stage_step++;
    if (p->shearing) orc_shear_wavevector(s, s->tremap + (gammaRK[0] + xiRK[0] + gammaRK[1]) * dt);
    orc_compute_dfield(s, s->stage_step);
    axpy(n, s->scr, s->dfields, s->fields, 1.0, gammaRK[2] * dt);
    s->t += s->dt;
    if (p->shearing) {
        s->tremap += s->dt;
        if (s->tremap > p->ly / (2.0 * p->shear * p->lx)) {
            orc_shift_time(s);
            orc_remap_all_fields(s);
        }
        orc_shear_wavevector(s, s->tremap);
    }
    s->stage_step++;
}

/* rkl.cu:466-585: stage count and coefficient recurrence */
int orc_rkl2_coeffs(double dt_current, double dt_sts, double *coef, int cap) {
    double scrh = dt_current / dt_sts;
    double s_str = 4.0 * (1.0 + 2.0 * scrh) / (1.0 + sqrt(9.0 + 16.0 * scrh));
    int s_RKL = 1 + (int)(s_str);
    double w1 = 4.0 / (s_RKL * s_RKL + s_RKL - 2.0);
    double mu_tilde_j = w1 / 3.0;
    double b_j, b_jm1, b_jm2, a_jm1;
    b_j = b_jm1 = b_jm2 = 1.0 / 3.0;
    a_jm1 = 1.0 - b_jm1;
    if (coef && cap >= 1) { coef[0] = 0; coef[1] = 0; coef[2] = mu_tilde_j; coef[3] = 0; }
    for (int stage = 2; stage <= s_RKL; stage++) {
        double mu_j = (2. * stage - 1.) / stage * b_j / b_jm1;
        mu_tilde_j = w1 * mu_j;
        double gamma_j = -a_jm1 * mu_tilde_j;
        double nu_j = -(stage - 1.) * b_j / (stage * b_jm2);
        if (coef && stage <= cap) {
            double *c = coef + 4 * (stage - 1);
            c[0] = mu_j; c[1] = nu_j; c[2] = mu_tilde_j; c[3] = gamma_j;
        }
        b_jm2 = b_jm1;
        b_jm1 = b_j;
        a_jm1 = 1.0 - b_jm1;
        b_j = 0.5 * (stage * stage + 3.0 * stage) / (stage * stage + 3.0 * stage + 2);
    }
    return s_RKL;
}

/* physics.cu:233-330 ParabolicTermsSTS into the compact dU arrays */
static void parabolic_terms_sts(orc_sim *s, double *dU) {
    const orc_params *p = &s->p;
    size_t N = s->ntc;
    if (p->incompressible)
        for (int c = 0; c < 3; c++)
            if (s->sts_pos[s->VEL + c] >= 0)
                nabla_scalar(s, CF(s, s->VEL + c), dU + 2 * N * s->sts_pos[s->VEL + c], p->nu, 0);
    if (p->mhd)
        for (int c = 0; c < 3; c++)
            if (s->sts_pos[s->MAG + c] >= 0)
                nabla_scalar(s, CF(s, s->MAG + c), dU + 2 * N * s->sts_pos[s->MAG + c], p->nu_m, 0);
    if (p->boussinesq || p->heat_equation) {
        if (s->sts_pos[s->TH] >= 0) {
            double *dTh = dU + 2 * N * s->sts_pos[s->TH];
            if (p->anisotropic_diffusion) {
                for (int c = 0; c < 3; c++)
                    if (s->sts_pos[s->MAG + c] >= 0)
                        complex2real_n(s, CF(s, s->MAG + c), RB(s, s->MAG + c), 1);
                anisotropic_heat_flux(s, dTh);
            } else {
                nabla_scalar(s, CF(s, s->TH), dTh, p->nu_th, 0);
            }
        }
    }
}

/* rkl.cu:437-591 compute_cycle_RKL2 */
void orc_rkl2_cycle(orc_sim *s) {
    size_t N = s->ntc;
    int ns = s->nsts;
    if (ns == 0) return;
    double dt_current = s->dt;
    /* the stage count follows dt / dt_sts without bound (rkl.cu:466-470) */
    int s_RKL = orc_rkl2_coeffs(dt_current, s->dt_sts, NULL, 0);
    double *coef = (double *)malloc(sizeof(double) * 4 * (size_t)(s_RKL + 1));
    orc_rkl2_coeffs(dt_current, s->dt_sts, coef, s_RKL);
    memset(s->dU0, 0, sizeof(double) * 2 * N * ns);
    memset(s->dU, 0, sizeof(double) * 2 * N * ns);
    parabolic_terms_sts(s, s->dU);
    double mu_tilde_1 = coef[2];
    for (int nv = 0; nv < ns; nv++) {
        int f = s->sts_index[nv];
        double *U = CF(s, f);
        double *Uc0 = s->Uc0 + 2 * N * nv, *Uc1 = s->Uc1 + 2 * N * nv;
        double *dU = s->dU + 2 * N * nv, *dU0 = s->dU0 + 2 * N * nv;
        memcpy(Uc0, U, sizeof(double) * 2 * N);
        memcpy(dU0, dU, sizeof(double) * 2 * N);
        memcpy(Uc1, U, sizeof(double) * 2 * N);
        axpy(N, Uc1, dU0, U, 1.0, mu_tilde_1 * dt_current);
    }
    for (int stage = 2; stage <= s_RKL; stage++) {
        const double *c = coef + 4 * (stage - 1);
        const double mu_j = c[0], nu_j = c[1], mu_tilde_j = c[2], gamma_j = c[3];
        memset(s->dU, 0, sizeof(double) * 2 * N * ns);
        parabolic_terms_sts(s, s->dU);
        for (int nv = 0; nv < ns; nv++) {
            int f = s->sts_index[nv];
            double *Uc = CF(s, f);
            double *Uc0 = s->Uc0 + 2 * N * nv, *Uc1 = s->Uc1 + 2 * N * nv;
            double *dU = s->dU + 2 * N * nv, *dU0 = s->dU0 + 2 * N * nv;
            double a = dt_current * mu_tilde_j, g = gamma_j * dt_current;
            /* rkl2_stage (cuda_kernels_generic.cu:160-172) */
#pragma omp parallel for schedule(static)
            for (long li = 0; li < (long)(2 * N); li++) {
                size_t i = (size_t)li;
                double Y = mu_j * Uc[i] + nu_j * Uc1[i];
                Uc1[i] = Uc[i];
                Uc[i] = Y + (1.0 - mu_j - nu_j) * Uc0[i] + a * dU[i] + g * dU0[i];
            }
        }
    }
    free(coef);
}

/* hydro_mhd_advance.cu:18-75 */
void orc_hydro_mhd_advance(orc_sim *s) {
    orc_compute_dt(s);
    if (s->p.supertimestepping && (s->step % 2) == 1) orc_rkl2_cycle(s);
    orc_rk3(s);
    if (s->p.supertimestepping && (s->step % 2) == 0) {
        orc_complex2real(s);
        orc_rkl2_cycle(s);
    }
    s->step += 1;
}

void orc_main_loop(orc_sim *s, int nsteps) {
    if (!s->loop_started) {
        /* supervisor.cu:171-175 */
        if (s->p.shearing) {
            orc_shift_time(s);
            orc_shear_wavevector(s, s->tremap);
        }
        s->loop_started = 1;
    }
    for (int n = 0; n < nsteps; n++) {
        orc_clean_divergence(s);
        orc_hydro_mhd_advance(s);
        if ((s->t - s->t_lastsnap + 1e-8) >= s->p.toutput_flow) s->t_lastsnap += s->p.toutput_flow;
    }
}

/* ------------------------------------------------------------------------- */
/* diagnostics                                                                */
/* ------------------------------------------------------------------------- */

static double energy_of(orc_sim *s, const double *v) {
    size_t N = s->ntc;
    int nzh = s->p.nz / 2 + 1;
    double tot = 0.0, sub = 0.0;
#pragma omp parallel for reduction(+ : tot) schedule(static)
    for (long li = 0; li < (long)N; li++) tot += v[2 * li] * v[2 * li] + v[2 * li + 1] * v[2 * li + 1];
    long rows = (long)s->p.nx * s->p.ny;
#pragma omp parallel for reduction(+ : sub) schedule(static)
    for (long r = 0; r < rows; r++) {
        size_t i = (size_t)r * nzh;
        sub += v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1];
    }
    double nt = (double)s->ntot;
    return tot / (nt * nt) - 0.5 * sub / (nt * nt);
}

/* spooky_output_functions.cu:23-55 computeEnergy */
double orc_energy(orc_sim *s, int field) { return energy_of(s, CF(s, field)); }

/* spooky_output_functions.cu:57-86 computeEnstrophy: Curl (cuda_kernels_generic.cu:275-286) */
double orc_enstrophy(orc_sim *s, int vf) {
    size_t N = s->ntc;
    const double *V = CF(s, vf);
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s);
    double *C = s->tmp;
#pragma omp parallel for schedule(static)
    for (long li = 0; li < (long)N; li++) {
        size_t i = (size_t)li;
        double ar = ky[i] * RE(V, 2 * N + i) - kz[i] * RE(V, N + i);
        double ai = ky[i] * IM(V, 2 * N + i) - kz[i] * IM(V, N + i);
        RE(C, i) = -ai; IM(C, i) = ar;
        ar = kz[i] * RE(V, i) - kx[i] * RE(V, 2 * N + i);
        ai = kz[i] * IM(V, i) - kx[i] * IM(V, 2 * N + i);
        RE(C, N + i) = -ai; IM(C, N + i) = ar;
        ar = kx[i] * RE(V, N + i) - ky[i] * RE(V, i);
        ai = kx[i] * IM(V, N + i) - ky[i] * IM(V, i);
        RE(C, 2 * N + i) = -ai; IM(C, 2 * N + i) = ar;
    }
    return energy_of(s, TMPC(s, 0)) + energy_of(s, TMPC(s, 1)) + energy_of(s, TMPC(s, 2));
}

/* spooky_output_functions.cu:188-235 twoFieldCorrelation on the real buffer */
double orc_correlation(orc_sim *s, int fa, int fb) {
    const double *a = RB(s, fa), *b = RB(s, fb);
    int nz = s->p.nz, nzp = 2 * (nz / 2 + 1);
    long rows = (long)s->p.nx * s->p.ny;
    double acc = 0.0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (long r = 0; r < rows; r++)
        for (int k = 0; k < nz; k++) acc += a[(size_t)r * nzp + k] * b[(size_t)r * nzp + k];
    return acc / (double)s->ntot;
}

/* Spectrum1d (cuda_kernels_generic.cu:441-471); bins at m >= nbins are dropped */
int orc_spectrum(orc_sim *s, int fa, int fb, double *out, int cap) {
    size_t N = s->ntc;
    int nzh = s->p.nz / 2 + 1;
    const double *v1 = CF(s, fa), *v2 = CF(s, fb);
    const double *kx = KX(s), *ky = KY(s), *kz = KZ(s);
    double ntotal = (double)s->p.nx * s->p.ny * s->p.nz;
    int nb = s->nbins;
    for (int m = 0; m < cap; m++) out[m] = 0.0;
    for (size_t i = 0; i < N; i++) {
        int kidx = (int)(i % nzh);
        double ar = v1[2 * i] / ntotal, ai = v1[2 * i + 1] / ntotal;
        double br = v2[2 * i] / ntotal, bi = v2[2 * i + 1] / ntotal;
        double pw = ar * br + ai * bi;
        if (kidx > 0) pw *= 2.0;
        double kabs = sqrt(kx[i] * kx[i] + ky[i] * ky[i] + kz[i] * kz[i]);
        int m = (int)(kabs / s->deltak + 0.5);
        if (m < nb && m < cap) out[m] += pw;
    }
    return nb;
}
