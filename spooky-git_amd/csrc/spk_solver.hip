// spk_solver.hip -- device state and time stepping of one SPOOKY run.
//
// One RHS evaluation (TimeStepping::compute_dfield, compute_dfield.cu:11-90):
//   k_inv_x      inverse x pass of the FP64 state (+ i k theta for conduction), scaled 1/N
//   pass y (+1)  inverse y pass, in place on W
//   k_zmid       z C2R of every input row -> real-space products -> z R2C (+ CFL maxima)
//   pass y (-1)  forward y pass, in place on P
//   pass x (-1)  forward x pass, in place on P
//   k_epilogue   i k derivatives, 2/3 mask, sources, diffusion, Leray projection and the
//                RK3 stage update, one pass over the spectral points
// W and P hold the bulk precision (FP64, or FP32 in FP32-bulk mode); the state,
// the RK3 scratch and every RKL2 array stay FP64.
#include "spk_internal.hpp"
#include "spk_kernels.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace spk {

// ------------------------------------------------------------------------
// compute_dt (compute_dt.cu:15-309): maxima reduction + CFL arithmetic on the device
// ------------------------------------------------------------------------
struct DtConst {
    int incompressible, heat_equation, mhd, boussinesq, aniso, rotating, shearing, supertimestepping;
    int safety_sts;
    int sts_v[3], sts_b[3], sts_th;
    double kxmax, kymax, kzmax;
    double nu, nu_m, nu_th, reynolds_ani, N2, OmegaT2, omega, shear;
    double cfl, cfl_par, cfl_sts, safety_source, t_final, toutput_flow;
    int nblocks;
};

__global__ __launch_bounds__(256) void k_dt(const double *__restrict__ pmax, DtConst c, TimeState *ts, Batch bt) {
    __shared__ double sm[6][256];
    pmax = rebase(bt, 0, pmax);
    ts = rebase(bt, 1, ts);
    if (const BoxRow *br = batch_row(bt)) {
        c.nu = br->nu;
        c.nu_m = br->nu_m;
        c.nu_th = br->nu_th;
        c.reynolds_ani = br->reynolds_ani;
    }
    const int tid = threadIdx.x;
    double mx[6] = {0, 0, 0, 0, 0, 0};
    if (c.incompressible) {
        for (int b = tid; b < c.nblocks; b += 256)
#pragma unroll
            for (int q = 0; q < 6; q++) mx[q] = fmax(mx[q], pmax[(long long)b * 6 + q]);
    }
#pragma unroll
    for (int q = 0; q < 6; q++) sm[q][tid] = mx[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
#pragma unroll
            for (int q = 0; q < 6; q++) sm[q][tid] = fmax(sm[q][tid], sm[q][tid + s]);
        __syncthreads();
    }
    if (tid != 0) return;
    double dt_hyp = 0.0, dt_par = 0.0, dt_sts = 0.0, dt = ts->dt;
    double gamma_v = 0.0, gamma_th = 0.0, gamma_par = 0.0, gamma_b = 0.0, gamma_tmp = 0.0, gamma_sts = 0.0;
    const double kxmax = c.kxmax, kymax = c.kymax, kzmax = c.kzmax, tremap = ts->tremap;
    if (c.heat_equation) {
        gamma_tmp = ((kxmax) * (kxmax) + kymax * kymax + kzmax * kzmax) * c.nu_th;
        if (!c.sts_th) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        dt_par = c.cfl * c.cfl_par / gamma_par;
        dt_sts = c.cfl_par / gamma_sts;
        dt = dt_par;
    }
    if (c.incompressible) {
        const double maxfx = sm[0][0], maxfy = sm[1][0], maxfz = sm[2][0];
        const double kxe = kxmax + fabs(tremap) * kymax;
        gamma_v = (kxe * maxfx + kymax * maxfy + kzmax * maxfz);
        if (c.rotating) gamma_v += fabs(c.omega) / c.safety_source;
        if (c.shearing) gamma_v += fabs(c.shear) / c.safety_source;
        gamma_tmp = kxe * kxe * c.nu;
        if (!c.sts_v[0]) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        gamma_tmp = kymax * kymax * c.nu;
        if (!c.sts_v[1]) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        gamma_tmp = kzmax * kzmax * c.nu;
        if (!c.sts_v[2]) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        if (c.boussinesq) {
            gamma_th += pow(fabs(c.N2), 0.5) / c.safety_source;
            if (c.aniso) {
                gamma_th += pow(fabs(c.OmegaT2), 0.5) / c.safety_source;
                gamma_tmp = (kxe * kxe + kymax * kymax + kzmax * kzmax) * (1. / c.reynolds_ani);
            } else {
                gamma_tmp = (kxe * kxe + kymax * kymax + kzmax * kzmax) * c.nu_th;
            }
            if (!c.sts_th) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        }
        if (c.mhd) {
            const double maxbx = sm[3][0], maxby = sm[4][0], maxbz = sm[5][0];
            gamma_b = kxe * maxbx + kymax * maxby + kzmax * maxbz;
            gamma_tmp = kxe * kxe * c.nu_m;
            if (!c.sts_b[0]) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
            gamma_tmp = kymax * kymax * c.nu_m;
            if (!c.sts_b[1]) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
            gamma_tmp = kzmax * kzmax * c.nu_m;
            if (!c.sts_b[2]) gamma_par += gamma_tmp; else gamma_sts += gamma_tmp;
        }
        dt_hyp = c.cfl / (gamma_v + gamma_th + gamma_b);
        dt_par = c.cfl_par / gamma_par;
        dt_sts = c.cfl_sts / gamma_sts;
        dt = (dt_hyp < dt_par) ? dt_hyp : dt_par;
        if (c.supertimestepping) {
            if (dt > dt_sts * c.safety_sts) dt = dt_sts * c.safety_sts;
        }
    }
    const double t = ts->t, tls = ts->t_lastsnap;
    if (t + dt > c.t_final) {
        dt = c.t_final - t;
    } else if (t + dt - tls > c.toutput_flow) {
        dt = c.toutput_flow - t + tls;
    }
    if (c.supertimestepping) {
        if (dt_sts > dt) dt_sts = dt;
    }
    ts->dt = dt;
    ts->dt_hyp = dt_hyp;
    ts->dt_par = dt_par;
    ts->dt_sts = dt_sts;
}

// clock updates: t += dt (RungeKutta3), step++ (HydroMHDAdvance), snapshot
// clock of CheckSnapshot (check_output.cu:67-83), tremap from the host
__global__ void k_clock(TimeState *ts, int add_dt, int add_step, int snap, double toutput_flow,
                        int set_tremap, double tremap, Batch bt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ts = rebase(bt, 0, ts);
    if (add_dt) ts->t += ts->dt;
    if (add_step) ts->step += 1.0;
    if (set_tremap) ts->tremap = tremap;
    if (snap && (ts->t - ts->t_lastsnap + 1e-8) >= toutput_flow) ts->t_lastsnap += toutput_flow;
}

__global__ void k_set_time(TimeState *ts, double t, double step, double tls, double tremap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ts->t = t;
    ts->step = step;
    ts->t_lastsnap = tls;
    ts->tremap = tremap;
}

// ------------------------------------------------------------------------
// CleanDivergence (cuda_kernels_generic.cu:288-313), in place on u and B
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_clean_div(cpx<double> *__restrict__ U, long long ntc, KGeom kg,
                                                   int f0, int f1, Batch bt) {
    U = rebase(bt, 0, U);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntc) return;
    const long long rowi = idx / kg.nzh;
    const int k = (int)(idx - rowi * kg.nzh);
    const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
    double kx, ky, kz, m;
    kvec_at(kg, i, j, k, kx, ky, kz, m);
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double ik2 = k2 > 0.0 ? 1.0 / k2 : 1.0;
    const double kk[3] = {kx, ky, kz};
    for (int g = 0; g < 2; g++) {
        const int f = g == 0 ? f0 : f1;
        if (f < 0) continue;
        cpx<double> X[3];
#pragma unroll
        for (int q = 0; q < 3; q++) X[q] = U[(f + q) * ntc + idx];
        const cpx<double> s = kdot3(kx, X[0], ky, X[1], kz, X[2]);
        const cpx<double> q0 = {-s.y, s.x};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const cpx<double> ikq = {-kk[q] * q0.y, kk[q] * q0.x};  // I k q0
            U[(f + q) * ntc + idx] = cpx<double>{X[q].x + ikq.x * ik2, X[q].y + ikq.y * ik2};
        }
    }
}

// The recurrence of rkl2_coeffs (host), operation for operation and without
// contraction into FMAs, so that both give the same doubles.
__global__ void k_rkl2_plan(const TimeState *__restrict__ ts, int smax, RklPlan *__restrict__ plan, Batch bt) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    ts = rebase(bt, 0, ts);
    plan = rebase(bt, 1, plan);
    const double scrh = ts->dt / ts->dt_sts;
    const double s_str = 4.0 * (1.0 + 2.0 * scrh) / (1.0 + sqrt(9.0 + 16.0 * scrh));
    // 1 + (int)s_str as on the host; past smax (or NaN) the plan is cut at smax
    int s_RKL = s_str < (double)smax ? 1 + (int)(s_str) : smax + 1;
    plan->overflow = s_RKL > smax;
    if (s_RKL > smax) s_RKL = smax;
    plan->s = s_RKL;
    const double w1 = 4.0 / (s_RKL * s_RKL + s_RKL - 2.0);
    double mu_tilde_j = w1 / 3.0;
    double b_j, b_jm1, b_jm2, a_jm1;
    b_j = b_jm1 = b_jm2 = 1.0 / 3.0;
    a_jm1 = 1.0 - b_jm1;
    for (int j = 0; j < 4 * smax; j++) plan->coef[j] = 0.0;
    plan->coef[2] = mu_tilde_j;
    for (int stage = 2; stage <= s_RKL; stage++) {
        const double mu_j = (2. * stage - 1.) / stage * b_j / b_jm1;
        mu_tilde_j = w1 * mu_j;
        const double gamma_j = -a_jm1 * mu_tilde_j;
        const double nu_j = -(stage - 1.) * b_j / (stage * b_jm2);
        double *c = plan->coef + 4 * (stage - 1);
        c[0] = mu_j;
        c[1] = nu_j;
        c[2] = mu_tilde_j;
        c[3] = gamma_j;
        b_jm2 = b_jm1;
        b_jm1 = b_j;
        a_jm1 = 1.0 - b_jm1;
        b_j = 0.5 * (stage * stage + 3.0 * stage) / (stage * stage + 3.0 * stage + 2);
    }
    plan->active[0] = 0;
    for (int j = 1; j <= smax; j++) plan->active[j] = j <= s_RKL;
}

// plan == nullptr: the host passes the stage's coefficients; otherwise they come
// from the device plan and the stage returns at once beyond its stage count
template <class T>
__global__ __launch_bounds__(256) void k_rkl2(const cpx<T> *__restrict__ Q, cpx<double> *__restrict__ U,
                                              cpx<double> *__restrict__ Uc0, cpx<double> *__restrict__ Uc1,
                                              cpx<double> *__restrict__ dU0, RklConst c, int stage, double mu,
                                              double nu, double dtmut, double gdt, const RklPlan *__restrict__ plan,
                                              const TimeState *__restrict__ ts, Batch bt) {
    Q = rebase(bt, 0, Q);
    U = rebase(bt, 1, U);
    Uc0 = rebase(bt, 2, Uc0);
    Uc1 = rebase(bt, 3, Uc1);
    dU0 = rebase(bt, 4, dU0);
    plan = rebase(bt, 5, plan);
    ts = rebase(bt, 6, ts);
    if (const BoxRow *br = batch_row(bt)) {
        c.nu = br->nu;
        c.nu_m = br->nu_m;
        c.nu_th = br->nu_th;
    }
    if (plan != nullptr) {
        if (!plan->active[stage]) return;
        const double *cf = plan->coef + 4 * (stage - 1);
        const double dt = ts->dt;
        mu = cf[0];
        nu = cf[1];
        dtmut = cf[2] * dt;
        gdt = cf[3] * dt;
    }
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= c.ntc) return;
    const long long N = c.ntc;
    const long long rowi = idx / c.kg.nzh;
    const int k = (int)(idx - rowi * c.kg.nzh);
    const int j = (int)(rowi % c.kg.ny), i = (int)(rowi / c.kg.ny);
    double kx, ky, kz, m;
    kvec_at(c.kg, i, j, k, kx, ky, kz, m);
    const double k2 = kx * kx + ky * ky + kz * kz;
    for (int n = 0; n < c.nsts; n++) {
        const long long ui = c.field[n] * N + idx, si = n * N + idx;
        const cpx<double> uc = U[ui];
        cpx<double> du;
        if (c.kind[n] == 3) {
            const cpx<double> q0 = cconv<double>(Q[idx]), q1 = cconv<double>(Q[N + idx]),
                              q2 = cconv<double>(Q[2 * N + idx]);
            du = mul_plus_im(m, kdot3(kx, q0, ky, q1, kz, q2));
        } else {
            const double a = c.kind[n] == 0 ? c.nu : (c.kind[n] == 1 ? c.nu_m : c.nu_th);
            du = cscale(uc, -a * k2);
        }
        if (stage == 1) {
            Uc0[si] = uc;
            dU0[si] = du;
            Uc1[si] = uc;
            U[ui] = cpx<double>{uc.x + dtmut * du.x, uc.y + dtmut * du.y};
        } else {
            const cpx<double> u1 = Uc1[si], u0 = Uc0[si], d0 = dU0[si];
            const double w = 1.0 - mu - nu;
            const cpx<double> Y = {mu * uc.x + nu * u1.x, mu * uc.y + nu * u1.y};
            Uc1[si] = uc;
            U[ui] = cpx<double>{Y.x + w * u0.x + dtmut * du.x + gdt * d0.x, Y.y + w * u0.y + dtmut * du.y + gdt * d0.y};
        }
    }
}

// ------------------------------------------------------------------------
// shearing-box remap (RemapComplexVec + MaskVector, cuda_kernels_generic.cu:386-423)
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_remap(const cpx<double> *__restrict__ v, cpx<double> *__restrict__ out,
                                               long long ntc, int nx, int ny, int nzh) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntc) return;
    const long long rowi = idx / nzh;
    const int k = (int)(idx - rowi * nzh);
    const int j = (int)(rowi % ny), i = (int)(rowi / ny);
    const int ix = i < nx / 2 ? i : i - nx;
    const int iy = j < ny / 2 ? j : j - ny;
    int nxt = ix + iy;
    if (nxt > -nx / 2 && nxt < nx / 2) {
        if (nxt < 0) nxt += nx;
        out[(long long)k + (long long)nzh * j + (long long)nzh * ny * nxt] = v[idx];
    }
}

__global__ __launch_bounds__(256) void k_mask_copy(const cpx<double> *__restrict__ in, cpx<double> *__restrict__ v,
                                                   long long ntc, KGeom kg) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntc) return;
    const long long rowi = idx / kg.nzh;
    const int k = (int)(idx - rowi * kg.nzh);
    const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
    double kx, ky, kz, m;
    kvec_at(kg, i, j, k, kx, ky, kz, m);
    v[idx] = cscale(in[idx], m);
}

__global__ __launch_bounds__(256) void k_zero(cpx<double> *__restrict__ v, long long n) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) v[idx] = cpx<double>{0.0, 0.0};
}

// ------------------------------------------------------------------------
// measurement reductions (spooky_output_functions.cu): fixed-order two-level sums
// ------------------------------------------------------------------------
constexpr int RED_BLOCKS = 1024;

__device__ __forceinline__ double block_sum(double v, double *sm) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sm[w];
    return s;
}

// kind 0: energy weight sum |v|^2 (computeEnergy); kind 1: enstrophy (Curl +
// computeEnergy); kind 2: energy of the gradient (computeDissipation)
__global__ __launch_bounds__(256) void k_spec_sum(const cpx<double> *__restrict__ U, long long ntc, KGeom kg,
                                                  int kind, double *__restrict__ part) {
    __shared__ double sm[4];
    double acc = 0.0;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < ntc;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long rowi = idx / kg.nzh;
        const int k = (int)(idx - rowi * kg.nzh);
        const double w = k == 0 ? 0.5 : 1.0;
        if (kind == 0) {
            const cpx<double> a = U[idx];
            acc += w * (a.x * a.x + a.y * a.y);
        } else if (kind == 2) {
            // sum of the energies of i k_n a (Gradient + computeEnergy, computeDissipation)
            const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
            double kx, ky, kz, m;
            kvec_at(kg, i, j, k, kx, ky, kz, m);
            const cpx<double> a = U[idx];
            acc += w * (kx * kx + ky * ky + kz * kz) * (a.x * a.x + a.y * a.y);
        } else {
            const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
            double kx, ky, kz, m;
            kvec_at(kg, i, j, k, kx, ky, kz, m);
            const cpx<double> V0 = U[idx], V1 = U[ntc + idx], V2 = U[2 * ntc + idx];
            const cpx<double> c0 = {-(ky * V2.y - kz * V1.y), ky * V2.x - kz * V1.x};
            const cpx<double> c1 = {-(kz * V0.y - kx * V2.y), kz * V0.x - kx * V2.x};
            const cpx<double> c2 = {-(kx * V1.y - ky * V0.y), kx * V1.x - ky * V0.x};
            acc += w * (c0.x * c0.x + c0.y * c0.y + c1.x * c1.x + c1.y * c1.y + c2.x * c2.x + c2.y * c2.y);
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// twoFieldCorrelation on physical points of padded real rows
__global__ __launch_bounds__(256) void k_real_dot(const double *__restrict__ a, const double *__restrict__ b,
                                                  long long nrows, int nz, int nzp, double *__restrict__ part) {
    __shared__ double sm[4];
    double acc = 0.0;
    const long long n = nrows * nz;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long r = p / nz;
        const long long o = r * nzp + (p - r * nz);
        acc += a[o] * b[o];
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_final_sum(const double *__restrict__ part, int n, double scale,
                                                    double *__restrict__ out) {
    __shared__ double sm[1024];
    const int tid = threadIdx.x;
    sm[tid] = tid < n ? part[tid] : 0.0;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (tid < s) sm[tid] += sm[tid + s];
        __syncthreads();
    }
    if (tid == 0) *out = sm[0] * scale;
}

// Spectrum1d (cuda_kernels_generic.cu:441-471); LDS histogram, then global atomics
__global__ __launch_bounds__(256) void k_spectrum(const cpx<double> *__restrict__ v1, const cpx<double> *__restrict__ v2,
                                                  long long ntc, KGeom kg, double ntotal, double deltak, int nbins,
                                                  double *__restrict__ out) {
    extern __shared__ double hist[];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0.0;
    __syncthreads();
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < ntc;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long rowi = idx / kg.nzh;
        const int k = (int)(idx - rowi * kg.nzh);
        const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
        double kx, ky, kz, m;
        kvec_at(kg, i, j, k, kx, ky, kz, m);
        const cpx<double> a = cscale(v1[idx], 1.0 / ntotal), b = cscale(v2[idx], 1.0 / ntotal);
        double pw = a.x * b.x + a.y * b.y;
        if (k > 0) pw *= 2.0;
        const double kabs = sqrt(kx * kx + ky * ky + kz * kz);
        const int bin = (int)(kabs / deltak + 0.5);
        if (bin < nbins) atomicAdd(&hist[bin], pw);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (hist[b] != 0.0) atomicAdd(&out[b], hist[b]);
}

__global__ void k_zero_d(double *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// ---- timevar diagnostics (spooky_output_functions.cu) ----

// spectral 3-vector from the state: kind 0 Helicity (i k x V / k^2, k^2 = 0 -> 1,
// cuda_kernels.cu Helicity), 1 Curl (i k x V), 2 Gradient of the scalar V[0] (i k V)
__global__ __launch_bounds__(256) void k_diag_vec(const cpx<double> *__restrict__ V, cpx<double> *__restrict__ Z,
                                                  long long ntc, KGeom kg, int kind) {
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < ntc;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long rowi = idx / kg.nzh;
        const int k = (int)(idx - rowi * kg.nzh);
        const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
        double kx, ky, kz, m;
        kvec_at(kg, i, j, k, kx, ky, kz, m);
        cpx<double> z0, z1, z2;
        if (kind == 2) {
            const cpx<double> a = V[idx];
            z0 = {-kx * a.y, kx * a.x};
            z1 = {-ky * a.y, ky * a.x};
            z2 = {-kz * a.y, kz * a.x};
        } else {
            const cpx<double> V0 = V[idx], V1 = V[ntc + idx], V2 = V[2 * ntc + idx];
            // i (k x V)
            const cpx<double> c0 = {ky * V2.x - kz * V1.x, ky * V2.y - kz * V1.y};
            const cpx<double> c1 = {kz * V0.x - kx * V2.x, kz * V0.y - kx * V2.y};
            const cpx<double> c2 = {kx * V1.x - ky * V0.x, kx * V1.y - ky * V0.y};
            double f = 1.0;
            if (kind == 0) {
                const double k2 = kx * kx + ky * ky + kz * kz;
                f = k2 > 0.0 ? 1.0 / k2 : 1.0;
            }
            z0 = {-f * c0.y, f * c0.x};
            z1 = {-f * c1.y, f * c1.x};
            z2 = {-f * c2.y, f * c2.x};
        }
        Z[idx] = z0;
        Z[ntc + idx] = z1;
        Z[2 * ntc + idx] = z2;
    }
}

// real-space 3-vector from B (padded rows, every value including padding):
// kind 0 Computebbstrat (B_i B_s / B^2, cuda_kernels.cu:385-398); kind 1
// ComputeBGradTheta + ComputeAnisotropicHeatFlux with chi = 1, OmegaT2 = 0
// (B_i (B . G) / B^2, cuda_kernels.cu:360-383); G: real grad theta
__global__ __launch_bounds__(256) void k_diag_bvec(const double *__restrict__ B, const double *__restrict__ G,
                                                   double *__restrict__ Z, long long n, int sd, int kind) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double b0 = B[i], b1 = B[n + i], b2 = B[2 * n + i];
        const double B2 = b0 * b0 + b1 * b1 + b2 * b2;
        double c;
        if (kind == 0) c = (sd == 0 ? b0 : (sd == 1 ? b1 : b2)) / B2;
        else c = (b0 * G[i] + b1 * G[n + i] + b2 * G[2 * n + i]) / B2;
        Z[i] = b0 * c;
        Z[n + i] = b1 * c;
        Z[2 * n + i] = b2 * c;
    }
}

// DivergenceMask (cuda_kernels_generic.cu:315-334): i mask (k . X), written
// (acc 0, flag 0 there) or added to Z (acc 1, flag 1 there)
template <class T>
__global__ __launch_bounds__(256) void k_diag_div(const cpx<T> *__restrict__ X, cpx<double> *__restrict__ Z,
                                                  long long ntc, KGeom kg, int acc) {
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < ntc;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long rowi = idx / kg.nzh;
        const int k = (int)(idx - rowi * kg.nzh);
        const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
        double kx, ky, kz, m;
        kvec_at(kg, i, j, k, kx, ky, kz, m);
        const cpx<double> d = kdot3(kx, cconv<double>(X[idx]), ky, cconv<double>(X[ntc + idx]), kz,
                                    cconv<double>(X[2 * ntc + idx]));
        const cpx<double> z0 = acc ? Z[idx] : cpx<double>{0.0, 0.0};
        Z[idx] = cpx<double>{z0.x - m * d.y, z0.y + m * d.x};
    }
}

// physical-point sums of a function of real B: kind 0 |b_z| (averagebz), 1 b_z^2
// (averagebz2) with b = B / (|B| + 1e-16) (bUnitvector); 2 acosf(|B_x| / (|B_xy| +
// 1e-16)) (averagephiB, AngleHorizPlane: single-precision acos as there)
__global__ __launch_bounds__(256) void k_diag_bsum(const double *__restrict__ B, long long fs, long long nrows, int nz,
                                                   int nzp, int kind, double *__restrict__ part) {
    __shared__ double sm[4];
    double acc = 0.0;
    const long long n = nrows * nz;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long r = p / nz;
        const long long o = r * nzp + (p - r * nz);
        const double bx = B[o], by = B[fs + o], bz = B[2 * fs + o];
        if (kind == 2) {
            const double nrm = 1e-16 + sqrt(bx * bx + by * by);
            acc += (double)acosf((float)(fabs(bx) / nrm));
        } else {
            const double nrm = 1e-16 + sqrt(bx * bx + by * by + bz * bz);
            const double u = bz / nrm;
            acc += kind == 0 ? fabs(u) : u * u;
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---- spectral transfer terms of the spectrum output (spooky_output_spectrum.cu:22-400) ----

// real-space products over padded rows (padding included, as there); inputs and
// outputs at a stride of n doubles per component:
// kind 0 ShearMatrix (cuda_kernels.cu:59-77): A_i A_j for ij = xx xy xz yy yz zz
// kind 1 MagneticEmf (cuda_kernels.cu:240-252): A x B with A = u, B = B
// kind 2 EnergyFluxVector (cuda_kernels.cu:277-290): A_i B with B a scalar
// kind 3 A_1 A_2, one output (B_y B_z)
// kind 4 ComputebGradTheta (user_defined_cuda_kernels.cu:50-60): B = real grad
//   theta, A = real B: (A . B) / (1e-16 + |A|), one output
__global__ __launch_bounds__(256) void k_diag_prod(const double *A, const double *B, double *Z, long long n, int kind) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double a0 = A[i], a1 = A[n + i], a2 = kind == 2 || kind == 3 ? 0.0 : A[2 * n + i];
        if (kind == 0) {
            Z[i] = a0 * a0;
            Z[n + i] = a0 * a1;
            Z[2 * n + i] = a0 * a2;
            Z[3 * n + i] = a1 * a1;
            Z[4 * n + i] = a1 * a2;
            Z[5 * n + i] = a2 * a2;
        } else if (kind == 1) {
            const double b0 = B[i], b1 = B[n + i], b2 = B[2 * n + i];
            Z[i] = -b1 * a2 + b2 * a1;
            Z[n + i] = b0 * a2 - b2 * a0;
            Z[2 * n + i] = -b0 * a1 + b1 * a0;
        } else if (kind == 2) {
            const double th = B[i];
            Z[i] = a0 * th;
            Z[n + i] = a1 * th;
            Z[2 * n + i] = A[2 * n + i] * th;
        } else if (kind == 3) {
            Z[i] = a1 * A[2 * n + i];
        } else {
            const double nrm = 1e-16 + sqrt(a0 * a0 + a1 * a1 + a2 * a2);
            Z[i] = (a0 * B[i] + a1 * B[n + i] + a2 * B[2 * n + i]) / nrm;
        }
    }
}

// spectral operators on transformed products, in place (Z may equal X: every
// element reads its inputs before it writes):
// kind 0 NonLinAdvection (cuda_kernels.cu:101-123): Z_i = -i mask k_j S_ij from
//   the 6 components of kind 0 above
// kind 1 MagneticShear (cuda_kernels.cu:257-272): Z = i mask k x emf
// kind 2 NonLinBoussinesqAdv (cuda_kernels.cu:293-303): Z = -i mask k . E
__global__ __launch_bounds__(256) void k_diag_sop(const cpx<double> *X, cpx<double> *Z, long long ntc, KGeom kg,
                                                  int kind) {
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < ntc;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long rowi = idx / kg.nzh;
        const int k = (int)(idx - rowi * kg.nzh);
        const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
        double kx, ky, kz, m;
        kvec_at(kg, i, j, k, kx, ky, kz, m);
        const cpx<double> x0 = X[idx], x1 = X[ntc + idx], x2 = X[2 * ntc + idx];
        if (kind == 0) {
            const cpx<double> x3 = X[3 * ntc + idx], x4 = X[4 * ntc + idx], x5 = X[5 * ntc + idx];
            const cpx<double> d0 = kdot3(kx, x0, ky, x1, kz, x2);
            const cpx<double> d1 = kdot3(kx, x1, ky, x3, kz, x4);
            const cpx<double> d2 = kdot3(kx, x2, ky, x4, kz, x5);
            Z[idx] = cpx<double>{m * d0.y, -m * d0.x};
            Z[ntc + idx] = cpx<double>{m * d1.y, -m * d1.x};
            Z[2 * ntc + idx] = cpx<double>{m * d2.y, -m * d2.x};
        } else if (kind == 1) {
            const cpx<double> c0 = {ky * x2.x - kz * x1.x, ky * x2.y - kz * x1.y};
            const cpx<double> c1 = {kz * x0.x - kx * x2.x, kz * x0.y - kx * x2.y};
            const cpx<double> c2 = {kx * x1.x - ky * x0.x, kx * x1.y - ky * x0.y};
            Z[idx] = cpx<double>{-m * c0.y, m * c0.x};
            Z[ntc + idx] = cpx<double>{-m * c1.y, m * c1.x};
            Z[2 * ntc + idx] = cpx<double>{-m * c2.y, m * c2.x};
        } else {
            const cpx<double> d = kdot3(kx, x0, ky, x1, kz, x2);
            Z[idx] = cpx<double>{m * d.y, -m * d.x};
        }
    }
}

// Large-scale forcing: U[v0 + f][idx[m]] += dt F[m][f] for the listed modes
__global__ void k_force(cpx<double> *__restrict__ U, long long ntc, int v0, const long long *__restrict__ idx,
                        const cpx<double> *__restrict__ F, int nmodes, const TimeState *__restrict__ ts) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmodes) return;
    const double dt = ts->dt;
    for (int f = 0; f < 3; f++) {
        cpx<double> &u = U[(long long)(v0 + f) * ntc + idx[m]];
        const cpx<double> a = F[3 * m + f];
        u = cpx<double>{u.x + dt * a.x, u.y + dt * a.y};
    }
}

// UnshearComplexVec (cuda_kernels_generic.cu:425-439) on a copy of nf fields:
// Z = U exp(-i ky c / 2), ky unsheared
__global__ __launch_bounds__(256) void k_unshear(const cpx<double> *__restrict__ U, cpx<double> *__restrict__ Z,
                                                 long long ntc, long long n, KGeom kg, double c) {
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long e = idx % ntc;
        const int j = (int)((e / kg.nzh) % kg.ny);
        const int jj = j < kg.ny / 2 ? j : j - kg.ny;
        const double ph = -0.5 * (kg.kfy * (double)jj) * c;
        const double cs = cos(ph), sn = sin(ph);
        const cpx<double> u = U[idx];
        Z[idx] = cpx<double>{u.x * cs - u.y * sn, u.x * sn + u.y * cs};
    }
}

}  // namespace spk

namespace spk {
bool fused_supported_nx(int nx);
bool rkl_fused_supported_nx(int nx);
int launch_pass_y_rkl(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int dir, void *buf, const KGeom &kg);
int launch_fwdx_rkl2(spk_ctx ctx, int bulk, int nx, const void *Q, void *U, void *Uc0, void *Uc1, void *dU0, void *W,
                     const RklConst &c, int stage, double mu, double nu, double dtmut, double gdt, int inv,
                     double scale, const void *tw, const RklPlan *plan, const TimeState *ts);
int launch_fwdx_epilogue(spk_ctx ctx, int bulk, int nx, int ly, int mode, int ef, const void *P, void *U, void *S,
                         void *W, const FusedConst &fc, const void *tw, const TimeState *ts, int ngroups);
bool plane_supported(const Grid3 &g, int prog);
bool plane_rkl_supported(const Grid3 &g);
int launch_plane(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int prog, const void *W, void *P,
                 const ZConst &zc, double *pmax);
int launch_plane_rkl(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, const void *G, const void *Breal, void *Q,
                     const KGeom &kg, const ZConst &zc);
}  // namespace spk

// ==========================================================================
// host side
// ==========================================================================
using namespace spk;

struct spk_solver_s {
    spk_ctx ctx;
    spk_params p;
    Grid3 g;
    int bulk;
    int ly;
    int nf, VX, VY, VZ, BX, BY, BZ, TH, VEL, MAG;
    int sts_pos[8], sts_index[8], nsts;
    int aniso_rhs, diff_mask;
    double kxmax, kymax, kzmax, kmax, deltak;
    int nbins;
    KGeom kg0;
    // RHS program
    int zprog, nin, nout, oT, oE, oQ, oA;
    InvProg inv;
    // RKL2 program (anisotropic conduction of theta)
    int rkl_aniso;
    InvProg inv_rkl;
    RklConst rkl;
    // device buffers
    cpx<double> *U, *S, *dU;
    double *realbuf;
    void *W, *P, *W2, *P2;
    void *Breal;        // real-space B rows (padded, bulk precision), made once per RKL2 cycle
    cpx<double> *diag;  // 6 FP64 fields of scratch for the timevar diagnostics, allocated at first use
    int rkl_cache_b;    // B is not an sts variable: its real-space rows stay valid for a whole cycle
    int rkl_device;     // RKL2 cycles planned on the device (no host synchronization per cycle)
    int rkl_smax;       // stage count at the sts cap: the stages a device-planned cycle launches
    RklPlan *rkl_plan;  // device
    int force_n;               // large-scale forcing: modes, their state indices and values (device)
    long long *force_idx;
    cpx<double> *force_val;
    InvProg inv_rkl_b, inv_rkl_g, inv_th;
    InvProg inv_rkl_t;  // fused RKL2 stage 1: i k_x theta, theta
    cpx<double> *rUc0, *rUc1, *rdU0;
    TimeState *ts;
    double *pmax, *part, *res, *spec;
    int zmid_blocks;
    Twiddles twb, tw64;
    // host mirror
    int step;
    double tremap;
    bool loop_started, products_valid;
    bool fused_w_ok;  // W holds the inverse x pass of the state after the last fused step
    const void *products_src;  // the state whose RHS products P holds when products_valid
    int fused_mode;      // caller's switch for k_fwdx_epilogue: 0 never, 1 autotuned, 2 always
    bool fused_faster;   // tile autotune: the fused back end beat the unfused one on this grid
    int fused_ef;        // points per thread of its x line
    int fused_ef_auto;   // the tile autotune's choice of fused_ef
    bool plane_on;       // caller's switch for the quasi-2D plane kernels (k_plane, k_plane_rkl)
    int wslot_b;         // W slot of B_x in the RHS program (-1: none)
    // where the next RKL2 cycle takes B's real-space rows from: 0 the state (inverse
    // x, y and z passes); 1 W's B slots, which hold B's inverse x and y passes;
    // 2 W's B slots, which hold its inverse x pass (set by the fused step around
    // its cycles, where W holds the current state)
    int breal_from_w;
};

namespace {

size_t esize(int prec) { return prec == SPK_F64 ? 16 : 8; }

template <class T>
const cpx<T> *twp(const Twiddles &tw, long long off) { return reinterpret_cast<const cpx<T> *>(tw.d) + off; }

KGeom kgeom(const spk_solver_s *s, double kxshift) {
    KGeom k = s->kg0;
    k.kxshift = kxshift;
    return k;
}

unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

// ---- dispatchers over the grid size ----

template <class T>
int launch_inv_x(spk_solver s, const InvProg &prog, int nslots, void *W, double kxshift) {
    const Grid3 &g = s->g;
    const double scale = 1.0 / ((double)g.nx * g.ny * g.nz);
    const cpx<T> *tw = twp<T>(s->twb, s->twb.off_x);
    KGeom kg = kgeom(s, kxshift);
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {s->U, W, s->ctx->gate}, &bt));
#define SPK_INV_CASE(NN)                                                                          \
    case NN: {                                                                                    \
        constexpr int TC = strided_tcols(NN);                                                     \
        dim3 grid((unsigned)(((long long)g.ny * g.nzh + TC - 1) / TC), nslots, nboxes(s->ctx));  \
        k_inv_x<T, NN><<<grid, TC * strided_tpl(NN), 0, s->ctx->stream>>>(s->U, (cpx<T> *)W,     \
                                                                           g.ntc, prog, kg, scale, tw, \
                                                                           s->ctx->gate, bt);          \
        break;                                                                                    \
    }
    switch (g.nx) {
        SPK_INV_CASE(2) SPK_INV_CASE(4) SPK_INV_CASE(8) SPK_INV_CASE(16) SPK_INV_CASE(32) SPK_INV_CASE(64)
        SPK_INV_CASE(128) SPK_INV_CASE(256) SPK_INV_CASE(512) SPK_INV_CASE(1024) SPK_INV_CASE(2048)
        default: set_error("NX=%d unsupported", g.nx); return SPK_ERR_ARG;
    }
#undef SPK_INV_CASE
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

template <class T, int NZ, bool BAT>
int launch_zmid_prog(int prog, dim3 grid, int nt, hipStream_t st, const cpx<T> *w, const cpx<T> *wr, cpx<T> *pp,
                     long long nrows, long long ntc, const cpx<T> *tz, const cpx<T> *tm, const ZConst &zc,
                     double *pmax, const int *gate, const Batch &bt) {
#define SPK_ZP(PR)                                                                                           \
    case PR:                                                                                                 \
        k_zmid<T, NZ, PR, BAT><<<grid, nt, 0, st>>>(w, wr, pp, nrows, ntc, tz, tm, zc, pmax, gate, bt); \
        break;
    switch (prog) {
        SPK_ZP(ZP_U)
        SPK_ZP(ZP_U | ZP_B)
        SPK_ZP(ZP_U | ZP_TH)
        SPK_ZP(ZP_U | ZP_B | ZP_TH)
        SPK_ZP(ZP_U | ZP_B | ZP_TH | ZP_G)
        SPK_ZP(ZP_B | ZP_G)
        SPK_ZP(ZP_B | ZP_G | ZP_BR)
        default: set_error("fused z program %d not instantiated", prog); return SPK_ERR_STATE;
    }
#undef SPK_ZP
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

template <class T, int NZ>
int launch_zmid_nz(spk_solver s, int prog, const void *W, void *P, double *pmax, unsigned *nblocks_out) {
    constexpr int M = NZ / 2;
    constexpr int RW = zmid_rw<T>(M);
    constexpr int NT = RW * (M / zmid_e<T>(M));
    const long long nrows = (long long)s->g.nx * s->g.ny;
    const unsigned nb = (unsigned)((nrows + RW - 1) / RW);
    if (nblocks_out) *nblocks_out = nb;
    const cpx<T> *tz = twp<T>(s->twb, s->twb.off_z), *tm = twp<T>(s->twb, s->twb.off_z + s->g.nz);
    ZConst zc{1.0 / s->p.reynolds_ani, s->p.OmegaT2, s->p.strat_direction};
    const cpx<T> *w = (const cpx<T> *)W;
    const cpx<T> *wr = (const cpx<T> *)s->Breal;
    cpx<T> *pp = (cpx<T> *)P;
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {W, s->Breal, P, pmax, s->ctx->gate}, &bt));
    const dim3 grid(nb, 1, nboxes(s->ctx));
    if (bt.rows)
        return launch_zmid_prog<T, NZ, true>(prog, grid, NT, s->ctx->stream, w, wr, pp, nrows, s->g.ntc, tz, tm, zc,
                                             pmax, s->ctx->gate, bt);
    return launch_zmid_prog<T, NZ, false>(prog, grid, NT, s->ctx->stream, w, wr, pp, nrows, s->g.ntc, tz, tm, zc,
                                          pmax, s->ctx->gate, bt);
}

template <class T>
int launch_zmid(spk_solver s, int prog, const void *W, void *P, double *pmax, unsigned *nblocks_out) {
    switch (s->g.nz) {
        case 2: return launch_zmid_nz<T, 2>(s, prog, W, P, pmax, nblocks_out);
        case 4: return launch_zmid_nz<T, 4>(s, prog, W, P, pmax, nblocks_out);
        case 8: return launch_zmid_nz<T, 8>(s, prog, W, P, pmax, nblocks_out);
        case 16: return launch_zmid_nz<T, 16>(s, prog, W, P, pmax, nblocks_out);
        case 32: return launch_zmid_nz<T, 32>(s, prog, W, P, pmax, nblocks_out);
        case 64: return launch_zmid_nz<T, 64>(s, prog, W, P, pmax, nblocks_out);
        case 128: return launch_zmid_nz<T, 128>(s, prog, W, P, pmax, nblocks_out);
        case 256: return launch_zmid_nz<T, 256>(s, prog, W, P, pmax, nblocks_out);
        case 512: return launch_zmid_nz<T, 512>(s, prog, W, P, pmax, nblocks_out);
        case 1024: return launch_zmid_nz<T, 1024>(s, prog, W, P, pmax, nblocks_out);
        default: set_error("NZ=%d unsupported", s->g.nz); return SPK_ERR_ARG;
    }
}

template <class T, int MODE>
int launch_epilogue_mode(spk_solver s, const EpiConst &c) {
    const unsigned nb = nblk(s->g.ntc);
    hipStream_t st = s->ctx->stream;
    const cpx<T> *P = (const cpx<T> *)s->P;
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {P, s->U, s->S, s->dU, s->ts}, &bt));
    const dim3 grid(nb, 1, nboxes(s->ctx));
    switch (s->ly) {
        case LY_HEAT: k_epilogue<T, LY_HEAT, MODE><<<grid, 256, 0, st>>>(P, s->U, s->S, s->dU, c, s->ts, bt); break;
        case LY_HD: k_epilogue<T, LY_HD, MODE><<<grid, 256, 0, st>>>(P, s->U, s->S, s->dU, c, s->ts, bt); break;
        case LY_HD_TH: k_epilogue<T, LY_HD_TH, MODE><<<grid, 256, 0, st>>>(P, s->U, s->S, s->dU, c, s->ts, bt); break;
        case LY_MHD: k_epilogue<T, LY_MHD, MODE><<<grid, 256, 0, st>>>(P, s->U, s->S, s->dU, c, s->ts, bt); break;
        default: k_epilogue<T, LY_MHD_TH, MODE><<<grid, 256, 0, st>>>(P, s->U, s->S, s->dU, c, s->ts, bt); break;
    }
    return SPK_OK;
}

// mode 0: store dU; 1, 2, 3: RK3 stage update
template <class T>
int launch_epilogue(spk_solver s, const EpiConst &c, int mode) {
    switch (mode) {
        case 0: SPK_TRY((launch_epilogue_mode<T, 0>(s, c))); break;
        case 1: SPK_TRY((launch_epilogue_mode<T, 1>(s, c))); break;
        case 2: SPK_TRY((launch_epilogue_mode<T, 2>(s, c))); break;
        default: SPK_TRY((launch_epilogue_mode<T, 3>(s, c))); break;
    }
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

// Wray's low-storage RK3 coefficients (hydro_mhd_advance.cu:79-80)
static const double gammaRK[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double xiRK[2] = {-17.0 / 60.0, -5.0 / 12.0};

// ---- pipeline pieces ----

ZConst zconst(spk_solver s) { return ZConst{1.0 / s->p.reynolds_ani, s->p.OmegaT2, s->p.strat_direction}; }

// On a quasi-2D grid the inverse y pass, the z kernel and the forward y pass of
// the products are one kernel (k_plane): P then leaves the front already
// transformed along y, and the back end starts at its x pass.
bool plane_front(spk_solver s) { return s->plane_on && s->nin > 0 && plane_supported(s->g, s->zprog); }
bool plane_rkl(spk_solver s) { return s->plane_on && plane_rkl_supported(s->g); }

// products of the RHS inputs in W (inverse x pass done): inverse y pass + z
// kernel, or the plane kernel (which also makes the forward y pass)
int rhs_mid(spk_solver s, bool with_max) {
    const Grid3 &g = s->g;
    s->fused_w_ok = false;  // W no longer holds the last fused step's inverse x pass as it left it
    double *pm = with_max ? s->pmax : nullptr;
    if (plane_front(s)) {
        SPK_TRY(launch_plane(s->ctx, g, s->twb, s->zprog, s->W, s->P, zconst(s), pm));
        s->zmid_blocks = g.nx;
        return SPK_OK;
    }
    PassArgs a{s->W, s->W, s->nin, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->twb, +1, s->bulk, s->bulk, a));
    unsigned nb = 0;
    if (s->bulk == SPK_F64) SPK_TRY(launch_zmid<double>(s, s->zprog, s->W, s->P, pm, &nb));
    else SPK_TRY(launch_zmid<float>(s, s->zprog, s->W, s->P, pm, &nb));
    s->zmid_blocks = (int)nb;
    return SPK_OK;
}

// inverse transforms of the RHS inputs and the fused z kernel: products in P
int rhs_front(spk_solver s, double kxshift, bool with_max) {
    if (s->nin == 0) return SPK_OK;
    if (s->bulk == SPK_F64) SPK_TRY(launch_inv_x<double>(s, s->inv, s->nin, s->W, kxshift));
    else SPK_TRY(launch_inv_x<float>(s, s->inv, s->nin, s->W, kxshift));
    return rhs_mid(s, with_max);
}

EpiConst make_epi(spk_solver s, double kxshift, double gam, double xi) {
    EpiConst c;
    memset(&c, 0, sizeof(c));
    c.kg = kgeom(s, kxshift);
    c.ntc = s->g.ntc;
    c.stratification = s->p.stratification;
    c.sd = s->p.strat_direction;
    c.shearing = s->p.shearing;
    c.rot_active = s->p.rotating && s->p.shearing;
    c.aniso_rhs = s->aniso_rhs;
    c.diff_mask = s->diff_mask;
    c.nu = s->p.nu; c.nu_m = s->p.nu_m; c.nu_th = s->p.nu_th; c.N2 = s->p.N2;
    c.shear = s->p.shear; c.omega = s->p.omega;
    c.oT = s->oT; c.oE = s->oE; c.oQ = s->oQ; c.oA = s->oA;
    c.gam = gam;
    c.xi = xi;
    return c;
}

// forward transforms of the products and the spectral epilogue
int rhs_back(spk_solver s, double kxshift, int mode, double gam, double xi) {
    const Grid3 &g = s->g;
    if (s->nout > 0) {
        PassArgs a{s->P, s->P, s->nout, g.ntc, g.ntc, 1.0};
        if (!plane_front(s)) SPK_TRY(launch_pass_y(s->ctx, g, s->twb, -1, s->bulk, s->bulk, a));
        SPK_TRY(launch_pass_x(s->ctx, g, s->twb, -1, s->bulk, s->bulk, a));
    }
    const EpiConst c = make_epi(s, kxshift, gam, xi);
    if (s->bulk == SPK_F64) return launch_epilogue<double>(s, c, mode);
    return launch_epilogue<float>(s, c, mode);
}

// ---- fused path: forward y pass, then forward x + epilogue + RK3 (+ projection
// + next inverse x) in one kernel; the next stage starts at the inverse y pass ----

// Each field group of k_fwdx_epilogue updates its own fields in place; with
// stratification (u_s <-> theta) the velocity and temperature groups share their
// blocks.  Shearing (sheared k between stages, remap) and the in-RHS conduction
// (its grad theta slots come from k_inv_x) stay on the unfused back end.
bool fused_allowed(spk_solver s) {
    return s->nin > 0 && !s->aniso_rhs && !s->p.shearing && fused_supported_nx(s->p.nx) && s->ly != LY_HEAT;
}

bool fused_ok(spk_solver s) {
    return fused_allowed(s) && (s->fused_mode == 2 || (s->fused_mode == 1 && s->fused_faster));
}

// grid rows of k_fwdx_epilogue: velocity (+ temperature when stratified),
// magnetic, temperature
int fused_groups(spk_solver s) {
    return 1 + (s->p.mhd ? 1 : 0) + (s->p.boussinesq && !s->p.stratification ? 1 : 0);
}

int rhs_back_fused(spk_solver s, int mode, double gam, double xi, int project, int inv) {
    const Grid3 &g = s->g;
    PassArgs a{s->P, s->P, s->nout, g.ntc, g.ntc, 1.0};
    if (!plane_front(s)) SPK_TRY(launch_pass_y(s->ctx, g, s->twb, -1, s->bulk, s->bulk, a));
    FusedConst fc;
    memset(&fc, 0, sizeof(fc));
    fc.e = make_epi(s, 0.0, gam, xi);
    fc.project = project;
    fc.inv = inv;
    fc.scale = 1.0 / ((double)g.nx * g.ny * g.nz);
    // the increment register's coefficients (FusedConst): stage 2 forms U1 = U0 + g1 R0
    // and S1 = U0 + (g1 + x1) R0, stage 3 forms S2 = U2 + x2 R1
    fc.cu = mode == 2 ? gammaRK[0] : 0.0;
    fc.cs = mode == 2 ? gammaRK[0] + xiRK[0] : (mode == 3 ? xiRK[1] : 0.0);
    const void *tw = (const char *)s->twb.d + (size_t)s->twb.off_x * (s->bulk == SPK_F64 ? 16 : 8);
    return launch_fwdx_epilogue(s->ctx, s->bulk, g.nx, s->ly, mode, s->fused_ef, s->P, s->U, s->S, s->W, fc, tw,
                                s->ts, fused_groups(s));
}

// inverse y pass + fused z kernel from a W that already holds the inverse x pass
int rhs_front_from_w(spk_solver s, bool with_max) { return rhs_mid(s, with_max); }

int dt_kernel(spk_solver s) {
    DtConst c;
    memset(&c, 0, sizeof(c));
    const spk_params &p = s->p;
    c.incompressible = p.incompressible && !p.heat_equation;
    c.heat_equation = p.heat_equation;
    c.mhd = p.mhd; c.boussinesq = p.boussinesq; c.aniso = p.anisotropic_diffusion;
    c.rotating = p.rotating; c.shearing = p.shearing; c.supertimestepping = p.supertimestepping;
    c.safety_sts = p.safety_sts;
    for (int q = 0; q < 3; q++) {
        c.sts_v[q] = c.incompressible && s->sts_pos[s->VX + q] >= 0;
        c.sts_b[q] = p.mhd && s->sts_pos[s->BX + q] >= 0;
    }
    c.sts_th = (p.boussinesq || p.heat_equation) && s->sts_pos[s->TH] >= 0;
    c.kxmax = s->kxmax; c.kymax = s->kymax; c.kzmax = s->kzmax;
    c.nu = p.nu; c.nu_m = p.nu_m; c.nu_th = p.nu_th; c.reynolds_ani = p.reynolds_ani;
    c.N2 = p.N2; c.OmegaT2 = p.OmegaT2; c.omega = p.omega; c.shear = p.shear;
    c.cfl = p.cfl; c.cfl_par = p.cfl_par; c.cfl_sts = p.cfl_sts; c.safety_source = p.safety_source;
    c.t_final = p.t_final; c.toutput_flow = p.toutput_flow;
    c.nblocks = s->zmid_blocks;
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {s->pmax, s->ts}, &bt));
    k_dt<<<dim3(1, 1, nboxes(s->ctx)), 256, 0, s->ctx->stream>>>(s->pmax, c, s->ts, bt);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int clock(spk_solver s, int add_dt, int add_step, int snap, int set_tremap, double tremap) {
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {s->ts}, &bt));
    k_clock<<<dim3(1, 1, nboxes(s->ctx)), 1, 0, s->ctx->stream>>>(s->ts, add_dt, add_step, snap, s->p.toutput_flow,
                                                                   set_tremap, tremap, bt);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int read_time(spk_solver s, TimeState *h) {
    SPK_HIP(hipMemcpyAsync(h, s->ts, sizeof(TimeState), hipMemcpyDeviceToHost, s->ctx->stream));
    SPK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SPK_OK;
}

int clean_divergence(spk_solver s) {
    const int f0 = (s->p.incompressible && !s->p.heat_equation) ? s->VEL : -1;
    const int f1 = s->p.mhd && f0 >= 0 ? s->MAG : -1;
    if (f0 < 0 && f1 < 0) return SPK_OK;
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {s->U}, &bt));
    k_clean_div<<<dim3(nblk(s->g.ntc), 1, nboxes(s->ctx)), 256, 0, s->ctx->stream>>>(
        s->U, s->g.ntc, kgeom(s, s->tremap * s->p.shear), f0, f1, bt);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int compute_dt(spk_solver s) {
    if (s->p.incompressible && !s->p.heat_equation) {
        SPK_TRY(rhs_front(s, s->tremap * s->p.shear, true));
        s->products_valid = true;
        s->products_src = s->U;
    }
    return dt_kernel(s);
}


int shift_time_host(spk_solver s, double t) {
    const spk_params &p = s->p;
    s->tremap = fmod(t + p.ly / (2.0 * p.shear * p.lx), p.ly / (p.shear * p.lx)) - p.ly / (2.0 * p.shear * p.lx);
    return SPK_OK;
}

int remap_all_fields(spk_solver s) {
    const Grid3 &g = s->g;
    cpx<double> *tmp = reinterpret_cast<cpx<double> *>(s->realbuf);
    for (int f = 0; f < s->nf; f++) {
        cpx<double> *v = s->U + (long long)f * g.ntc;
        k_zero<<<nblk(g.ntc), 256, 0, s->ctx->stream>>>(tmp, g.ntc);
        k_remap<<<nblk(g.ntc), 256, 0, s->ctx->stream>>>(v, tmp, g.ntc, g.nx, g.ny, g.nzh);
        k_mask_copy<<<nblk(g.ntc), 256, 0, s->ctx->stream>>>(tmp, v, g.ntc, kgeom(s, 0.0));
        SPK_LAUNCH_CHECK();
    }
    return SPK_OK;
}

int rk3(spk_solver s) {
    const spk_params &p = s->p;
    const bool shear = p.shearing;
    double dt = 0.0;
    if (shear) {
        TimeState h;
        SPK_TRY(read_time(s, &h));
        dt = h.dt;
    }
    const double ks0 = s->tremap * p.shear;
    // stage 1
    if (!s->products_valid || s->products_src != s->U) SPK_TRY(rhs_front(s, ks0, false));
    s->products_valid = false;
    SPK_TRY(rhs_back(s, ks0, 1, gammaRK[0], xiRK[0]));
    // stage 2
    const double ks1 = shear ? (s->tremap + gammaRK[0] * dt) * p.shear : ks0;
    SPK_TRY(rhs_front(s, ks1, false));
    SPK_TRY(rhs_back(s, ks1, 2, gammaRK[1], xiRK[1]));
    // stage 3
    const double ks2 = shear ? (s->tremap + (gammaRK[0] + xiRK[0] + gammaRK[1]) * dt) * p.shear : ks0;
    SPK_TRY(rhs_front(s, ks2, false));
    SPK_TRY(rhs_back(s, ks2, 3, gammaRK[2], 0.0));
    SPK_TRY(clock(s, 1, 0, 0, 0, 0.0));
    if (shear) {
        TimeState h;
        SPK_TRY(read_time(s, &h));
        s->tremap += dt;
        if (s->tremap > p.ly / (2.0 * p.shear * p.lx)) {
            SPK_TRY(shift_time_host(s, h.t));
            SPK_TRY(remap_all_fields(s));
        }
        SPK_TRY(clock(s, 0, 0, 0, 1, s->tremap));
    }
    return SPK_OK;
}

// RKL2 coefficient recurrence (rkl.cu:466-585); returns the stage count
int rkl2_coeffs(double dt_current, double dt_sts, std::vector<double> &coef) {
    const double scrh = dt_current / dt_sts;
    const double s_str = 4.0 * (1.0 + 2.0 * scrh) / (1.0 + sqrt(9.0 + 16.0 * scrh));
    const int s_RKL = 1 + (int)(s_str);
    const double w1 = 4.0 / (s_RKL * s_RKL + s_RKL - 2.0);
    double mu_tilde_j = w1 / 3.0;
    double b_j, b_jm1, b_jm2, a_jm1;
    b_j = b_jm1 = b_jm2 = 1.0 / 3.0;
    a_jm1 = 1.0 - b_jm1;
    coef.assign(4 * (size_t)(s_RKL > 0 ? s_RKL : 1), 0.0);
    coef[2] = mu_tilde_j;
    for (int stage = 2; stage <= s_RKL; stage++) {
        const double mu_j = (2. * stage - 1.) / stage * b_j / b_jm1;
        mu_tilde_j = w1 * mu_j;
        const double gamma_j = -a_jm1 * mu_tilde_j;
        const double nu_j = -(stage - 1.) * b_j / (stage * b_jm2);
        double *c = &coef[4 * (size_t)(stage - 1)];
        c[0] = mu_j; c[1] = nu_j; c[2] = mu_tilde_j; c[3] = gamma_j;
        b_jm2 = b_jm1;
        b_jm1 = b_j;
        a_jm1 = 1.0 - b_jm1;
        b_j = 0.5 * (stage * stage + 3.0 * stage) / (stage * stage + 3.0 * stage + 2);
    }
    return s_RKL;
}

template <class T>
int rkl2_stage_launch(spk_solver s, int stage, double mu, double nu, double dtmut, double gdt) {
    RklConst c = s->rkl;
    c.kg = kgeom(s, s->tremap * s->p.shear);
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {s->P2, s->U, s->rUc0, s->rUc1, s->rdU0}, &bt));
    k_rkl2<T><<<dim3(nblk(s->g.ntc), 1, nboxes(s->ctx)), 256, 0, s->ctx->stream>>>(
        (const cpx<T> *)s->P2, s->U, s->rUc0, s->rUc1, s->rdU0, c, stage, mu, nu, dtmut, gdt, nullptr, nullptr, bt);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int rkl2_front(spk_solver s) {
    // anisotropic flux of theta from i k theta and B (AnisotropicHeatFlux,
    // anisotropic_conduction.cu:10-62), into P2
    const Grid3 &g = s->g;
    const double ks = s->tremap * s->p.shear;
    if (s->bulk == SPK_F64) SPK_TRY(launch_inv_x<double>(s, s->inv_rkl, 6, s->W2, ks));
    else SPK_TRY(launch_inv_x<float>(s, s->inv_rkl, 6, s->W2, ks));
    PassArgs a{s->W2, s->W2, 6, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->twb, +1, s->bulk, s->bulk, a));
    if (s->bulk == SPK_F64) SPK_TRY(launch_zmid<double>(s, ZP_B | ZP_G, s->W2, s->P2, nullptr, nullptr));
    else SPK_TRY(launch_zmid<float>(s, ZP_B | ZP_G, s->W2, s->P2, nullptr, nullptr));
    PassArgs b{s->P2, s->P2, 3, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->twb, -1, s->bulk, s->bulk, b));
    SPK_TRY(launch_pass_x(s->ctx, g, s->twb, -1, s->bulk, s->bulk, b));
    return SPK_OK;
}

// Real-space B rows for a whole RKL2 cycle (B is not an sts variable there):
// inverse x, inverse y and z C2R of the three B fields into Breal.
int rkl2_b_real(spk_solver s) {
    const Grid3 &g = s->g;
    if (s->breal_from_w != 0 && s->wslot_b >= 0) {
        // B's inverse x pass is already in W (the step's RHS program): no pass over
        // the state, and W itself is left as it is
        const void *wb = (const char *)s->W + (size_t)s->wslot_b * (size_t)g.ntc * esize(s->bulk);
        if (s->breal_from_w == 2) {
            PassArgs a{wb, s->W2, 3, g.ntc, g.ntc, 1.0};
            SPK_TRY(launch_pass_y(s->ctx, g, s->twb, +1, s->bulk, s->bulk, a));
            wb = s->W2;
        }
        PassArgs z{wb, s->Breal, 3, g.ntc, g.ntc, 1.0};
        return launch_pass_z_c2r(s->ctx, g, s->twb, s->bulk, s->bulk, z);
    }
    const double ks = s->tremap * s->p.shear;
    if (s->bulk == SPK_F64) SPK_TRY(launch_inv_x<double>(s, s->inv_rkl_b, 3, s->W2, ks));
    else SPK_TRY(launch_inv_x<float>(s, s->inv_rkl_b, 3, s->W2, ks));
    PassArgs a{s->W2, s->W2, 3, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->twb, +1, s->bulk, s->bulk, a));
    PassArgs z{s->W2, s->Breal, 3, g.ntc, g.ntc, 1.0};
    return launch_pass_z_c2r(s->ctx, g, s->twb, s->bulk, s->bulk, z);
}

// As rkl2_front with B taken from Breal: only i k theta is transformed per stage.
int rkl2_front_cached(spk_solver s) {
    const Grid3 &g = s->g;
    const double ks = s->tremap * s->p.shear;
    const size_t fb = (size_t)g.ntc * esize(s->bulk);
    void *wg = (char *)s->W2 + 3 * fb;  // the i k theta slots of the fused z program
    if (s->bulk == SPK_F64) SPK_TRY(launch_inv_x<double>(s, s->inv_rkl_g, 3, wg, ks));
    else SPK_TRY(launch_inv_x<float>(s, s->inv_rkl_g, 3, wg, ks));
    PassArgs a{wg, wg, 3, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->twb, +1, s->bulk, s->bulk, a));
    if (s->bulk == SPK_F64) SPK_TRY(launch_zmid<double>(s, ZP_B | ZP_G | ZP_BR, s->W2, s->P2, nullptr, nullptr));
    else SPK_TRY(launch_zmid<float>(s, ZP_B | ZP_G | ZP_BR, s->W2, s->P2, nullptr, nullptr));
    PassArgs b{s->P2, s->P2, 3, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->twb, -1, s->bulk, s->bulk, b));
    SPK_TRY(launch_pass_x(s->ctx, g, s->twb, -1, s->bulk, s->bulk, b));
    return SPK_OK;
}

// The fused RKL2 stage (spk_rklfused.hip): anisotropic conduction of theta as
// the only sts variable, B's real-space rows cached for the cycle, on a grid the
// fused x kernel is built for, and the fused back end not switched off.
bool rkl_fused(spk_solver s) {
    return s->rkl_aniso && s->rkl_cache_b && s->nsts == 1 && !s->p.shearing && rkl_fused_supported_nx(s->p.nx) &&
           s->fused_mode != 0;
}

// the part of a fused RKL2 stage before its forward x pass: (stage 1 only) the
// inverse x pass of i k_x theta and theta, then the inverse y pass that makes
// the three gradient components, the flux in the z kernel and the forward y
// pass that folds k_y F_y + k_z F_z into one slot
int rkl2_front_fused(spk_solver s, bool with_inv_x) {
    const Grid3 &g = s->g;
    const size_t fb = (size_t)g.ntc * esize(s->bulk);
    void *wg = (char *)s->W2 + 3 * fb;  // the grad theta slots of the fused z program
    if (with_inv_x) {
        if (s->bulk == SPK_F64) SPK_TRY(launch_inv_x<double>(s, s->inv_rkl_t, 2, wg, 0.0));
        else SPK_TRY(launch_inv_x<float>(s, s->inv_rkl_t, 2, wg, 0.0));
    }
    const KGeom kg = kgeom(s, 0.0);
    if (plane_rkl(s)) return launch_plane_rkl(s->ctx, g, s->twb, wg, s->Breal, s->P2, kg, zconst(s));
    SPK_TRY(launch_pass_y_rkl(s->ctx, g, s->twb, +1, wg, kg));
    if (s->bulk == SPK_F64) SPK_TRY(launch_zmid<double>(s, ZP_B | ZP_G | ZP_BR, s->W2, s->P2, nullptr, nullptr));
    else SPK_TRY(launch_zmid<float>(s, ZP_B | ZP_G | ZP_BR, s->W2, s->P2, nullptr, nullptr));
    return launch_pass_y_rkl(s->ctx, g, s->twb, -1, s->P2, kg);
}

// forward x pass of the flux + stage update + (inv) the next stage's inverse x pass
int rkl2_back_fused(spk_solver s, int stage, double mu, double nu, double dtmut, double gdt, int inv,
                    const RklPlan *plan) {
    const Grid3 &g = s->g;
    const size_t fb = (size_t)g.ntc * esize(s->bulk);
    RklConst c = s->rkl;
    c.kg = kgeom(s, 0.0);
    const void *tw = (const char *)s->twb.d + (size_t)s->twb.off_x * esize(s->bulk);
    return launch_fwdx_rkl2(s->ctx, s->bulk, g.nx, s->P2, s->U, s->rUc0, s->rUc1, s->rdU0, (char *)s->W2 + 3 * fb, c,
                            stage, mu, nu, dtmut, gdt, inv, 1.0 / ((double)g.nx * g.ny * g.nz), tw, plan,
                            plan ? s->ts : nullptr);
}

int rkl2_cycle(spk_solver s) {
    if (s->nsts == 0) return SPK_OK;
    TimeState h;
    SPK_TRY(read_time(s, &h));
    std::vector<double> coef;
    const int sr = rkl2_coeffs(h.dt, h.dt_sts, coef);
    if (s->rkl_aniso && s->rkl_cache_b) SPK_TRY(rkl2_b_real(s));
    if (rkl_fused(s)) {
        for (int stage = 1; stage <= sr; stage++) {
            SPK_TRY(rkl2_front_fused(s, stage == 1));
            const double *c = &coef[4 * (size_t)(stage - 1)];
            SPK_TRY(rkl2_back_fused(s, stage, c[0], c[1], c[2] * h.dt, c[3] * h.dt, stage < sr, nullptr));
        }
        return SPK_OK;
    }
    for (int stage = 1; stage <= sr; stage++) {
        if (s->rkl_aniso) SPK_TRY(s->rkl_cache_b ? rkl2_front_cached(s) : rkl2_front(s));
        const double *c = &coef[4 * (size_t)(stage - 1)];
        const double mu = c[0], nu = c[1], mut = c[2], gam = c[3];
        const double dtmut = stage == 1 ? mut * h.dt : h.dt * mut;
        if (s->bulk == SPK_F64) SPK_TRY(rkl2_stage_launch<double>(s, stage, mu, nu, dtmut, gam * h.dt));
        else SPK_TRY(rkl2_stage_launch<float>(s, stage, mu, nu, dtmut, gam * h.dt));
    }
    return SPK_OK;
}

// rkl2_cycle with the stage count and coefficients planned on the device:
// RKL_SMAX-bounded stages are launched (rkl_smax, the stage count at the sts cap
// dt = safety_sts dt_sts) and those beyond the step's count return at once.
int rkl2_cycle_device(spk_solver s) {
    if (s->nsts == 0) return SPK_OK;
    hipStream_t st = s->ctx->stream;
    Batch bp;
    SPK_TRY(make_batch(s->ctx, {s->ts, s->rkl_plan}, &bp));
    k_rkl2_plan<<<dim3(1, 1, nboxes(s->ctx)), 64, 0, st>>>(s->ts, s->rkl_smax, s->rkl_plan, bp);
    SPK_LAUNCH_CHECK();
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {s->P2, s->U, s->rUc0, s->rUc1, s->rdU0, s->rkl_plan, s->ts}, &bt));
    const dim3 grid(nblk(s->g.ntc), 1, nboxes(s->ctx));
    if (s->rkl_aniso && s->rkl_cache_b) SPK_TRY(rkl2_b_real(s));
    if (rkl_fused(s)) {
        for (int stage = 1; stage <= s->rkl_smax; stage++) {
            s->ctx->gate = &s->rkl_plan->active[stage];  // a device address; never read on the host
            const int rc = rkl2_front_fused(s, stage == 1);
            s->ctx->gate = nullptr;
            SPK_TRY(rc);
            SPK_TRY(rkl2_back_fused(s, stage, 0.0, 0.0, 0.0, 0.0, stage < s->rkl_smax, s->rkl_plan));
        }
        return SPK_OK;
    }
    RklConst c = s->rkl;
    c.kg = kgeom(s, s->tremap * s->p.shear);
    for (int stage = 1; stage <= s->rkl_smax; stage++) {
        if (s->rkl_aniso) {
            s->ctx->gate = &s->rkl_plan->active[stage];  // a device address; never read on the host
            const int rc = s->rkl_cache_b ? rkl2_front_cached(s) : rkl2_front(s);
            s->ctx->gate = nullptr;
            SPK_TRY(rc);
        }
        if (s->bulk == SPK_F64)
            k_rkl2<double><<<grid, 256, 0, st>>>((const cpx<double> *)s->P2, s->U, s->rUc0, s->rUc1, s->rdU0, c,
                                                 stage, 0.0, 0.0, 0.0, 0.0, s->rkl_plan, s->ts, bt);
        else
            k_rkl2<float><<<grid, 256, 0, st>>>((const cpx<float> *)s->P2, s->U, s->rUc0, s->rUc1, s->rdU0, c,
                                                stage, 0.0, 0.0, 0.0, 0.0, s->rkl_plan, s->ts, bt);
        SPK_LAUNCH_CHECK();
    }
    return SPK_OK;
}

int hydro_mhd_advance(spk_solver s) {
    const bool sts = s->p.supertimestepping && s->nsts > 0;
    auto cycle = [&]() { return s->rkl_device ? rkl2_cycle_device(s) : rkl2_cycle(s); };
    SPK_TRY(compute_dt(s));
    if (sts && (s->step % 2) == 1) SPK_TRY(cycle());
    SPK_TRY(rk3(s));
    if (sts && (s->step % 2) == 0) SPK_TRY(cycle());
    s->step += 1;
    return clock(s, 0, 1, 0, 0, 0.0);
}

// Tile autotune of the back end of one RK3 stage on this grid: the fused kernel
// with 4 or 8 points per thread against the unfused forward x pass + epilogue +
// next inverse x pass (+ a third of the projection pass).  Zero state, zero
// products and a zero step keep every buffer at zero.
int tune_fused(spk_solver s) {
    s->fused_ef = 4;
    s->fused_ef_auto = 4;
    s->fused_faster = false;
    if (!fused_allowed(s)) return SPK_OK;
    const Grid3 &g = s->g;
    hipStream_t st = s->ctx->stream;
    const size_t fb = (size_t)g.ntc * esize(s->bulk);
    SPK_HIP(hipMemsetAsync(s->P, 0, fb * s->nout, st));
    SPK_HIP(hipMemsetAsync(s->W, 0, fb * s->nin, st));
    FusedConst fc;
    memset(&fc, 0, sizeof(fc));
    fc.e = make_epi(s, 0.0, 0.0, 0.0);
    fc.inv = 1;
    fc.scale = 1.0 / ((double)g.nx * g.ny * g.nz);
    const void *tw = (const char *)s->twb.d + (size_t)s->twb.off_x * (s->bulk == SPK_F64 ? 16 : 8);
    hipEvent_t e0, e1;
    SPK_HIP(hipEventCreate(&e0));
    SPK_HIP(hipEventCreate(&e1));
    auto time_it = [&](auto &&launch, float *ms) -> int {
        for (int r = 0; r < 4; r++) {
            if (r == 1) SPK_HIP(hipEventRecord(e0, st));
            SPK_TRY(launch());
        }
        SPK_HIP(hipEventRecord(e1, st));
        SPK_HIP(hipEventSynchronize(e1));
        SPK_HIP(hipEventElapsedTime(ms, e0, e1));
        return SPK_OK;
    };
    float best = 1e30f, t_unf = 0.f, t_cl = 0.f;
    int rc = SPK_OK;
    for (int ef : {4, 8}) {
        float ms = 0.f;
        rc = time_it([&] {
            return launch_fwdx_epilogue(s->ctx, s->bulk, g.nx, s->ly, 2, ef, s->P, s->U, s->S, s->W, fc, tw, s->ts,
                                        fused_groups(s));
        }, &ms);
        if (rc != SPK_OK) break;
        if (ms < best) {
            best = ms;
            s->fused_ef = ef;
        }
    }
    if (rc == SPK_OK) {
        const EpiConst c = make_epi(s, 0.0, 0.0, 0.0);
        PassArgs ap{s->P, s->P, s->nout, g.ntc, g.ntc, 1.0};
        rc = time_it([&] {
            SPK_TRY(launch_pass_x(s->ctx, g, s->twb, -1, s->bulk, s->bulk, ap));
            SPK_TRY(s->bulk == SPK_F64 ? launch_epilogue<double>(s, c, 2) : launch_epilogue<float>(s, c, 2));
            return s->bulk == SPK_F64 ? launch_inv_x<double>(s, s->inv, s->nin, s->W, 0.0)
                                      : launch_inv_x<float>(s, s->inv, s->nin, s->W, 0.0);
        }, &t_unf);
    }
    if (rc == SPK_OK) rc = time_it([&] { return clean_divergence(s); }, &t_cl);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SPK_TRY(rc);
    s->fused_faster = best < t_unf + t_cl / 3.0f;
    s->fused_ef_auto = s->fused_ef;
    return SPK_OK;
}

// One main-loop iteration on the fused path (HydroMHDAdvance,
// hydro_mhd_advance.cu:18-75).  w_valid: W holds the inverse x pass of the
// current, already projected state (written by the previous step's stage 3);
// otherwise CleanFieldDivergence and the inverse x pass run first.  The RKL2
// cycle of a super-time-stepped run sits where the unfused path has it: on odd
// steps after compute_dt, whose real fields (here the stage-1 products) it
// leaves as they were, and on even steps after RK3, after which the changed
// fields' inverse x pass in W is made again (theta alone) or left to the next
// step (s->fused_w_ok).
int step_fused(spk_solver s, bool w_valid) {
    if (!w_valid) {
        SPK_TRY(clean_divergence(s));
        if (s->bulk == SPK_F64) SPK_TRY(launch_inv_x<double>(s, s->inv, s->nin, s->W, 0.0));
        else SPK_TRY(launch_inv_x<float>(s, s->inv, s->nin, s->W, 0.0));
    }
    const bool sts = s->p.supertimestepping && s->nsts > 0;
    auto cycle = [&]() { return s->rkl_device ? rkl2_cycle_device(s) : rkl2_cycle(s); };
    // compute_dt: the stage-1 transforms carry the CFL maxima
    SPK_TRY(rhs_front_from_w(s, true));
    SPK_TRY(dt_kernel(s));
    // W holds the inverse x pass of the current state (and, after the separate
    // inverse y pass, its y pass too): the cycle takes B's real rows from there
    auto cycle_from_w = [&](int mode) {
        s->breal_from_w = mode;
        const int rc = cycle();
        s->breal_from_w = 0;
        return rc;
    };
    if (sts && (s->step % 2) == 1) SPK_TRY(cycle_from_w(plane_front(s) ? 2 : 1));
    SPK_TRY(rhs_back_fused(s, 1, gammaRK[0], xiRK[0], 0, 1));
    SPK_TRY(rhs_front_from_w(s, false));
    SPK_TRY(rhs_back_fused(s, 2, gammaRK[1], xiRK[1], 0, 1));
    SPK_TRY(rhs_front_from_w(s, false));
    SPK_TRY(rhs_back_fused(s, 3, gammaRK[2], 0.0, 1, 1));
    SPK_TRY(clock(s, 1, 0, 0, 0, 0.0));
    s->fused_w_ok = true;
    if (sts && (s->step % 2) == 0) {
        SPK_TRY(cycle_from_w(2));  // W: stage 3's inverse x pass of the new state
        if (s->nsts == 1 && s->sts_index[0] == s->TH) {
            // theta's W slot is its field index (the RHS program's slot order)
            const Grid3 &g = s->g;
            void *wt = (char *)s->W + (size_t)s->TH * (size_t)g.ntc * esize(s->bulk);
            if (s->bulk == SPK_F64) SPK_TRY(launch_inv_x<double>(s, s->inv_th, 1, wt, 0.0));
            else SPK_TRY(launch_inv_x<float>(s, s->inv_th, 1, wt, 0.0));
        } else {
            s->fused_w_ok = false;  // u or B changed: project and transform at the next step
        }
    }
    s->step += 1;
    s->products_valid = false;
    return clock(s, 0, 1, 0, 0, 0.0);
}

void free_all(spk_solver s) {
    void *bufs[] = {s->U,    s->S,    s->dU,   s->realbuf,  s->W,  s->P,    s->W2,   s->P2,  s->Breal,     s->diag,
                    s->rUc0, s->rUc1, s->rdU0, s->rkl_plan, s->ts, s->pmax, s->part, s->res, s->spec, s->force_idx,
                    s->force_val};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    twiddles_destroy(&s->twb);
    twiddles_destroy(&s->tw64);
}

template <class TPtr>
int dalloc(TPtr **p, size_t bytes) {
    hipError_t e = hipMalloc((void **)p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return SPK_ERR_ALLOC;
    }
    return SPK_OK;
}

int create(spk_ctx ctx, const spk_params *pin, int bulk, spk_solver s) {
    const spk_params &p = *pin;
    s->ctx = ctx;
    s->p = p;
    s->bulk = bulk;
    SPK_REQUIRE(is_pow2(p.nx) && is_pow2(p.ny) && is_pow2(p.nz) && p.nx >= 2 && p.ny >= 2 && p.nz >= 2 &&
                    p.nx <= 2048 && p.ny <= 2048 && p.nz <= 1024,
                SPK_ERR_ARG, "unsupported grid %dx%dx%d (powers of two, NX,NY <= 2048, NZ <= 1024)", p.nx, p.ny, p.nz);
    SPK_REQUIRE(bulk == SPK_F64 || bulk == SPK_F32, SPK_ERR_ARG, "bad bulk precision %d", bulk);
    SPK_REQUIRE(p.heat_equation || p.incompressible, SPK_ERR_ARG, "neither incompressible nor heat_equation set");
    SPK_REQUIRE(!p.anisotropic_diffusion || (p.mhd && p.boussinesq), SPK_ERR_ARG,
                "anisotropic_diffusion needs mhd and boussinesq (the flux is along B)");
    SPK_REQUIRE(p.strat_direction >= 0 && p.strat_direction <= 2, SPK_ERR_ARG, "bad strat_direction");
    SPK_REQUIRE(!p.shearing || p.shear != 0.0, SPK_ERR_ARG, "shearing needs a nonzero shear");
    s->g = make_grid(p.nx, p.ny, p.nz);
    // Parameters::popVariablesGrid (parameters.cu:568-640)
    s->VX = s->VY = s->VZ = s->BX = s->BY = s->BZ = s->TH = s->VEL = s->MAG = 0;
    if (p.heat_equation) {
        s->TH = 0; s->nf = 1; s->ly = LY_HEAT;
    } else {
        s->VX = 0; s->VY = 1; s->VZ = 2; s->VEL = 0; s->nf = 3; s->ly = LY_HD;
        if (p.mhd) {
            s->BX = 3; s->BY = 4; s->BZ = 5; s->MAG = 3; s->nf = 6; s->ly = LY_MHD;
            if (p.boussinesq) { s->TH = 6; s->nf = 7; s->ly = LY_MHD_TH; }
        } else if (p.boussinesq) {
            s->TH = 3; s->nf = 4; s->ly = LY_HD_TH;
        }
    }
    const bool inc = !p.heat_equation;
    const bool has_th = p.heat_equation || p.boussinesq;
    // RKLegendre ctor (rkl.cu:40-57): sts variables
    s->nsts = 0;
    for (int f = 0; f < 8; f++) s->sts_pos[f] = -1;
    for (int f = 0; f < s->nf; f++)
        if (p.supertimestepping && (p.sts_mask >> f & 1)) {
            s->sts_index[s->nsts] = f;
            s->sts_pos[f] = s->nsts++;
        }
    s->diff_mask = 0;
    for (int f = 0; f < s->nf; f++)
        if (s->sts_pos[f] < 0) s->diff_mask |= 1 << f;
    s->aniso_rhs = has_th && p.anisotropic_diffusion && s->sts_pos[s->TH] < 0;
    // wavevector.cu:70-110
    double kxmax = 2.0 * M_PI / p.lx * ((p.nx / 2) - 1);
    double kymax = 2.0 * M_PI / p.ly * ((p.ny / 2) - 1);
    double kzmax = 2.0 * M_PI / p.lz * ((p.nz / 2) - 1);
    if (p.with2d) kzmax = 0.0;
    KGeom &kg = s->kg0;
    kg.nx = p.nx; kg.ny = p.ny; kg.nzh = s->g.nzh;
    kg.kfx = (2.0 * M_PI) / p.lx;
    kg.kfy = (2.0 * M_PI) / p.ly;
    kg.kfz = p.with2d ? 0.0 : (2.0 * M_PI) / p.lz;
    kg.thx = 2.0 / 3.0 * kxmax; kg.thy = 2.0 / 3.0 * kymax; kg.thz = 2.0 / 3.0 * kzmax;
    kg.antialias = p.antialiasing;
    kg.kxshift = 0.0;
    if (p.antialiasing) {
        kxmax = (2.0 / 3.0) * kxmax;
        kymax = (2.0 / 3.0) * kymax;
        kzmax = (2.0 / 3.0) * kzmax;
    }
    s->kxmax = kxmax; s->kymax = kymax; s->kzmax = kzmax;
    s->kmax = pow(kxmax * kxmax + kymax * kymax + kzmax * kzmax, 0.5);
    double lmin = p.lx;
    if (p.ly < lmin) lmin = p.ly;
    if (p.lz < lmin) lmin = p.lz;
    s->deltak = 2.0 * M_PI / lmin;
    s->nbins = (int)(s->kmax / s->deltak + 0.5);
    // RHS program
    memset(&s->inv, 0, sizeof(s->inv));
    int ns = 0;
    s->zprog = 0;
    if (inc) {
        for (int q = 0; q < 3; q++) s->inv.s[ns++] = InvSlot{s->VX + q, -1};
        s->zprog |= ZP_U;
        if (p.mhd) {
            for (int q = 0; q < 3; q++) s->inv.s[ns++] = InvSlot{s->BX + q, -1};
            s->zprog |= ZP_B;
        }
        if (p.boussinesq) {
            s->inv.s[ns++] = InvSlot{s->TH, -1};
            s->zprog |= ZP_TH;
        }
        if (s->aniso_rhs) {
            for (int q = 0; q < 3; q++) s->inv.s[ns++] = InvSlot{s->TH, q};
            s->zprog |= ZP_G;
        }
    }
    s->nin = ns;
    s->wslot_b = -1;
    for (int q = 0; q < ns; q++)
        if (s->inv.s[q].src == s->BX && s->inv.s[q].deriv < 0 && s->wslot_b < 0) s->wslot_b = q;
    s->oT = 0;
    s->oE = inc ? 5 : 0;
    s->oQ = s->oE + (inc && p.mhd ? 3 : 0);
    s->oA = s->oQ + (inc && p.boussinesq ? 3 : 0);
    s->nout = inc ? s->oA + (s->aniso_rhs ? 3 : 0) : 0;
    // RKL2 program
    s->rkl_aniso = has_th && p.anisotropic_diffusion && s->sts_pos[s->TH] >= 0;
    memset(&s->inv_rkl, 0, sizeof(s->inv_rkl));
    for (int q = 0; q < 3; q++) s->inv_rkl.s[q] = InvSlot{s->BX + q, -1};
    for (int q = 0; q < 3; q++) s->inv_rkl.s[3 + q] = InvSlot{s->TH, q};
    memset(&s->inv_rkl_b, 0, sizeof(s->inv_rkl_b));
    memset(&s->inv_rkl_g, 0, sizeof(s->inv_rkl_g));
    memset(&s->inv_th, 0, sizeof(s->inv_th));
    memset(&s->inv_rkl_t, 0, sizeof(s->inv_rkl_t));
    s->inv_rkl_t.s[0] = InvSlot{s->TH, 0};
    s->inv_rkl_t.s[1] = InvSlot{s->TH, -1};
    s->inv_th.s[0] = InvSlot{s->TH, -1};
    for (int q = 0; q < 3; q++) s->inv_rkl_b.s[q] = InvSlot{s->BX + q, -1};
    for (int q = 0; q < 3; q++) s->inv_rkl_g.s[q] = InvSlot{s->TH, q};
    memset(&s->rkl, 0, sizeof(s->rkl));
    s->rkl.ntc = s->g.ntc;
    s->rkl.nsts = s->nsts;
    s->rkl.nu = p.nu; s->rkl.nu_m = p.nu_m; s->rkl.nu_th = p.nu_th;
    for (int n = 0; n < s->nsts; n++) {
        const int f = s->sts_index[n];
        s->rkl.field[n] = f;
        if (has_th && f == s->TH) s->rkl.kind[n] = p.anisotropic_diffusion ? 3 : 2;
        else if (p.mhd && f >= s->BX && f < s->BX + 3 && inc) s->rkl.kind[n] = 1;
        else s->rkl.kind[n] = 0;
    }
    // buffers
    const long long N = s->g.ntc;
    const size_t f64 = (size_t)N * 16, fb = (size_t)N * esize(bulk);
    SPK_TRY(dalloc(&s->U, f64 * s->nf));
    SPK_TRY(dalloc(&s->S, f64 * s->nf));
    SPK_TRY(dalloc(&s->dU, f64 * s->nf));
    SPK_TRY(dalloc(&s->realbuf, f64 * s->nf));
    SPK_TRY(dalloc(&s->W, fb * (s->nin > 0 ? s->nin : 1)));
    SPK_TRY(dalloc(&s->P, fb * (s->nout > 0 ? s->nout : 1)));
    if (s->rkl_aniso) {
        SPK_TRY(dalloc(&s->W2, fb * 6));
        SPK_TRY(dalloc(&s->P2, fb * 3));
        s->rkl_cache_b = s->sts_pos[s->BX] < 0 && s->sts_pos[s->BX + 1] < 0 && s->sts_pos[s->BX + 2] < 0;
        if (s->rkl_cache_b) SPK_TRY(dalloc(&s->Breal, fb * 3));
    }
    if (s->nsts > 0) {
        SPK_TRY(dalloc(&s->rUc0, f64 * s->nsts));
        SPK_TRY(dalloc(&s->rUc1, f64 * s->nsts));
        SPK_TRY(dalloc(&s->rdU0, f64 * s->nsts));
        SPK_TRY(dalloc(&s->rkl_plan, sizeof(RklPlan)));
        // stage count at r = dt / dt_sts = safety_sts (the cap of compute_dt), with a
        // margin for the rounding of that quotient
        const double r = (double)p.safety_sts * (1.0 + 1e-12);
        s->rkl_smax = 1 + (int)(4.0 * (1.0 + 2.0 * r) / (1.0 + sqrt(9.0 + 16.0 * r)));
        SPK_HIP(hipMemsetAsync(s->rkl_plan, 0, sizeof(RklPlan), ctx->stream));
    }
    SPK_TRY(dalloc(&s->ts, sizeof(TimeState)));
    const long long rows = (long long)p.nx * p.ny;
    SPK_TRY(dalloc(&s->pmax, sizeof(double) * 6 * (size_t)(rows + 1)));
    SPK_TRY(dalloc(&s->part, sizeof(double) * RED_BLOCKS));
    SPK_TRY(dalloc(&s->res, sizeof(double) * 16));
    SPK_TRY(dalloc(&s->spec, sizeof(double) * (size_t)(s->nbins + 1)));
    SPK_TRY(twiddles_create(ctx, s->g, bulk, &s->twb));
    SPK_TRY(twiddles_create(ctx, s->g, SPK_F64, &s->tw64));
    SPK_HIP(hipMemsetAsync(s->U, 0, f64 * s->nf, ctx->stream));
    SPK_HIP(hipMemsetAsync(s->S, 0, f64 * s->nf, ctx->stream));
    SPK_HIP(hipMemsetAsync(s->dU, 0, f64 * s->nf, ctx->stream));
    SPK_HIP(hipMemsetAsync(s->realbuf, 0, f64 * s->nf, ctx->stream));
    TimeState h{};
    SPK_HIP(hipMemcpyAsync(s->ts, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    SPK_HIP(hipStreamSynchronize(ctx->stream));
    s->step = 0;
    s->tremap = 0.0;
    s->loop_started = false;
    s->products_valid = false;
    s->fused_mode = 1;
    s->plane_on = true;
    s->breal_from_w = 0;
    s->zmid_blocks = 0;
    return tune_fused(s);
}

int reduce_final(spk_solver s, int nparts, double scale, double *out) {
    k_final_sum<<<1, 1024, 0, s->ctx->stream>>>(s->part, nparts, scale, s->res);
    SPK_LAUNCH_CHECK();
    SPK_HIP(hipMemcpyAsync(out, s->res, sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
    SPK_HIP(hipStreamSynchronize(s->ctx->stream));
    return SPK_OK;
}

unsigned red_grid(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < RED_BLOCKS ? (b > 0 ? b : 1) : RED_BLOCKS);
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================

// Every entry point taking a solver checks it and, since it may change the
// state or the transform buffers, drops the continuation of the last fused
// step (the inverse x pass it left in W); only spk_main_loop_resume keeps it.
// entry points that read the state and write neither it nor W (times, queries,
// diagnostics into their own scratch and the real-space buffer)
#define SPK_CHECK_SOLVER_RO(s) SPK_REQUIRE((s) != nullptr, SPK_ERR_ARG, "%s: null solver", __func__)
#define SPK_CHECK_SOLVER(s)                                                              \
    do {                                                                                 \
        SPK_REQUIRE((s) != nullptr, SPK_ERR_ARG, "%s: null solver", __func__);           \
        (s)->fused_w_ok = false;                                                         \
    } while (0)

extern "C" int spk_solver_create(spk_ctx ctx, const spk_params *p, int bulk, spk_solver *out) {
    SPK_REQUIRE(ctx && p && out, SPK_ERR_ARG, "spk_solver_create: null argument");
    SPK_HIP(hipSetDevice(ctx->device));
    spk_solver s = new spk_solver_s();
    memset((void *)s, 0, sizeof(*s));
    int r = create(ctx, p, bulk, s);
    if (r != SPK_OK) {
        free_all(s);
        delete s;
        return r;
    }
    ctx_retain(ctx);
    *out = s;
    return SPK_OK;
}

extern "C" int spk_solver_destroy(spk_solver s) {
    if (!s) return SPK_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    free_all(s);
    spk_ctx c = s->ctx;
    delete s;
    ctx_release(c);
    return SPK_OK;
}

extern "C" int spk_solver_info(spk_solver s, int *vars, size_t *ntc) {
    SPK_CHECK_SOLVER_RO(s);
    if (vars) {
        const int v[10] = {s->VX, s->VY, s->VZ, s->BX, s->BY, s->BZ, s->TH, s->VEL, s->MAG, s->nf};
        memcpy(vars, v, sizeof(v));
    }
    if (ntc) *ntc = (size_t)s->g.ntc;
    return SPK_OK;
}

extern "C" int spk_solver_buffers(spk_solver s, void **fields, void **dfields, void **realbuf) {
    SPK_CHECK_SOLVER(s);
    if (fields) *fields = s->U;
    if (dfields) *dfields = s->dU;
    if (realbuf) *realbuf = s->realbuf;
    return SPK_OK;
}

extern "C" int spk_solver_get_time(spk_solver s, double *out8) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(out8, SPK_ERR_ARG, "spk_solver_get_time: null out");
    TimeState h;
    SPK_TRY(read_time(s, &h));
    const double o[8] = {h.t, h.dt, h.tremap, h.dt_hyp, h.dt_par, h.dt_sts, h.t_lastsnap, h.step};
    memcpy(out8, o, sizeof(o));
    return SPK_OK;
}

extern "C" int spk_solver_set_time(spk_solver s, double t, int step, double t_lastsnap) {
    SPK_CHECK_SOLVER(s);
    s->step = step;
    k_set_time<<<1, 1, 0, s->ctx->stream>>>(s->ts, t, (double)step, t_lastsnap, s->tremap);
    SPK_LAUNCH_CHECK();
    s->products_valid = false;
    return SPK_OK;
}

extern "C" int spk_clean_divergence(spk_solver s) {
    SPK_CHECK_SOLVER(s);
    s->products_valid = false;
    return clean_divergence(s);
}

extern "C" int spk_complex2real(spk_solver s) {
    SPK_CHECK_SOLVER_RO(s);
    const Grid3 &g = s->g;
    PassArgs a{s->U, s->realbuf, s->nf, g.ntc, g.ntc, 1.0 / ((double)g.nx * g.ny * g.nz)};
    SPK_TRY(launch_pass_x(s->ctx, g, s->tw64, +1, SPK_F64, SPK_F64, a));
    PassArgs b{s->realbuf, s->realbuf, s->nf, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->tw64, +1, SPK_F64, SPK_F64, b));
    return launch_pass_z_c2r(s->ctx, g, s->tw64, SPK_F64, SPK_F64, b);
}

extern "C" int spk_compute_dt(spk_solver s) {
    SPK_CHECK_SOLVER(s);
    return compute_dt(s);
}

extern "C" int spk_compute_dfield(spk_solver s) {
    SPK_CHECK_SOLVER(s);
    const double ks = s->tremap * s->p.shear;
    SPK_TRY(rhs_front(s, ks, false));
    s->products_valid = false;
    return rhs_back(s, ks, 0, 0.0, 0.0);
}

extern "C" int spk_rk3(spk_solver s) {
    SPK_CHECK_SOLVER(s);
    return rk3(s);
}

extern "C" int spk_rkl2_cycle(spk_solver s) {
    SPK_CHECK_SOLVER(s);
    return rkl2_cycle(s);
}

extern "C" int spk_hydro_mhd_advance(spk_solver s) {
    SPK_CHECK_SOLVER(s);
    return hydro_mhd_advance(s);
}

namespace {
// resume: the first step may start from the inverse x pass the previous fused
// step left in W (spk_main_loop_resume); otherwise every call starts from a fresh
// projection and inverse x pass
int main_loop(spk_solver s, int nsteps, bool resume) {
    if (!s->loop_started) {
        // supervisor.cu:171-175
        if (s->p.shearing) {
            TimeState h;
            SPK_TRY(read_time(s, &h));
            SPK_TRY(shift_time_host(s, h.t));
            SPK_TRY(clock(s, 0, 0, 0, 1, s->tremap));
        }
        s->loop_started = true;
    }
    const bool fused = fused_ok(s);
    for (int n = 0; n < nsteps; n++) {
        if (fused) {
            // forcing changes the state after the fused back end made the next step's
            // inverse x pass, so that pass is made again from the state
            SPK_TRY(step_fused(s, (n > 0 || resume) && s->force_n == 0 && s->fused_w_ok));
        } else {
            SPK_TRY(clean_divergence(s));
            s->products_valid = false;
            SPK_TRY(hydro_mhd_advance(s));
        }
        if (s->force_n > 0) {
            SPK_REQUIRE(!s->ctx->rows, SPK_ERR_STATE, "spk_main_loop: forcing is not batched");
            k_force<<<(s->force_n + 255) / 256, 256, 0, s->ctx->stream>>>(s->U, s->g.ntc, s->VX, s->force_idx,
                                                                          s->force_val, s->force_n, s->ts);
            SPK_LAUNCH_CHECK();
        }
        SPK_TRY(clock(s, 0, 0, 1, 0, 0.0));
    }
    return SPK_OK;
}
}  // namespace

extern "C" int spk_main_loop(spk_solver s, int nsteps) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(nsteps >= 0, SPK_ERR_ARG, "spk_main_loop: negative step count");
    return main_loop(s, nsteps, false);
}

extern "C" int spk_main_loop_resume(spk_solver s, int nsteps) {
    SPK_REQUIRE(s != nullptr, SPK_ERR_ARG, "spk_main_loop_resume: null solver");
    SPK_REQUIRE(nsteps >= 0, SPK_ERR_ARG, "spk_main_loop_resume: negative step count");
    return main_loop(s, nsteps, true);
}

extern "C" int spk_solver_set_forcing(spk_solver s, int nmodes, const long long *index, const double *values) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(nmodes >= 0 && (nmodes == 0 || (index && values)), SPK_ERR_ARG, "spk_solver_set_forcing: bad arguments");
    SPK_REQUIRE(nmodes == 0 || (s->p.incompressible && !s->p.heat_equation), SPK_ERR_ARG,
                "spk_solver_set_forcing: forcing acts on the velocity of an incompressible run");
    for (int m = 0; m < nmodes; m++)
        SPK_REQUIRE(index[m] >= 0 && index[m] < s->g.ntc, SPK_ERR_ARG, "spk_solver_set_forcing: mode index %lld",
                    index[m]);
    SPK_HIP(hipStreamSynchronize(s->ctx->stream));
    if (s->force_idx) (void)hipFree(s->force_idx);
    if (s->force_val) (void)hipFree(s->force_val);
    s->force_idx = nullptr;
    s->force_val = nullptr;
    s->force_n = 0;
    if (nmodes == 0) return SPK_OK;
    SPK_TRY(dalloc(&s->force_idx, sizeof(long long) * (size_t)nmodes));
    SPK_TRY(dalloc(&s->force_val, sizeof(cpx<double>) * 3 * (size_t)nmodes));
    SPK_HIP(hipMemcpyAsync(s->force_idx, index, sizeof(long long) * (size_t)nmodes, hipMemcpyHostToDevice,
                           s->ctx->stream));
    SPK_HIP(hipMemcpyAsync(s->force_val, values, sizeof(double) * 6 * (size_t)nmodes, hipMemcpyHostToDevice,
                           s->ctx->stream));
    SPK_HIP(hipStreamSynchronize(s->ctx->stream));
    s->force_n = nmodes;
    return SPK_OK;
}

extern "C" int spk_energy(spk_solver s, int field, double *out) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(field >= 0 && field < s->nf && out, SPK_ERR_ARG, "spk_energy: bad field %d", field);
    const unsigned nb = red_grid(s->g.ntc);
    k_spec_sum<<<nb, 256, 0, s->ctx->stream>>>(s->U + (long long)field * s->g.ntc, s->g.ntc, s->kg0, 0, s->part);
    SPK_LAUNCH_CHECK();
    const double nt = (double)s->g.nx * s->g.ny * s->g.nz;
    return reduce_final(s, (int)nb, 1.0 / (nt * nt), out);
}

extern "C" int spk_enstrophy(spk_solver s, int f0, double *out) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(f0 >= 0 && f0 + 2 < s->nf && out, SPK_ERR_ARG, "spk_enstrophy: bad field %d", f0);
    const unsigned nb = red_grid(s->g.ntc);
    k_spec_sum<<<nb, 256, 0, s->ctx->stream>>>(s->U + (long long)f0 * s->g.ntc, s->g.ntc,
                                               kgeom(s, s->tremap * s->p.shear), 1, s->part);
    SPK_LAUNCH_CHECK();
    const double nt = (double)s->g.nx * s->g.ny * s->g.nz;
    return reduce_final(s, (int)nb, 1.0 / (nt * nt), out);
}

extern "C" int spk_correlation(spk_solver s, int fa, int fb, double *out) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(fa >= 0 && fa < s->nf && fb >= 0 && fb < s->nf && out, SPK_ERR_ARG, "spk_correlation: bad fields");
    const long long rows = (long long)s->g.nx * s->g.ny;
    const unsigned nb = red_grid(rows * s->g.nz);
    const long long fs = 2 * s->g.ntc;
    k_real_dot<<<nb, 256, 0, s->ctx->stream>>>(s->realbuf + fa * fs, s->realbuf + fb * fs, rows, s->g.nz, 2 * s->g.nzh,
                                               s->part);
    SPK_LAUNCH_CHECK();
    return reduce_final(s, (int)nb, 1.0 / ((double)s->g.nx * s->g.ny * s->g.nz), out);
}

namespace {

// computeSpectrum1d (spooky_output_spectrum.cu:424-444) of two spectral fields
int spectrum_of(spk_solver s, const cpx<double> *a, const cpx<double> *b, double *out, int cap) {
    const int nb = s->nbins;
    SPK_REQUIRE(nb <= 16384, SPK_ERR_ARG, "spectrum: %d bins exceed the LDS histogram", nb);
    k_zero_d<<<1, 256, 0, s->ctx->stream>>>(s->spec, nb + 1);
    const unsigned g = red_grid(s->g.ntc);
    k_spectrum<<<g, 256, sizeof(double) * (size_t)(nb + 1), s->ctx->stream>>>(
        a, b, s->g.ntc, kgeom(s, s->tremap * s->p.shear), (double)s->g.nx * s->g.ny * s->g.nz, s->deltak, nb, s->spec);
    SPK_LAUNCH_CHECK();
    std::vector<double> h((size_t)nb + 1);
    SPK_HIP(hipMemcpyAsync(h.data(), s->spec, sizeof(double) * (size_t)(nb + 1), hipMemcpyDeviceToHost, s->ctx->stream));
    SPK_HIP(hipStreamSynchronize(s->ctx->stream));
    for (int i = 0; i < nb && i < cap; i++) out[i] = h[i];
    return SPK_OK;
}

}  // namespace

extern "C" int spk_spectrum(spk_solver s, int fa, int fb, double *out, int cap, int *nbins) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(fa >= 0 && fa < s->nf && fb >= 0 && fb < s->nf && out && cap >= 0, SPK_ERR_ARG,
                "spk_spectrum: bad arguments");
    if (nbins) *nbins = s->nbins;
    return spectrum_of(s, s->U + (long long)fa * s->g.ntc, s->U + (long long)fb * s->g.ntc, out, cap);
}

extern "C" int spk_spectrum_bins(spk_solver s, int *nbins, double *deltak) {
    SPK_CHECK_SOLVER_RO(s);
    if (nbins) *nbins = s->nbins;
    if (deltak) *deltak = s->deltak;
    return SPK_OK;
}

namespace {

// FP64 3D inverse transform of n complex fields (1/N folded in) into padded real rows
int diag_c2r(spk_solver s, const void *in, void *out, int n) {
    const Grid3 &g = s->g;
    PassArgs a{in, out, n, g.ntc, g.ntc, 1.0 / ((double)g.nx * g.ny * g.nz)};
    SPK_TRY(launch_pass_x(s->ctx, g, s->tw64, +1, SPK_F64, SPK_F64, a));
    PassArgs b{out, out, n, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->tw64, +1, SPK_F64, SPK_F64, b));
    return launch_pass_z_c2r(s->ctx, g, s->tw64, SPK_F64, SPK_F64, b);
}

// FP64 3D forward transform of n padded real fields into complex fields
int diag_r2c(spk_solver s, const void *in, void *out, int n) {
    const Grid3 &g = s->g;
    PassArgs a{in, out, n, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_z_r2c(s->ctx, g, s->tw64, SPK_F64, SPK_F64, a));
    PassArgs b{out, out, n, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->tw64, -1, SPK_F64, SPK_F64, b));
    return launch_pass_x(s->ctx, g, s->tw64, -1, SPK_F64, SPK_F64, b);
}

// twoFieldCorrelation of padded real fields: sum over physical points / N, added to *acc
int diag_dot(spk_solver s, const double *a, const double *b, double *acc) {
    const long long rows = (long long)s->g.nx * s->g.ny;
    const unsigned nb = red_grid(rows * s->g.nz);
    k_real_dot<<<nb, 256, 0, s->ctx->stream>>>(a, b, rows, s->g.nz, 2 * s->g.nzh, s->part);
    SPK_LAUNCH_CHECK();
    double v;
    SPK_TRY(reduce_final(s, (int)nb, 1.0 / ((double)s->g.nx * s->g.ny * s->g.nz), &v));
    *acc += v;
    return SPK_OK;
}

// 6 FP64 fields of scratch, allocated at the first diagnostic that needs them
int need_diag(spk_solver s) {
    if (!s->diag) {
        hipError_t e = hipMalloc((void **)&s->diag, (size_t)s->g.ntc * 16 * 6);
        if (e != hipSuccess) {
            s->diag = nullptr;
            set_error("hipMalloc(diagnostic scratch): %s", hipGetErrorString(e));
            return SPK_ERR_ALLOC;
        }
    }
    return SPK_OK;
}

int diag_spec_sum(spk_solver s, const cpx<double> *u, int kind, const KGeom &kg, double *out) {
    const unsigned nb = red_grid(s->g.ntc);
    k_spec_sum<<<nb, 256, 0, s->ctx->stream>>>(u, s->g.ntc, kg, kind, s->part);
    SPK_LAUNCH_CHECK();
    const double nt = (double)s->g.nx * s->g.ny * s->g.nz;
    return reduce_final(s, (int)nb, 1.0 / (nt * nt), out);
}

}  // namespace

// WriteTimevarOutput (spooky_output_timevar.cu:48-204), one quantity by name
extern "C" int spk_timevar(spk_solver s, const char *name, double *out) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(name && out, SPK_ERR_ARG, "spk_timevar: null argument");
    const std::string v(name);
    const spk_params &p = s->p;
    const Grid3 &g = s->g;
    const long long fr = 2 * g.ntc;  // doubles per padded real field
    const KGeom kg = kgeom(s, s->tremap * p.shear);
    const double *R = s->realbuf;
    const int VX = s->VX, BX = s->BX, TH = s->TH;
    auto energy = [&](int f, double *o) { return diag_spec_sum(s, s->U + (long long)f * g.ntc, 0, kg, o); };
    auto corr = [&](int a, int b, double *o) {
        *o = 0.0;
        return diag_dot(s, R + a * fr, R + b * fr, o);
    };
    const unsigned gb = red_grid(g.ntc);
    hipStream_t st = s->ctx->stream;
    double a = 0.0, b = 0.0, c = 0.0;
    const bool inc = p.incompressible, mhd = p.mhd, th = p.boussinesq || p.heat_equation;
    if (inc && (v == "ev" || v == "Kx" || v == "Ky" || v == "Kz")) {
        if (v != "ev") return energy(VX + (v[1] - 'x'), out);
        SPK_TRY(energy(VX, &a)); SPK_TRY(energy(VX + 1, &b)); SPK_TRY(energy(VX + 2, &c));
        *out = a + b + c;
        return SPK_OK;
    }
    if (mhd && (v == "em" || v == "Mx" || v == "My" || v == "Mz")) {
        if (v != "em") return energy(BX + (v[1] - 'x'), out);
        SPK_TRY(energy(BX, &a)); SPK_TRY(energy(BX + 1, &b)); SPK_TRY(energy(BX + 2, &c));
        *out = a + b + c;
        return SPK_OK;
    }
    if (th && v == "et") return energy(TH, out);
    if (inc && v == "vxvy") return corr(VX, VX + 1, out);
    if (inc && v == "vxvz") return corr(VX, VX + 2, out);
    if (inc && v == "vyvz") return corr(VX + 1, VX + 2, out);
    if (mhd && v == "bxby") return corr(BX, BX + 1, out);
    if (mhd && v == "bxbz") return corr(BX, BX + 2, out);
    if (mhd && v == "bybz") return corr(BX + 1, BX + 2, out);
    if (p.boussinesq && v == "thvx") return corr(TH, VX, out);
    if (p.boussinesq && v == "thvz") return corr(TH, VX + 2, out);
    if (inc && v == "w2") return diag_spec_sum(s, s->U + (long long)VX * g.ntc, 1, kg, out);
    if (mhd && v == "j2") return diag_spec_sum(s, s->U + (long long)BX * g.ntc, 1, kg, out);
    if (mhd && v == "hc") {
        *out = 0.0;
        for (int q = 0; q < 3; q++) SPK_TRY(diag_dot(s, R + (VX + q) * fr, R + (BX + q) * fr, out));
        return SPK_OK;
    }
    if ((inc && v == "hv") || (mhd && v == "hm")) {
        // computeHelicity (spooky_output_functions.cu:88-107)
        const int f0 = v == "hv" ? VX : BX;
        SPK_TRY(need_diag(s));
        k_diag_vec<<<gb, 256, 0, st>>>(s->U + (long long)f0 * g.ntc, s->diag, g.ntc, kg, 0);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_c2r(s, s->diag, s->diag, 3));
        const double *h = (const double *)s->diag;
        *out = 0.0;
        for (int q = 0; q < 3; q++) SPK_TRY(diag_dot(s, h + q * fr, R + (f0 + q) * fr, out));
        return SPK_OK;
    }
    if (mhd && (v == "thetaB" || v == "bz2" || v == "phiB")) {
        // averagebz / averagebz2 / averagephiB (spooky_output_functions.cu:340-395)
        const int kind = v == "thetaB" ? 0 : (v == "bz2" ? 1 : 2);
        const long long rows = (long long)g.nx * g.ny;
        const unsigned nb = red_grid(rows * g.nz);
        k_diag_bsum<<<nb, 256, 0, st>>>(R + BX * fr, fr, rows, g.nz, 2 * g.nzh, kind, s->part);
        SPK_LAUNCH_CHECK();
        return reduce_final(s, (int)nb, 1.0 / ((double)g.nx * g.ny * g.nz), out);
    }
    if (p.boussinesq && inc && v == "pv") {
        // potentialVorticity (spooky_output_functions.cu:109-140)
        SPK_TRY(need_diag(s));
        k_diag_vec<<<gb, 256, 0, st>>>(s->U + (long long)VX * g.ntc, s->diag, g.ntc, kg, 1);
        k_diag_vec<<<gb, 256, 0, st>>>(s->U + (long long)TH * g.ntc, s->diag + 3 * g.ntc, g.ntc, kg, 2);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_c2r(s, s->diag, s->diag, 6));
        const double *h = (const double *)s->diag;
        *out = 0.0;
        for (int q = 0; q < 3; q++) SPK_TRY(diag_dot(s, h + q * fr, h + (3 + q) * fr, out));
        return SPK_OK;
    }
    if (th && v == "dissgradT") {
        if (!(p.anisotropic_diffusion && mhd))  // computeDissipation (spooky_output_functions.cu:142-160)
            return diag_spec_sum(s, s->U + (long long)TH * g.ntc, 2, kg, out);
        // - computeAnisoDissipation (spooky_output_functions.cu:228-252, anisotropic_conduction.cu:107-156)
        SPK_TRY(need_diag(s));
        k_diag_vec<<<gb, 256, 0, st>>>(s->U + (long long)TH * g.ntc, s->diag, g.ntc, kg, 2);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_c2r(s, s->diag, s->diag, 3));
        double *h = (double *)s->diag;
        k_diag_bvec<<<red_grid(fr), 256, 0, st>>>(R + BX * fr, h, h + 3 * fr, fr, p.strat_direction, 1);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, h + 3 * fr, s->diag + 3 * g.ntc, 3));
        k_diag_div<<<gb, 256, 0, st>>>(s->diag + 3 * g.ntc, s->diag, g.ntc, kg, 0);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_c2r(s, s->diag, s->diag, 1));
        double d = 0.0;
        SPK_TRY(diag_dot(s, h, R + TH * fr, &d));
        *out = -d;
        return SPK_OK;
    }
    if (p.anisotropic_diffusion && p.boussinesq && v == "fluxbbgradT") {
        // computeAnisoInjection (spooky_output_functions.cu:254-276, anisotropic_conduction.cu:69-104)
        SPK_TRY(need_diag(s));
        double *h = (double *)s->diag;
        k_diag_bvec<<<red_grid(fr), 256, 0, st>>>(R + BX * fr, nullptr, h, fr, p.strat_direction, 0);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, h, s->diag, 3));
        k_diag_div<<<gb, 256, 0, st>>>(s->diag, s->diag + 3 * g.ntc, g.ntc, kg, 0);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_c2r(s, s->diag + 3 * g.ntc, s->diag + 3 * g.ntc, 1));
        *out = 0.0;
        return diag_dot(s, h + 3 * fr, R + TH * fr, out);
    }
    set_error("spk_timevar: '%s' is not a timevar quantity of this configuration", name);
    return SPK_ERR_ARG;
}

// WriteSpectrumOutput (spooky_output_spectrum.cu:22-400), one spectrum by name
extern "C" int spk_spectrum_named(spk_solver s, const char *name, double *out, int cap, int *nbins) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(name && out && cap >= 0, SPK_ERR_ARG, "spk_spectrum_named: bad arguments");
    if (nbins) *nbins = s->nbins;
    const std::string v(name);
    const spk_params &p = s->p;
    const Grid3 &g = s->g;
    const long long fr = 2 * g.ntc;  // doubles per padded real field
    const KGeom kg = kgeom(s, s->tremap * p.shear);
    const double *R = s->realbuf;
    const int VX = s->VX, BX = s->BX, TH = s->TH;
    const bool inc = p.incompressible, mhd = p.mhd, bq = p.boussinesq;
    auto F = [&](int f) -> const cpx<double> * { return s->U + (long long)f * g.ntc; };
    // shell spectra of two state fields; bxbz pairs B_x with B_y as the reference
    // does (spooky_output_spectrum.cu:153-158)
    struct Pair {
        const char *n;
        bool on;
        int a, b;
    };
    const Pair pairs[] = {
        {"Kx", inc, VX, VX},         {"Ky", inc, VX + 1, VX + 1},     {"Kz", inc, VX + 2, VX + 2},
        {"Mx", mhd, BX, BX},         {"My", mhd, BX + 1, BX + 1},     {"Mz", mhd, BX + 2, BX + 2},
        {"Eth", bq || p.heat_equation, TH, TH},
        {"vxvy", inc, VX, VX + 1},   {"vyvz", inc, VX + 1, VX + 2},   {"vxvz", inc, VX, VX + 2},
        {"bxby", mhd, BX, BX + 1},   {"bybz", mhd, BX + 1, BX + 2},   {"bxbz", mhd, BX, BX + 1},
        {"thvx", bq, TH, VX},        {"thvy", bq, TH, VX + 1},        {"thvz", bq, TH, VX + 2},
    };
    for (const Pair &q : pairs)
        if (q.on && v == q.n) return spectrum_of(s, F(q.a), F(q.b), out, cap);
    // transfer spectra: a state field against a transformed nonlinear term
    auto comp = [&](const char *base) -> int {
        const size_t nb = strlen(base);
        if (v.size() == nb + 1 && v.compare(0, nb, base) == 0 && v[nb] >= 'x' && v[nb] <= 'z') return v[nb] - 'x';
        return -1;
    };
    hipStream_t st = s->ctx->stream;
    const unsigned gb = red_grid(g.ntc), gr = red_grid(fr);
    int c;
    if ((inc && (c = comp("advec_")) >= 0) || (inc && mhd && (c = comp("lorentz_")) >= 0)) {
        // NonLinearAdvection (hydro.cu:78-105) of real u (advec) or real B (lorentz),
        // against u
        SPK_TRY(need_diag(s));
        double *Dr = (double *)s->diag;
        k_diag_prod<<<gr, 256, 0, st>>>(R + (v[0] == 'a' ? VX : BX) * fr, nullptr, Dr, fr, 0);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, Dr, s->diag, 6));
        k_diag_sop<<<gb, 256, 0, st>>>(s->diag, s->diag, g.ntc, kg, 0);
        SPK_LAUNCH_CHECK();
        return spectrum_of(s, F(VX + c), s->diag + c * g.ntc, out, cap);
    }
    if (inc && mhd && (c = comp("emfpower_")) >= 0) {
        // CurlEMF (mhd.cu:20-62) against B
        SPK_TRY(need_diag(s));
        double *Dr = (double *)s->diag;
        k_diag_prod<<<gr, 256, 0, st>>>(R + VX * fr, R + BX * fr, Dr, fr, 1);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, Dr, s->diag, 3));
        k_diag_sop<<<gb, 256, 0, st>>>(s->diag, s->diag, g.ntc, kg, 1);
        SPK_LAUNCH_CHECK();
        return spectrum_of(s, F(BX + c), s->diag + c * g.ntc, out, cap);
    }
    if (bq && inc && v == "th_advec") {
        // AdvectTemperature (boussinesq.cu:10-42) against theta
        SPK_TRY(need_diag(s));
        double *Dr = (double *)s->diag;
        k_diag_prod<<<gr, 256, 0, st>>>(R + VX * fr, R + TH * fr, Dr, fr, 2);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, Dr, s->diag, 3));
        k_diag_sop<<<gb, 256, 0, st>>>(s->diag, s->diag, g.ntc, kg, 2);
        SPK_LAUNCH_CHECK();
        return spectrum_of(s, F(TH), s->diag, out, cap);
    }
    if (mhd && (c = comp("helicity_")) >= 0) {
        // MagneticHelicity (mhd.cu:65-76) against B
        SPK_TRY(need_diag(s));
        k_diag_vec<<<gb, 256, 0, st>>>(F(BX), s->diag, g.ntc, kg, 0);
        SPK_LAUNCH_CHECK();
        return spectrum_of(s, F(BX + c), s->diag + c * g.ntc, out, cap);
    }
    if (p.anisotropic_diffusion && v == "th_aniso_diss") {
        // AnisotropicDissipation (anisotropic_conduction.cu:107-156). Its last step
        // (DivergenceMask, flag 1) adds to a scratch array that, at this point of
        // WriteSpectrumOutput, still holds the unmasked transform of B_y B_z left
        // there by the lorentz term (hydro.cu:95-101 on real B); the sum is mirrored.
        SPK_TRY(need_diag(s));
        double *Dr = (double *)s->diag;
        k_diag_vec<<<gb, 256, 0, st>>>(F(TH), s->diag, g.ntc, kg, 2);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_c2r(s, s->diag, s->diag, 3));
        k_diag_bvec<<<gr, 256, 0, st>>>(R + BX * fr, Dr, Dr + 3 * fr, fr, p.strat_direction, 1);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, Dr + 3 * fr, s->diag + 3 * g.ntc, 3));
        k_diag_prod<<<gr, 256, 0, st>>>(R + BX * fr, nullptr, Dr, fr, 3);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, Dr, s->diag, 1));
        k_diag_div<<<gb, 256, 0, st>>>(s->diag + 3 * g.ntc, s->diag, g.ntc, kg, 1);
        SPK_LAUNCH_CHECK();
        return spectrum_of(s, F(TH), s->diag, out, cap);
    }
    if (p.anisotropic_diffusion && v == "th_aniso_inj") {
        // AnisotropicInjection (anisotropic_conduction.cu:69-104)
        SPK_TRY(need_diag(s));
        double *Dr = (double *)s->diag;
        k_diag_bvec<<<gr, 256, 0, st>>>(R + BX * fr, nullptr, Dr, fr, p.strat_direction, 0);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_r2c(s, Dr, s->diag, 3));
        k_diag_div<<<gb, 256, 0, st>>>(s->diag, s->diag + 3 * g.ntc, g.ntc, kg, 0);
        SPK_LAUNCH_CHECK();
        return spectrum_of(s, F(TH), s->diag + 3 * g.ntc, out, cap);
    }
    set_error("spk_spectrum_named: '%s' is not a spectrum of this configuration", name);
    return SPK_ERR_ARG;
}

// WriteUserTimevarOutput (user_output_timevar.cu:18-76): the user quantities the
// reference ships
extern "C" int spk_user_timevar(spk_solver s, const char *name, double *out) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(name && out, SPK_ERR_ARG, "spk_user_timevar: null argument");
    const std::string v(name);
    const spk_params &p = s->p;
    const Grid3 &g = s->g;
    if (p.anisotropic_diffusion && v == "kpartheta") {
        // computekpartheta (user_output_functions.cu:31-58): <(b . grad theta)^2> / (2 E_theta)
        const long long fr = 2 * g.ntc;
        const KGeom kg = kgeom(s, s->tremap * p.shear);
        hipStream_t st = s->ctx->stream;
        SPK_TRY(need_diag(s));
        double *Dr = (double *)s->diag;
        k_diag_vec<<<red_grid(g.ntc), 256, 0, st>>>(s->U + (long long)s->TH * g.ntc, s->diag, g.ntc, kg, 2);
        SPK_LAUNCH_CHECK();
        SPK_TRY(diag_c2r(s, s->diag, s->diag, 3));
        k_diag_prod<<<red_grid(fr), 256, 0, st>>>(s->realbuf + s->BX * fr, Dr, Dr + 3 * fr, fr, 4);
        SPK_LAUNCH_CHECK();
        double c2 = 0.0, e = 0.0;
        SPK_TRY(diag_dot(s, Dr + 3 * fr, Dr + 3 * fr, &c2));
        SPK_TRY(diag_spec_sum(s, s->U + (long long)s->TH * g.ntc, 0, kg, &e));
        *out = c2 / (2.0 * e);
        return SPK_OK;
    }
    // customFunction returns 0 (user_output_functions.cu:60-84); uservar2 is 0
    if (v == "uservar1" || v == "uservar2") {
        *out = 0.0;
        return SPK_OK;
    }
    set_error("spk_user_timevar: '%s' is not a user quantity of this configuration", name);
    return SPK_ERR_ARG;
}

// the real fields a snapshot holds (check_output.cu:70-84): the inverse transform
// of the state, unsheared first in a shearing box (UnshearOutput,
// unshear_output.cu:21-54: phase exp(-i ky c / 2), c = Lx S fmod(t, 2 Ly / (S Lx)))
extern "C" int spk_snapshot_real(spk_solver s) {
    SPK_CHECK_SOLVER_RO(s);
    const spk_params &p = s->p;
    if (!p.shearing) return spk_complex2real(s);
    const Grid3 &g = s->g;
    TimeState h;
    SPK_TRY(read_time(s, &h));
    const double tvel = fmod(h.t, 2.0 * p.ly / (p.shear * p.lx));
    const long long n = (long long)s->nf * g.ntc;
    k_unshear<<<red_grid(n), 256, 0, s->ctx->stream>>>(s->U, reinterpret_cast<cpx<double> *>(s->realbuf), g.ntc, n,
                                                       s->kg0, tvel * p.lx * p.shear);
    SPK_LAUNCH_CHECK();
    s->products_valid = false;
    return diag_c2r(s, s->realbuf, s->realbuf, s->nf);
}

extern "C" int spk_kernel_timings(spk_solver s, int reps, double *out_ms, int cap, int *count) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(reps >= 1 && out_ms && cap >= 0, SPK_ERR_ARG, "spk_kernel_timings: bad arguments");
    constexpr int NK = 9;
    double ms[NK];
    for (int k = 0; k < NK; k++) ms[k] = -1.0;
    if (count) *count = NK;
    const Grid3 &g = s->g;
    hipStream_t st = s->ctx->stream;
    const size_t fbytes = (size_t)g.ntc * 16 * s->nf;
    // keep the state: the epilogue and the projection write it
    SPK_HIP(hipMemcpyAsync(s->realbuf, s->U, fbytes, hipMemcpyDeviceToDevice, st));
    hipEvent_t e0, e1;
    SPK_HIP(hipEventCreate(&e0));
    SPK_HIP(hipEventCreate(&e1));
    const double ks = s->tremap * s->p.shear;
    const bool inc = s->nin > 0;
    auto timed = [&](int k, auto &&launch) -> int {
        SPK_TRY(launch());  // warm
        SPK_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) SPK_TRY(launch());
        SPK_HIP(hipEventRecord(e1, st));
        SPK_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        SPK_HIP(hipEventElapsedTime(&t, e0, e1));
        ms[k] = (double)t / reps;
        return SPK_OK;
    };
    int rc = SPK_OK;
    do {
        if (inc) {
            rc = timed(0, [&] {
                return s->bulk == SPK_F64 ? launch_inv_x<double>(s, s->inv, s->nin, s->W, ks)
                                          : launch_inv_x<float>(s, s->inv, s->nin, s->W, ks);
            });
            if (rc) break;
            PassArgs aw{s->W, s->W, s->nin, g.ntc, g.ntc, 1.0};
            rc = timed(1, [&] { return launch_pass_y(s->ctx, g, s->twb, +1, s->bulk, s->bulk, aw); });
            if (rc) break;
            unsigned nb = 0;
            rc = timed(2, [&] {
                return s->bulk == SPK_F64 ? launch_zmid<double>(s, s->zprog, s->W, s->P, s->pmax, &nb)
                                          : launch_zmid<float>(s, s->zprog, s->W, s->P, s->pmax, &nb);
            });
            if (rc) break;
            s->zmid_blocks = (int)nb;
            PassArgs ap{s->P, s->P, s->nout, g.ntc, g.ntc, 1.0};
            rc = timed(3, [&] { return launch_pass_y(s->ctx, g, s->twb, -1, s->bulk, s->bulk, ap); });
            if (rc) break;
            rc = timed(4, [&] { return launch_pass_x(s->ctx, g, s->twb, -1, s->bulk, s->bulk, ap); });
            if (rc) break;
        }
        EpiConst c;
        memset(&c, 0, sizeof(c));
        c.kg = kgeom(s, ks);
        c.ntc = g.ntc;
        c.stratification = s->p.stratification;
        c.sd = s->p.strat_direction;
        c.shearing = s->p.shearing;
        c.rot_active = s->p.rotating && s->p.shearing;
        c.aniso_rhs = s->aniso_rhs;
        c.diff_mask = s->diff_mask;
        c.nu = s->p.nu; c.nu_m = s->p.nu_m; c.nu_th = s->p.nu_th; c.N2 = s->p.N2;
        c.shear = s->p.shear; c.omega = s->p.omega;
        c.oT = s->oT; c.oE = s->oE; c.oQ = s->oQ; c.oA = s->oA;
        c.gam = 0.0;  // U is restored from the copy after timing
        c.xi = 0.0;
        rc = timed(5, [&] {
            return s->bulk == SPK_F64 ? launch_epilogue<double>(s, c, 2) : launch_epilogue<float>(s, c, 2);
        });
        if (rc) break;
        if (inc) {
            rc = timed(6, [&] { return clean_divergence(s); });
            if (rc) break;
        }
        rc = timed(7, [&] { return dt_kernel(s); });
        if (rc) break;
        if (fused_allowed(s)) {
            FusedConst fc;
            memset(&fc, 0, sizeof(fc));
            fc.e = c;
            fc.project = 0;
            fc.inv = 1;
            fc.scale = 1.0 / ((double)g.nx * g.ny * g.nz);
            const void *tw = (const char *)s->twb.d + (size_t)s->twb.off_x * (s->bulk == SPK_F64 ? 16 : 8);
            rc = timed(8, [&] {
                return launch_fwdx_epilogue(s->ctx, s->bulk, g.nx, s->ly, 2, s->fused_ef, s->P, s->U, s->S, s->W, fc,
                                            tw, s->ts, fused_groups(s));
            });
        }
    } while (0);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SPK_TRY(rc);
    SPK_HIP(hipMemcpyAsync(s->U, s->realbuf, fbytes, hipMemcpyDeviceToDevice, st));
    SPK_HIP(hipStreamSynchronize(st));
    s->products_valid = false;
    for (int k = 0; k < NK && k < cap; k++) out_ms[k] = ms[k];
    return SPK_OK;
}

namespace spk {
void solver_views(spk_solver s, SolverViews *v) {
    s->fused_w_ok = false;  // the caller may change the state through the views
    v->U = s->U;
    v->realbuf = s->realbuf;
    v->ts = s->ts;
    v->kg = kgeom(s, s->tremap * s->p.shear);
    v->VX = s->VX;
    v->ntc = s->g.ntc;
    v->nx = s->g.nx;
    v->ny = s->g.ny;
    v->nz = s->g.nz;
    v->ctx = s->ctx;
    v->incompressible = s->p.incompressible && !s->p.heat_equation;
}
spk_ctx solver_ctx(spk_solver s) { return s->ctx; }
int solver_host_step(spk_solver s) { return s->step; }
// host-side step count after n steps that ran from a replayed graph
void solver_advance_host(spk_solver s, int n) {
    s->step += n;
    s->products_valid = false;
}

void solver_box_row(spk_solver s, BoxRow *row, size_t *sizes) {
    memset(row, 0, sizeof(*row));
    for (int k = 0; k < BOX_SLOTS; k++) sizes[k] = 0;
    const size_t f64 = (size_t)s->g.ntc * 16, fb = (size_t)s->g.ntc * esize(s->bulk);
    auto put = [&](int k, const void *p, size_t n) {
        row->p[k] = static_cast<char *>(const_cast<void *>(p));
        sizes[k] = p ? n : 0;
    };
    put(SL_U, s->U, f64 * s->nf);
    put(SL_S, s->S, f64 * s->nf);
    put(SL_DU, s->dU, f64 * s->nf);
    put(SL_REAL, s->realbuf, f64 * s->nf);
    put(SL_W, s->W, fb * (s->nin > 0 ? s->nin : 1));
    put(SL_P, s->P, fb * (s->nout > 0 ? s->nout : 1));
    put(SL_W2, s->W2, fb * 6);
    put(SL_P2, s->P2, fb * 3);
    put(SL_BREAL, s->Breal, fb * 3);
    put(SL_TS, s->ts, sizeof(TimeState));
    put(SL_PMAX, s->pmax, sizeof(double) * 6 * (size_t)((long long)s->p.nx * s->p.ny + 1));
    put(SL_UC0, s->rUc0, f64 * s->nsts);
    put(SL_UC1, s->rUc1, f64 * s->nsts);
    put(SL_DU0, s->rdU0, f64 * s->nsts);
    put(SL_PLAN, s->rkl_plan, sizeof(RklPlan));
    row->nu = s->p.nu;
    row->nu_m = s->p.nu_m;
    row->nu_th = s->p.nu_th;
    row->reynolds_ani = s->p.reynolds_ani;
}

bool solver_same_shape(spk_solver a, spk_solver b) {
    spk_params pa = a->p, pb = b->p;
    pa.nu = pb.nu = 0.0;
    pa.nu_m = pb.nu_m = 0.0;
    pa.nu_th = pb.nu_th = 0.0;
    pa.reynolds_ani = pb.reynolds_ani = 0.0;
    const bool ints = pa.nx == pb.nx && pa.ny == pb.ny && pa.nz == pb.nz && pa.incompressible == pb.incompressible &&
                      pa.heat_equation == pb.heat_equation && pa.boussinesq == pb.boussinesq &&
                      pa.stratification == pb.stratification && pa.strat_direction == pb.strat_direction &&
                      pa.mhd == pb.mhd && pa.anisotropic_diffusion == pb.anisotropic_diffusion &&
                      pa.supertimestepping == pb.supertimestepping && pa.shearing == pb.shearing &&
                      pa.rotating == pb.rotating && pa.with2d == pb.with2d && pa.antialiasing == pb.antialiasing &&
                      pa.sts_mask == pb.sts_mask && pa.safety_sts == pb.safety_sts;
    const bool dbls = pa.lx == pb.lx && pa.ly == pb.ly && pa.lz == pb.lz && pa.N2 == pb.N2 &&
                      pa.OmegaT2 == pb.OmegaT2 && pa.omega == pb.omega && pa.shear == pb.shear && pa.cfl == pb.cfl &&
                      pa.cfl_par == pb.cfl_par && pa.cfl_sts == pb.cfl_sts && pa.safety_source == pb.safety_source &&
                      pa.t_initial == pb.t_initial && pa.t_final == pb.t_final && pa.toutput_flow == pb.toutput_flow;
    return ints && dbls && a->bulk == b->bulk && a->rkl_device == b->rkl_device && a->force_n == 0 &&
           b->force_n == 0 && !a->p.shearing && (a->W2 != nullptr) == (b->W2 != nullptr) &&
           (a->Breal != nullptr) == (b->Breal != nullptr) && a->plane_on == b->plane_on;
}

bool solver_w_ok(spk_solver s) { return s->fused_w_ok; }
void solver_set_w_ok(spk_solver s, bool v) { s->fused_w_ok = v; }
bool solver_continues(spk_solver s) {
    const bool sts_uv = s->p.supertimestepping && s->nsts > 0 && !(s->nsts == 1 && s->sts_index[0] == s->TH);
    return fused_ok(s) && s->force_n == 0 && !sts_uv;
}

int solver_follow_back_end(spk_solver s, spk_solver lead) {
    s->fused_mode = lead->fused_mode;
    s->fused_faster = lead->fused_faster;
    s->fused_ef = lead->fused_ef;
    s->fused_ef_auto = lead->fused_ef_auto;
    s->plane_on = lead->plane_on;
    return SPK_OK;
}
}  // namespace spk

extern "C" int spk_solver_set_rkl_device(spk_solver s, int on) {
    SPK_CHECK_SOLVER(s);
    if (!on) {
        s->rkl_device = 0;
        return SPK_OK;
    }
    const spk_params &p = s->p;
    SPK_REQUIRE(s->nsts > 0 && p.incompressible && !p.heat_equation, SPK_ERR_ARG,
                "spk_solver_set_rkl_device: needs super-time-stepping of an incompressible run (the sts cap of "
                "compute_dt bounds the stage count only there)");
    SPK_REQUIRE(s->rkl_smax <= RKL_SMAX, SPK_ERR_ARG,
                "spk_solver_set_rkl_device: safety_sts = %d allows %d stages; a device plan holds %d", p.safety_sts,
                s->rkl_smax, RKL_SMAX);
    s->rkl_device = 1;
    return SPK_OK;
}

extern "C" int spk_solver_rkl_plan(spk_solver s, int *stages, int *overflow, int *smax) {
    SPK_CHECK_SOLVER_RO(s);
    SPK_REQUIRE(s->rkl_plan, SPK_ERR_ARG, "spk_solver_rkl_plan: no super-time-stepping in this run");
    int h[2];
    SPK_HIP(hipMemcpyAsync(h, s->rkl_plan, sizeof(h), hipMemcpyDeviceToHost, s->ctx->stream));
    SPK_HIP(hipStreamSynchronize(s->ctx->stream));
    if (stages) *stages = h[0];
    if (overflow) *overflow = h[1];
    if (smax) *smax = s->rkl_smax;
    return SPK_OK;
}

extern "C" int spk_solver_set_fused(spk_solver s, int mode) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(mode >= 0 && mode <= 2, SPK_ERR_ARG, "spk_solver_set_fused: mode %d", mode);
    s->fused_mode = mode;
    return SPK_OK;
}

extern "C" int spk_solver_set_fused_points(spk_solver s, int points) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(points == 0 || points == 4 || points == 8, SPK_ERR_ARG,
                "spk_solver_set_fused_points: %d (4, 8, or 0 for the autotuned tile)", points);
    s->fused_ef = points == 0 ? s->fused_ef_auto : points;
    return SPK_OK;
}

extern "C" int spk_solver_set_plane(spk_solver s, int on) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(on == 0 || on == 1, SPK_ERR_ARG, "spk_solver_set_plane: %d (0 or 1)", on);
    s->plane_on = on != 0;
    s->products_valid = false;  // P's layout after the front depends on the switch
    return SPK_OK;
}

extern "C" int spk_solver_plane(spk_solver s, int *rk3, int *rkl2) {
    SPK_CHECK_SOLVER_RO(s);
    if (rk3) *rk3 = plane_front(s) ? 1 : 0;
    if (rkl2) *rkl2 = plane_rkl(s) && s->rkl_aniso && rkl_fused(s) ? 1 : 0;
    return SPK_OK;
}

extern "C" int spk_solver_pipeline(spk_solver s, int *fused, int *points) {
    SPK_CHECK_SOLVER_RO(s);
    if (fused) *fused = fused_ok(s) ? 1 : 0;
    if (points) *points = s->fused_ef;
    return SPK_OK;
}

// ==========================================================================
// C ABI on caller-owned buffers
// ==========================================================================
// The reference's Physics, TimeStepping and RKLegendre methods take the field
// arrays as arguments (physics.hpp:18-42, timestepping.hpp:39-43, rkl.hpp:42) and
// its host classes own every device array (Supervisor::spookyGpuAlloc,
// supervisor.hpp:74-78).  The *_ptr entries run the solver's kernels on arrays the
// caller owns: for the duration of the call the solver's own state arrays are
// replaced by the caller's, and every kernel is enqueued with the caller's
// addresses.  The solver keeps only its transform buffers, twiddles and
// reduction partials.

__global__ __launch_bounds__(256) void k_maxabs(const double *__restrict__ r, long long nrows, int nz, int nzp,
                                                unsigned long long *__restrict__ out) {
    __shared__ double sm[4];
    double m = 0.0;
    const long long n = nrows * nz;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        const long long row = p / nz;
        m = fmax(m, fabs(r[row * nzp + (p - row * nz)]));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, sm[w]);
        // non-negative doubles order as their bit patterns
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

__global__ void k_zero_u64(unsigned long long *p) {
    if (threadIdx.x == 0) *p = 0ull;
}

namespace {

struct Borrow {
    spk_solver s;
    cpx<double> *U, *S, *dU, *uc0, *uc1, *du0;
    double *rb;
    Borrow(spk_solver s_, const void *u, void *scr, void *du, double *r)
        : s(s_), U(s_->U), S(s_->S), dU(s_->dU), uc0(s_->rUc0), uc1(s_->rUc1), du0(s_->rdU0), rb(s_->realbuf) {
        if (u) s->U = static_cast<cpx<double> *>(const_cast<void *>(u));
        if (scr) s->S = static_cast<cpx<double> *>(scr);
        if (du) s->dU = static_cast<cpx<double> *>(du);
        if (r) s->realbuf = r;
    }
    ~Borrow() {
        s->U = U; s->S = S; s->dU = dU; s->realbuf = rb;
        s->rUc0 = uc0; s->rUc1 = uc1; s->rdU0 = du0;
    }
};

bool aligned16(const void *p) { return p != nullptr && ((uintptr_t)p & 15u) == 0; }

#define SPK_CHECK_PTR(p, what) \
    SPK_REQUIRE(aligned16(p), SPK_ERR_ARG, "%s: %s is null or not 16-byte aligned", __func__, what)

int complex2real_into(spk_solver s, const void *in, double *out, int nf) {
    const Grid3 &g = s->g;
    PassArgs a{in, out, nf, g.ntc, g.ntc, 1.0 / ((double)g.nx * g.ny * g.nz)};
    SPK_TRY(launch_pass_x(s->ctx, g, s->tw64, +1, SPK_F64, SPK_F64, a));
    PassArgs b{out, out, nf, g.ntc, g.ntc, 1.0};
    SPK_TRY(launch_pass_y(s->ctx, g, s->tw64, +1, SPK_F64, SPK_F64, b));
    return launch_pass_z_c2r(s->ctx, g, s->tw64, SPK_F64, SPK_F64, b);
}

// transform buffers of the anisotropic flux, made at the first call that needs them
int need_flux_buffers(spk_solver s) {
    if (s->W2 && s->P2) return SPK_OK;
    const size_t fb = (size_t)s->g.ntc * esize(s->bulk);
    if (!s->W2) SPK_TRY(dalloc(&s->W2, fb * 6));
    if (!s->P2) SPK_TRY(dalloc(&s->P2, fb * 3));
    return SPK_OK;
}

}  // namespace

extern "C" int spk_compute_dt_ptr(spk_solver s, const void *complex_fields, double *real_buffer) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    Borrow b(s, complex_fields, nullptr, nullptr, real_buffer);
    SPK_TRY(compute_dt(s));
    if (real_buffer) SPK_TRY(complex2real_into(s, complex_fields, real_buffer, s->nf));
    return SPK_OK;
}

extern "C" int spk_compute_dfield_ptr(spk_solver s, const void *complex_fields, double *real_buffer,
                                      void *complex_dfields) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    SPK_CHECK_PTR(complex_dfields, "complex_dfields");
    (void)real_buffer;  // the products are formed from complex_fields
    Borrow b(s, complex_fields, nullptr, complex_dfields, nullptr);
    const double ks = s->tremap * s->p.shear;
    SPK_TRY(rhs_front(s, ks, false));
    s->products_valid = false;
    return rhs_back(s, ks, 0, 0.0, 0.0);
}

extern "C" int spk_rk3_ptr(spk_solver s, void *complex_fields, double *real_buffer, void *scratch) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    SPK_REQUIRE(scratch == nullptr || aligned16(scratch), SPK_ERR_ARG, "spk_rk3_ptr: scratch not 16-byte aligned");
    Borrow b(s, complex_fields, scratch, nullptr, real_buffer);
    return rk3(s);
}

extern "C" int spk_hydro_mhd_advance_ptr(spk_solver s, void *complex_fields, double *real_buffer, void *scratch) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    SPK_REQUIRE(scratch == nullptr || aligned16(scratch), SPK_ERR_ARG,
                "spk_hydro_mhd_advance_ptr: scratch not 16-byte aligned");
    Borrow b(s, complex_fields, scratch, nullptr, real_buffer);
    return hydro_mhd_advance(s);
}

extern "C" int spk_clean_divergence_ptr(spk_solver s, void *complex_fields) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    Borrow b(s, complex_fields, nullptr, nullptr, nullptr);
    return clean_divergence(s);
}

extern "C" int spk_project(spk_solver s, void *complex_vec) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_vec, "complex_vec");
    Batch bt;
    SPK_TRY(make_batch(s->ctx, {complex_vec}, &bt));
    k_clean_div<<<nblk(s->g.ntc), 256, 0, s->ctx->stream>>>(static_cast<cpx<double> *>(complex_vec), s->g.ntc,
                                                            kgeom(s, s->tremap * s->p.shear), 0, -1, bt);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

extern "C" int spk_complex2real_ptr(spk_solver s, const void *complex_in, double *real_out, int num_fields) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_in, "complex_in");
    SPK_CHECK_PTR(real_out, "real_out");
    SPK_REQUIRE(num_fields >= 1, SPK_ERR_ARG, "spk_complex2real_ptr: num_fields %d", num_fields);
    return complex2real_into(s, complex_in, real_out, num_fields);
}

extern "C" int spk_rkl2_cycle_ptr(spk_solver s, void *complex_fields, double *real_buffer) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    Borrow b(s, complex_fields, nullptr, nullptr, real_buffer);
    return rkl2_cycle(s);
}

extern "C" int spk_rkl2_stage(spk_solver s, int stage, const double *coef, double dt, void *complex_fields,
                              void *uc0, void *uc1, void *du0) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(s->nsts > 0, SPK_ERR_STATE, "spk_rkl2_stage: no sts variables in this run");
    SPK_REQUIRE(stage >= 1 && coef, SPK_ERR_ARG, "spk_rkl2_stage: bad stage %d or null coef", stage);
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    SPK_CHECK_PTR(uc0, "uc0");
    SPK_CHECK_PTR(uc1, "uc1");
    SPK_CHECK_PTR(du0, "du0");
    Borrow b(s, complex_fields, nullptr, nullptr, nullptr);
    s->rUc0 = static_cast<cpx<double> *>(uc0);
    s->rUc1 = static_cast<cpx<double> *>(uc1);
    s->rdU0 = static_cast<cpx<double> *>(du0);
    if (s->rkl_aniso) SPK_TRY(rkl2_front(s));
    if (s->bulk == SPK_F64) return rkl2_stage_launch<double>(s, stage, coef[0], coef[1], coef[2] * dt, coef[3] * dt);
    return rkl2_stage_launch<float>(s, stage, coef[0], coef[1], coef[2] * dt, coef[3] * dt);
}

extern "C" int spk_aniso_flux(spk_solver s, const void *complex_fields, double *real_buffer, void *complex_dtheta) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(s->p.anisotropic_diffusion, SPK_ERR_STATE, "spk_aniso_flux: anisotropic_diffusion is off");
    SPK_CHECK_PTR(complex_fields, "complex_fields");
    SPK_CHECK_PTR(complex_dtheta, "complex_dtheta");
    (void)real_buffer;  // B and grad theta are transformed from complex_fields
    SPK_TRY(need_flux_buffers(s));
    Borrow b(s, complex_fields, nullptr, nullptr, nullptr);
    SPK_TRY(rkl2_front(s));
    KGeom kg = kgeom(s, s->tremap * s->p.shear);
    cpx<double> *z = static_cast<cpx<double> *>(complex_dtheta);
    const unsigned gb = red_grid(s->g.ntc);
    if (s->bulk == SPK_F64)
        k_diag_div<<<gb, 256, 0, s->ctx->stream>>>((const cpx<double> *)s->P2, z, s->g.ntc, kg, 1);
    else
        k_diag_div<<<gb, 256, 0, s->ctx->stream>>>((const cpx<float> *)s->P2, z, s->g.ntc, kg, 1);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

extern "C" int spk_maxabs(spk_solver s, const double *real_field, double *out_dev) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(real_field && out_dev && ((uintptr_t)out_dev & 7u) == 0, SPK_ERR_ARG, "spk_maxabs: bad pointers");
    unsigned long long *o = reinterpret_cast<unsigned long long *>(out_dev);
    k_zero_u64<<<1, 64, 0, s->ctx->stream>>>(o);
    const long long rows = (long long)s->g.nx * s->g.ny;
    k_maxabs<<<red_grid(rows * s->g.nz), 256, 0, s->ctx->stream>>>(real_field, rows, s->g.nz, 2 * s->g.nzh, o);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

extern "C" int spk_energy_ptr(spk_solver s, const void *complex_field, double *out) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_field, "complex_field");
    SPK_REQUIRE(out, SPK_ERR_ARG, "spk_energy_ptr: null out");
    const unsigned nb = red_grid(s->g.ntc);
    k_spec_sum<<<nb, 256, 0, s->ctx->stream>>>(static_cast<const cpx<double> *>(complex_field), s->g.ntc, s->kg0, 0,
                                               s->part);
    SPK_LAUNCH_CHECK();
    const double nt = (double)s->g.nx * s->g.ny * s->g.nz;
    return reduce_final(s, (int)nb, 1.0 / (nt * nt), out);
}

extern "C" int spk_enstrophy_ptr(spk_solver s, const void *complex_vec, double *out) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_vec, "complex_vec");
    SPK_REQUIRE(out, SPK_ERR_ARG, "spk_enstrophy_ptr: null out");
    const unsigned nb = red_grid(s->g.ntc);
    k_spec_sum<<<nb, 256, 0, s->ctx->stream>>>(static_cast<const cpx<double> *>(complex_vec), s->g.ntc,
                                               kgeom(s, s->tremap * s->p.shear), 1, s->part);
    SPK_LAUNCH_CHECK();
    const double nt = (double)s->g.nx * s->g.ny * s->g.nz;
    return reduce_final(s, (int)nb, 1.0 / (nt * nt), out);
}

extern "C" int spk_correlation_ptr(spk_solver s, const double *real_a, const double *real_b, double *out) {
    SPK_CHECK_SOLVER(s);
    SPK_REQUIRE(real_a && real_b && out, SPK_ERR_ARG, "spk_correlation_ptr: null argument");
    const long long rows = (long long)s->g.nx * s->g.ny;
    const unsigned nb = red_grid(rows * s->g.nz);
    k_real_dot<<<nb, 256, 0, s->ctx->stream>>>(real_a, real_b, rows, s->g.nz, 2 * s->g.nzh, s->part);
    SPK_LAUNCH_CHECK();
    return reduce_final(s, (int)nb, 1.0 / ((double)s->g.nx * s->g.ny * s->g.nz), out);
}

extern "C" int spk_spectrum_ptr(spk_solver s, const void *complex_a, const void *complex_b, double *out, int cap,
                                int *nbins) {
    SPK_CHECK_SOLVER(s);
    SPK_CHECK_PTR(complex_a, "complex_a");
    SPK_CHECK_PTR(complex_b, "complex_b");
    SPK_REQUIRE(out && cap >= 0, SPK_ERR_ARG, "spk_spectrum_ptr: bad output");
    if (nbins) *nbins = s->nbins;
    return spectrum_of(s, static_cast<const cpx<double> *>(complex_a), static_cast<const cpx<double> *>(complex_b),
                       out, cap);
}

// ==========================================================================
// Energy of a spectral field without allocating
// ==========================================================================
// spk_energy_noalloc is spk_energy_ptr for callers that cannot allocate device
// memory.  Its reduction partials live in `scratch`, a static device array
// every GPU holds once for the life of the process.  In host code the array's
// name is not a device address: noalloc_setup looks the address up on each
// device once (spk_ctx_create calls it with the device current) and the entry
// point passes that address to the kernels.  Every call on one GPU shares that
// GPU's array; calls that overlap on one GPU race.  Each call writes partials
// 0 .. nb-1 before its final sum reads exactly those, so nothing needs
// clearing between calls.

static __device__ __align__(256) double scratch[SPK_NOALLOC_PARTS];
static_assert(SPK_NOALLOC_PARTS >= RED_BLOCKS, "the scratch array holds every reduction partial");

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

}  // namespace

namespace spk {

int noalloc_setup(int device) {
    SPK_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, SPK_ERR_ARG, "noalloc_setup: device %d", device);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return SPK_OK;
    int cur = -1;
    SPK_HIP(hipGetDevice(&cur));
    SPK_REQUIRE(cur == device, SPK_ERR_STATE, "noalloc_setup: device %d is not current (%d is)", device, cur);
    void *p = nullptr;
    SPK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SPK_REQUIRE(p != nullptr, SPK_ERR_HIP, "noalloc_setup: no address for the scratch array on device %d", device);
    g_noalloc_ptr[device] = static_cast<double *>(p);
    return SPK_OK;
}

}  // namespace spk

extern "C" int spk_energy_noalloc(spk_ctx ctx, int nx, int ny, int nz, const void *complex_field, double *out_dev) {
    SPK_REQUIRE(ctx, SPK_ERR_ARG, "spk_energy_noalloc: null ctx");
    SPK_REQUIRE(nx > 0 && ny > 0 && nz > 0, SPK_ERR_ARG, "spk_energy_noalloc: bad grid %d x %d x %d", nx, ny, nz);
    SPK_REQUIRE(complex_field && ((uintptr_t)complex_field & 15u) == 0, SPK_ERR_ARG,
                "spk_energy_noalloc: complex_field must be a 16-byte aligned device pointer");
    SPK_REQUIRE(out_dev && ((uintptr_t)out_dev & 7u) == 0, SPK_ERR_ARG,
                "spk_energy_noalloc: out_dev must be an 8-byte aligned device pointer");
    SPK_REQUIRE(ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES, SPK_ERR_ARG, "spk_energy_noalloc: device %d",
                ctx->device);
    double *part;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        part = g_noalloc_ptr[ctx->device];
    }
    SPK_REQUIRE(part, SPK_ERR_STATE, "spk_energy_noalloc: scratch address of device %d not set up", ctx->device);
    const Grid3 g = make_grid(nx, ny, nz);
    const unsigned nb = red_grid(g.ntc);
    SPK_REQUIRE(nb <= (unsigned)SPK_NOALLOC_PARTS, SPK_ERR_ARG,
                "spk_energy_noalloc: %u reduction partials exceed the %d-entry scratch array", nb, SPK_NOALLOC_PARTS);
    KGeom kg{};
    kg.nx = nx;
    kg.ny = ny;
    kg.nzh = g.nzh;
    k_spec_sum<<<nb, 256, 0, ctx->stream>>>(static_cast<const cpx<double> *>(complex_field), g.ntc, kg, 0, part);
    SPK_LAUNCH_CHECK();
    const double nt = (double)nx * ny * nz;
    k_final_sum<<<1, 1024, 0, ctx->stream>>>(part, (int)nb, 1.0 / (nt * nt), out_dev);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}
