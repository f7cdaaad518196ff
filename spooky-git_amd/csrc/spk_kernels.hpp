// spk_kernels.hpp -- device kernels of the solver's hot path (one RHS evaluation
// = inverse x pass -> inverse y pass -> fused z kernel -> forward y pass ->
// forward x pass -> spectral epilogue).
#pragma once

#include "spk_rows.hpp"

namespace spk {

// Device-resident clock of one run (TimeStepping members current_time,
// current_dt, tremap, dt_hyp, dt_par, dt_sts; InputOutput::t_lastsnap; step).
struct TimeState {
    double t, dt, tremap, dt_hyp, dt_par, dt_sts, t_lastsnap, step;
};

// Spectral grid geometry: k is computed from the index instead of loaded from
// kvec/mask arrays (wavevector.cu:50-96 restated per point).
struct KGeom {
    int nx, ny, nzh;
    double kfx, kfy, kfz;    // 2 pi / L_i (kfz = 0 with with2d)
    double thx, thy, thz;    // dealias thresholds (2/3) * kmax_i before rescaling
    int antialias;
    double kxshift;          // tremap * shear: sheared kx = kx0 + kxshift * ky (ShearWavevector)
};

// what a module outside the solver needs of it (the closure coupling)
struct SolverViews {
    cpx<double> *U;
    double *realbuf;
    TimeState *ts;
    KGeom kg;
    int VX;
    long long ntc;
    int nx, ny, nz;
    spk_ctx ctx;
    bool incompressible;
};
void solver_views(spk_solver s, SolverViews *v);

__device__ __forceinline__ void kvec_at(const KGeom &g, int i, int j, int k, double &kx, double &ky,
                                        double &kz, double &m) {
    const int ii = i < g.nx / 2 ? i : i - g.nx;  // fmod(i + NX/2, NX) - NX/2
    const int jj = j < g.ny / 2 ? j : j - g.ny;
    const double kx0 = g.kfx * (double)ii;
    ky = g.kfy * (double)jj;
    kz = g.kfz * (double)k;
    kx = kx0 + g.kxshift * ky;
    m = 1.0;
    if (g.antialias && (fabs(kx0) > g.thx || fabs(ky) > g.thy || fabs(kz) > g.thz)) m = 0.0;
}

// ------------------------------------------------------------------------
// RKL2 stage (rkl.cu:496-585, rkl2_stage cuda_kernels_generic.cu:160-172) fused
// with ParabolicTermsSTS (physics.cu:233-330) for every sts variable
// ------------------------------------------------------------------------
struct RklConst {
    KGeom kg;
    long long ntc;
    int nsts;
    int field[8];    // state index of sts variable n
    int kind[8];     // 0: -nu k^2, 1: -nu_m k^2, 2: -nu_th k^2, 3: div of anisotropic flux
    double nu, nu_m, nu_th;
};

// Device-sized RKL2 cycle: the stage count and coefficient table of the
// recurrence (rkl.cu:466-585) computed on the device from the step's dt and
// dt_sts, so that a cycle needs no host synchronization and can be captured in
// a graph.  RKL_SMAX stages bound what a plan can hold.
constexpr int RKL_SMAX = 64;
struct RklPlan {
    int s, overflow;
    int active[RKL_SMAX + 1];   // active[j] = 1 for stages 1 <= j <= s
    double coef[4 * RKL_SMAX];  // mu, nu, mu~, gamma~ of stage j at 4 (j - 1)
};

// ------------------------------------------------------------------------
// inverse x pass from the FP64 state, with optional i*k_d derivative
// (Gradient, cuda_kernels_generic.cu:249-260) and the 1/N scaling of c2r_fft
// ------------------------------------------------------------------------
struct InvSlot {
    int src;    // state field index
    int deriv;  // -1: the field itself; 0,1,2: i*k_{x,y,z} * field
};
constexpr int MAX_SLOTS = 12;
struct InvProg {
    InvSlot s[MAX_SLOTS];
};

template <class T, int N>
__global__ __launch_bounds__(strided_tcols(N) * strided_tpl(N)) void k_inv_x(
    const cpx<double> *__restrict__ U, cpx<T> *__restrict__ W, long long fs, InvProg prog, KGeom kg,
    double scale, const cpx<T> *__restrict__ tw, const int *__restrict__ gate, Batch bt) {
    U = rebase(bt, 0, U);
    W = rebase(bt, 1, W);
    gate = rebase(bt, 2, gate);
    if (gate != nullptr && *gate == 0) return;
    constexpr int E = strided_e(N);
    constexpr int TPL = N / E;
    constexpr int TC = strided_tcols(N);
    constexpr int LL = lds_len(N);
    __shared__ cpx<T> lds[line_needs_lds<N>() ? TC * LL : 1];

    const long long ncols = (long long)kg.ny * kg.nzh;
    const int tid = threadIdx.x;
    const int c = tid % TC, t = tid / TC;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const long long col = (long long)tile * TC + c;
    const bool active = col < ncols;
    const InvSlot sl = prog.s[blockIdx.y];
    const int j = active ? (int)(col / kg.nzh) : 0;
    const int k = active ? (int)(col - (long long)j * kg.nzh) : 0;
    const cpx<double> *src = U + (long long)sl.src * fs + col;

    cpx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
        const int n = t + i * TPL;
        cpx<double> a = active ? src[(long long)n * ncols] : cpx<double>{0, 0};
        if (sl.deriv >= 0) {
            double kx, ky, kz, m;
            kvec_at(kg, n, j, k, kx, ky, kz, m);
            const double kd = sl.deriv == 0 ? kx : (sl.deriv == 1 ? ky : kz);
            a = cpx<double>{-(kd * a.y), kd * a.x};
        }
        v[i] = cconv<T>(cscale(a, scale));
    }
    LineFFT<T, N, E, +1>::run(v, lds + c * LL, t, tw);
    if (active) {
        cpx<T> *dst = W + (long long)blockIdx.y * fs + col;
#pragma unroll
        for (int i = 0; i < E; i++) dst[(long long)(t + i * TPL) * ncols] = v[i];
    }
}

// ------------------------------------------------------------------------
// fused z kernel: C2R of the input rows, real-space products, R2C of the
// products.  Real-space fields never reach HBM.
// ------------------------------------------------------------------------
// ZP_BR: the B rows come in real space (z values of a padded real row, read as
// NZ/2 complex pairs) from a second buffer instead of through a C2R -- B does
// not change during an RKL2 cycle, so its real-space rows are made once per cycle
enum : int { ZP_U = 1, ZP_B = 2, ZP_TH = 4, ZP_G = 8, ZP_BR = 16 };

template <int PROG>
struct ZProg {
    static constexpr bool U = PROG & ZP_U, B = PROG & ZP_B, TH = PROG & ZP_TH, G = PROG & ZP_G;
    static constexpr bool BR = (PROG & ZP_BR) && B;
    // inputs: u(3) | B(3) | theta | grad theta(3)
    static constexpr int IU = 0, IB = IU + (U ? 3 : 0), ITH = IB + (B ? 3 : 0), IG = ITH + (TH ? 1 : 0);
    static constexpr int NIN = IG + (G ? 3 : 0);
    // outputs: Basdevant T(5) | emf(3) | u theta(3) | anisotropic flux(3)
    static constexpr int OT = 0, OE = OT + (U ? 5 : 0), OQ = OE + ((U && B) ? 3 : 0),
                         OA = OQ + ((U && TH) ? 3 : 0);
    static constexpr int NOUT = OA + ((G && B) ? 3 : 0);
};

struct ZConst {
    double chi, OmegaT2;  // anisotropic conduction: 1/reynolds_ani, mti_frequency_squared
    int sd;               // strat_direction
};

// Product q at one real-space point; a[] holds the NIN input values there.
// Expressions follow the reference kernels term by term:
//   TracelessShearMatrix[MHD] (cuda_kernels.cu:39-55, 216-234), MagneticEmf
//   (cuda_kernels.cu:240-252), EnergyFluxVector (:277-289), ComputeBGradTheta
//   + ComputeAnisotropicHeatFlux (:360-383).
template <int PROG, class T>
__device__ __forceinline__ T zproduct(int q, const T *a, const ZConst &zc) {
    using Z = ZProg<PROG>;
    const T two = (T)2.0, three = (T)3.0;
    if constexpr (Z::U) {
        const T V0 = a[Z::IU], V1 = a[Z::IU + 1], V2 = a[Z::IU + 2];
        if (q < Z::OE) {
            if constexpr (Z::B) {
                const T B0 = a[Z::IB], B1 = a[Z::IB + 1], B2 = a[Z::IB + 2];
                switch (q - Z::OT) {
                    case 0: return (two * V0 * V0 - V1 * V1 - V2 * V2 - two * B0 * B0 + B1 * B1 + B2 * B2) / three;
                    case 1: return V0 * V1 - B0 * B1;
                    case 2: return V0 * V2 - B0 * B2;
                    case 3: return (-V0 * V0 + two * V1 * V1 - V2 * V2 + B0 * B0 - two * B1 * B1 + B2 * B2) / three;
                    default: return V1 * V2 - B1 * B2;
                }
            } else {
                switch (q - Z::OT) {
                    case 0: return (two * V0 * V0 - V1 * V1 - V2 * V2) / three;
                    case 1: return V0 * V1;
                    case 2: return V0 * V2;
                    case 3: return (-V0 * V0 + two * V1 * V1 - V2 * V2) / three;
                    default: return V1 * V2;
                }
            }
        }
        if constexpr (Z::B) {
            if (q < Z::OQ) {
                const T B0 = a[Z::IB], B1 = a[Z::IB + 1], B2 = a[Z::IB + 2];
                switch (q - Z::OE) {
                    case 0: return -B1 * V2 + B2 * V1;
                    case 1: return B0 * V2 - B2 * V0;
                    default: return -B0 * V1 + B1 * V0;
                }
            }
        }
        if constexpr (Z::TH) {
            if (q < Z::OA) {
                const T th = a[Z::ITH];
                const int c = q - Z::OQ;
                return (c == 0 ? V0 : (c == 1 ? V1 : V2)) * th;
            }
        }
    }
    if constexpr (Z::G && Z::B) {
        const T B0 = a[Z::IB], B1 = a[Z::IB + 1], B2 = a[Z::IB + 2];
        const T bg = B0 * a[Z::IG] + B1 * a[Z::IG + 1] + B2 * a[Z::IG + 2];
        const T Bsq = B0 * B0 + B1 * B1 + B2 * B2;
        const T bs = zc.sd == 0 ? B0 : (zc.sd == 1 ? B1 : B2);
        const int c = q - Z::OA;
        const T Bc = c == 0 ? B0 : (c == 1 ? B1 : B2);
        return (T)zc.chi * Bc * (bg + (T)zc.OmegaT2 * bs) / Bsq;
    }
    return (T)0;
}

// points per thread in the fused z kernel: all NIN input rows stay in registers.
// 4 in both precisions (FP32 with 8 points per thread ran the 256^3 FP32-bulk
// step 2% slower: 4.05 against 3.95 ms)
template <class T>
__host__ __device__ constexpr int zmid_e(int m) {
    return line_e(m, 4);
}
template <class T>
__host__ __device__ constexpr int zmid_rw(int m) {
    return (m / zmid_e<T>(m)) >= 256 ? 1 : 256 / (m / zmid_e<T>(m));
}

// BAT: a batched launch (blockIdx.z is the box); single-box launches use the
// BAT = false instantiation, whose code has no table lookups
template <class T, int NZ, int PROG, bool BAT = false>
__global__ __launch_bounds__(256) void k_zmid(const cpx<T> *__restrict__ W, const cpx<T> *__restrict__ Wr,
                                              cpx<T> *__restrict__ P, long long nrows, long long fs,
                                              const cpx<T> *__restrict__ tw_nz,
                                              const cpx<T> *__restrict__ tw_m, ZConst zc,
                                              double *__restrict__ pmax, const int *__restrict__ gate, Batch bt) {
    if constexpr (BAT) {
        W = rebase(bt, 0, W);
        Wr = rebase(bt, 1, Wr);
        P = rebase(bt, 2, P);
        pmax = rebase(bt, 3, pmax);
        gate = rebase(bt, 4, gate);
        zc.chi = 1.0 / batch_row(bt)->reynolds_ani;
    }
    if (gate != nullptr && *gate == 0) return;
    using Z = ZProg<PROG>;
    constexpr int M = NZ / 2;
    constexpr int NZH = M + 1;
    constexpr int E = zmid_e<T>(M);
    using Row = ZRow<T, M, E, true>;
    constexpr int RW = zmid_rw<T>(M);
    constexpr int NT = RW * Row::TPL;
    __shared__ cpx<T> lds[RW * Row::LL];
    __shared__ double smax[6][(NT + 63) / 64];

    const int tid = threadIdx.x;
    const int rl = tid / Row::TPL, t = tid % Row::TPL;
    const long long row = (long long)blockIdx.x * RW + rl;
    const bool active = row < nrows;
    cpx<T> *L = lds + rl * Row::LL;

    // every input row is loaded before the first transform: one memory latency
    // per block instead of one per field (the LDS fences inside a transform keep
    // later loads from being issued early)
    cpx<T> in[Z::NIN][E];
    cpx<T> nyq[Z::NIN];
#pragma unroll
    for (int f = 0; f < Z::NIN; f++) {
        if (Z::BR && f >= Z::IB && f < Z::IB + 3) {
            const cpx<T> *r = Wr + (f - Z::IB) * fs + row * NZH;
#pragma unroll
            for (int i = 0; i < E; i++) in[f][i] = active ? r[t + i * Row::TPL] : cpx<T>{(T)0, (T)0};
        } else {
            Row::load(W + f * fs + row * NZH, active, t, in[f], nyq[f]);
        }
    }
    // Each thread's twiddles are read once, here, and kept in registers for all
    // NIN + NOUT row transforms: a twiddle load issued after the stores of a
    // previous row transform would wait for those stores (vector memory
    // operations retire in issue order), and twiddle reads from LDS collide in
    // banks.
    typename Row::RowTw rt;
    Row::load_tw(rt, t, tw_nz, tw_m);
#pragma unroll
    for (int f = 0; f < Z::NIN; f++)
        if (!(Z::BR && f >= Z::IB && f < Z::IB + 3)) Row::c2r_tw(in[f], nyq[f], active, L, t, rt, (T)1);

    if constexpr (Z::U) {
        if (pmax) {
            // max |u_i|, |B_i| over the block's physical points (compute_dt.cu:90-106)
            double mx[6];
#pragma unroll
            for (int c = 0; c < 6; c++) mx[c] = 0.0;
            if (active) {
#pragma unroll
                for (int i = 0; i < E; i++) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        mx[c] = fmax(mx[c], fmax(fabs((double)in[Z::IU + c][i].x), fabs((double)in[Z::IU + c][i].y)));
                        if constexpr (Z::B)
                            mx[3 + c] = fmax(mx[3 + c], fmax(fabs((double)in[Z::IB + c][i].x), fabs((double)in[Z::IB + c][i].y)));
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 6; c++) {
                double vmax = mx[c];
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o, 64));
                if ((tid & 63) == 0) smax[c][tid >> 6] = vmax;
            }
            __syncthreads();
            if (tid < 6) {
                double vmax = 0.0;
                for (int w = 0; w < (NT + 63) / 64; w++) vmax = fmax(vmax, smax[tid][w]);
                pmax[(long long)blockIdx.x * 6 + tid] = vmax;
            }
        }
    }

#pragma unroll
    for (int q = 0; q < Z::NOUT; q++) {
        cpx<T> v[E];
#pragma unroll
        for (int i = 0; i < E; i++) {
            T ae[Z::NIN], ao[Z::NIN];
#pragma unroll
            for (int f = 0; f < Z::NIN; f++) {
                ae[f] = in[f][i].x;
                ao[f] = in[f][i].y;
            }
            v[i] = cpx<T>{zproduct<PROG>(q, ae, zc), zproduct<PROG>(q, ao, zc)};
        }
        Row::r2c_tw(v, active, P + q * fs + row * NZH, L, t, rt);
    }
}

// ------------------------------------------------------------------------
// spectral epilogue: one RHS (compute_dfield.cu:11-90) per point, then the
// RK3 stage update (hydro_mhd_advance.cu:114-164) or a plain dU store
// ------------------------------------------------------------------------

// Field layouts of Parameters::popVariablesGrid (parameters.cu:568-640)
enum : int { LY_HEAT = 0, LY_HD = 1, LY_HD_TH = 2, LY_MHD = 3, LY_MHD_TH = 4 };
template <int LY>
struct Layout {
    static constexpr bool INC = LY != LY_HEAT;
    static constexpr bool MHD = LY == LY_MHD || LY == LY_MHD_TH;
    static constexpr bool HASTH = LY == LY_HEAT || LY == LY_HD_TH || LY == LY_MHD_TH;
    static constexpr bool BOUSS = INC && HASTH;
    static constexpr int VX = 0, BX = 3;
    static constexpr int TH = LY == LY_HEAT ? 0 : (MHD ? 6 : 3);
    static constexpr int NF = LY == LY_HEAT ? 1 : 3 + (MHD ? 3 : 0) + (HASTH ? 1 : 0);
};

struct EpiConst {
    KGeom kg;
    long long ntc;
    int stratification, sd, shearing, rot_active;
    int aniso_rhs;   // theta gets div of the anisotropic flux products
    int diff_mask;   // bit f: explicit diffusion of field f in this RHS (not an sts variable)
    double nu, nu_m, nu_th, N2, shear, omega;
    int oT, oE, oQ, oA;
    double gam, xi;  // gammaRK, xiRK of the stage
};

__device__ __forceinline__ cpx<double> cadd(cpx<double> a, cpx<double> b) { return {a.x + b.x, a.y + b.y}; }
// -I m (sum): (m * sum.y, -m * sum.x)
__device__ __forceinline__ cpx<double> mul_minus_im(double m, cpx<double> s) { return {m * s.y, -m * s.x}; }
// I m (sum): (-m * sum.y, m * sum.x)
__device__ __forceinline__ cpx<double> mul_plus_im(double m, cpx<double> s) { return {-m * s.y, m * s.x}; }
// a*X + b*Y + c*Z with real coefficients
__device__ __forceinline__ cpx<double> kdot3(double a, cpx<double> X, double b, cpx<double> Y, double c, cpx<double> Z) {
    return {a * X.x + b * Y.x + c * Z.x, a * X.y + b * Y.y + c * Z.y};
}
// d += f * u when field q is diffused explicitly in this RHS
__device__ __forceinline__ cpx<double> diffuse(int mask, int q, cpx<double> d, cpx<double> u, double f) {
    const double g = (mask >> q & 1) ? f : 0.0;
    return (mask >> q & 1) ? cpx<double>{d.x + g * u.x, d.y + g * u.y} : d;
}

template <class T, int LY, int MODE>
__global__ __launch_bounds__(256) void k_epilogue(const cpx<T> *__restrict__ P, cpx<double> *__restrict__ U,
                                                  cpx<double> *__restrict__ S, cpx<double> *__restrict__ dU,
                                                  EpiConst c, const TimeState *__restrict__ ts, Batch bt) {
    using L = Layout<LY>;
    P = rebase(bt, 0, P);
    U = rebase(bt, 1, U);
    S = rebase(bt, 2, S);
    dU = rebase(bt, 3, dU);
    ts = rebase(bt, 4, ts);
    if (const BoxRow *br = batch_row(bt)) {
        c.nu = br->nu;
        c.nu_m = br->nu_m;
        c.nu_th = br->nu_th;
    }
    constexpr int NF = L::NF, VX = L::VX, BX = L::BX, TH = L::TH;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= c.ntc) return;
    const long long N = c.ntc;
    const long long rowi = idx / c.kg.nzh;
    const int k = (int)(idx - rowi * c.kg.nzh);
    const int j = (int)(rowi % c.kg.ny);
    const int i = (int)(rowi / c.kg.ny);
    double kx, ky, kz, m;
    kvec_at(c.kg, i, j, k, kx, ky, kz, m);
    const double k2 = kx * kx + ky * ky + kz * kz;

    cpx<double> d[NF], u[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        d[f] = cpx<double>{0.0, 0.0};
        u[f] = U[f * N + idx];
    }
    auto prod = [&](int q) { return cconv<double>(P[(long long)q * N + idx]); };

    // HyperbolicTerms (physics.cu:30-103)
    if constexpr (L::INC) {
        // NonLinHydroAdv (cuda_kernels.cu:82-98), T_zz = -T_xx - T_yy
        const cpx<double> Txx = prod(c.oT), Txy = prod(c.oT + 1), Txz = prod(c.oT + 2), Tyy = prod(c.oT + 3),
                          Tyz = prod(c.oT + 4);
        d[VX] = mul_minus_im(m, kdot3(kx, Txx, ky, Txy, kz, Txz));
        d[VX + 1] = mul_minus_im(m, kdot3(kx, Txy, ky, Tyy, kz, Tyz));
        const cpx<double> sxy = {Txx.x + Tyy.x, Txx.y + Tyy.y};
        d[VX + 2] = mul_minus_im(m, kdot3(kx, Txz, ky, Tyz, -kz, sxy));
    }
    if constexpr (L::MHD) {
        // MagneticShear (cuda_kernels.cu:257-272)
        const cpx<double> E0 = prod(c.oE), E1 = prod(c.oE + 1), E2 = prod(c.oE + 2);
        d[BX] = mul_plus_im(m, cpx<double>{ky * E2.x - kz * E1.x, ky * E2.y - kz * E1.y});
        d[BX + 1] = mul_plus_im(m, cpx<double>{kz * E0.x - kx * E2.x, kz * E0.y - kx * E2.y});
        d[BX + 2] = mul_plus_im(m, cpx<double>{kx * E1.x - ky * E0.x, kx * E1.y - ky * E0.y});
    }
    if constexpr (L::BOUSS) {
        // NonLinBoussinesqAdv (cuda_kernels.cu:293-304)
        d[TH] = mul_minus_im(m, kdot3(kx, prod(c.oQ), ky, prod(c.oQ + 1), kz, prod(c.oQ + 2)));
    }
    // SourceTerms (physics.cu:105-142)
    if constexpr (L::BOUSS) {
        if (c.stratification) {
            // BoussinesqStrat (cuda_kernels.cu:310-325)
            // one-hot blend keeps u[] in registers (no select between addresses)
            const double w0 = c.sd == 0 ? 1.0 : 0.0, w1 = c.sd == 1 ? 1.0 : 0.0, w2 = c.sd == 2 ? 1.0 : 0.0;
            const cpx<double> us = {w0 * u[VX].x + w1 * u[VX + 1].x + w2 * u[VX + 2].x,
                                    w0 * u[VX].y + w1 * u[VX + 1].y + w2 * u[VX + 2].y};
            const cpx<double> mth = {-u[TH].x, -u[TH].y};
            if (c.sd == 0) d[VX] = cadd(d[VX], mth);
            else if (c.sd == 1) d[VX + 1] = cadd(d[VX + 1], mth);
            else d[VX + 2] = cadd(d[VX + 2], mth);
            d[TH] = cadd(d[TH], cscale(us, c.N2));
        }
    }
    if constexpr (L::INC) {
        if (c.shearing) {
            // ShearingFlow (cuda_kernels.cu:332-340)
            d[VX + 1] = cadd(d[VX + 1], cscale(u[VX], c.shear));
            if constexpr (L::MHD) d[BX + 1] = cadd(d[BX + 1], cscale(u[BX], -c.shear));
        }
        if (c.rot_active) {
            // CoriolisForce (cuda_kernels.cu:344-355)
            const double w2 = 2.0 * c.omega;
            d[VX] = cadd(d[VX], cscale(u[VX + 1], w2));
            d[VX + 1] = cpx<double>{d[VX + 1].x - w2 * u[VX].x, d[VX + 1].y - w2 * u[VX].y};
        }
        // ParabolicTerms (physics.cu:145-230), nablaOpScalar (cuda_kernels_generic.cu:198-214)
        const double f = -c.nu * k2;
        d[VX] = diffuse(c.diff_mask, VX, d[VX], u[VX], f);
        d[VX + 1] = diffuse(c.diff_mask, VX + 1, d[VX + 1], u[VX + 1], f);
        d[VX + 2] = diffuse(c.diff_mask, VX + 2, d[VX + 2], u[VX + 2], f);
    }
    if constexpr (L::MHD) {
        const double f = -c.nu_m * k2;
        d[BX] = diffuse(c.diff_mask, BX, d[BX], u[BX], f);
        d[BX + 1] = diffuse(c.diff_mask, BX + 1, d[BX + 1], u[BX + 1], f);
        d[BX + 2] = diffuse(c.diff_mask, BX + 2, d[BX + 2], u[BX + 2], f);
    }
    if constexpr (L::HASTH) {
        if (c.aniso_rhs) {
            // DivergenceMask flag 1 (cuda_kernels_generic.cu:315-334)
            d[TH] = cadd(d[TH], mul_plus_im(m, kdot3(kx, prod(c.oA), ky, prod(c.oA + 1), kz, prod(c.oA + 2))));
        } else if (c.diff_mask >> TH & 1) {
            const cpx<double> t = cscale(u[TH], -c.nu_th * k2);
            d[TH] = (LY == LY_HEAT) ? t : cadd(d[TH], t);
        }
    }
    // GradPseudoPressure[Shearing] (cuda_kernels.cu:128-197)
    if constexpr (L::INC) {
        cpx<double> s = kdot3(kx, d[VX], ky, d[VX + 1], kz, d[VX + 2]);
        if (c.shearing) {
            const double sk = c.shear * ky;
            s = cpx<double>{sk * u[VX].x + kx * d[VX].x + ky * d[VX + 1].x + kz * d[VX + 2].x,
                            sk * u[VX].y + kx * d[VX].y + ky * d[VX + 1].y + kz * d[VX + 2].y};
        }
        const cpx<double> div = {-s.y, s.x};
        const double ik2 = k2 > 0.0 ? 1.0 / k2 : 1.0;
        const double ax = kx * ik2, ay = ky * ik2, az = kz * ik2;
        d[VX] = cadd(d[VX], cpx<double>{-ax * div.y, ax * div.x});
        d[VX + 1] = cadd(d[VX + 1], cpx<double>{-ay * div.y, ay * div.x});
        d[VX + 2] = cadd(d[VX + 2], cpx<double>{-az * div.y, az * div.x});
    }

    if constexpr (MODE == 0) {
#pragma unroll
        for (int f = 0; f < NF; f++) dU[f * N + idx] = d[f];
    } else {
        const double dt = ts->dt;
        const double a = c.gam * dt, b = c.xi * dt;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            // axpyComplex (cuda_kernels_generic.cu:137-144) with a = 1
            const cpx<double> base = (MODE == 1) ? u[f] : S[f * N + idx];
            const cpx<double> un = {base.x + a * d[f].x, base.y + a * d[f].y};
            U[f * N + idx] = un;
            if constexpr (MODE < 3) S[f * N + idx] = cpx<double>{un.x + b * d[f].x, un.y + b * d[f].y};
        }
    }
}

// ------------------------------------------------------------------------
// fused forward x pass + spectral epilogue + RK3 update (+ stage-3
// CleanFieldDivergence + the next stage's inverse x pass).  One block row per
// field group (blockIdx.y): velocity (products T_xx..T_yz), magnetic (emf),
// temperature (u theta).  After each product's x line FFT its contribution to
// k.T, k x emf or k.(u theta) is accumulated in registers, so the transformed
// products never go back to HBM.  The groups run in different blocks and update
// their fields in place, so no group reads another group's state (stratified
// runs take the temperature and velocity groups in the same blocks).  Not used with shearing
// (sheared k between stages, remap) or in-RHS conduction (its grad theta slots
// come from k_inv_x).
// ------------------------------------------------------------------------
// EF: points per thread of the x line (4 or 8, chosen per grid by the solver's
// tile autotune); TPL = N/EF threads per line, TCOLS columns per 256-thread block
__host__ __device__ constexpr int fx_e(int n, int ef) { return line_e(n, ef); }
__host__ __device__ constexpr int fx_tpl(int n, int ef) { return n / fx_e(n, ef); }
// at most 64 columns: a lane past the last column repeats the last column's
// loads and stores, which is safe only within the wave that owns that column
__host__ __device__ constexpr int fx_tcols(int n, int ef) {
    return fx_tpl(n, ef) >= 512 ? 1 : (512 / fx_tpl(n, ef) < 64 ? 512 / fx_tpl(n, ef) : 64);
}

// RK3 registers of the fused path.  Wray's scheme (hydro_mhd_advance.cu:114-164)
//   U1 = U0 + g1 dt L0, S1 = U1 + x1 dt L0, U2 = S1 + g2 dt L1, S2 = U2 + x2 dt L1,
//   U3 = S2 + g3 dt L2
// is kept with the second register holding the stage increment R = dt L in the
// bulk precision instead of S in FP64:
//   stage 1: reads U0, writes R0 (U1 = U0 + g1 R0 is formed where it is needed)
//   stage 2: reads U0, R0; current state U1 = U0 + cu R0 (cu = g1), base
//            S1 = U0 + cs R0 (cs = g1 + x1); writes U2 = S1 + g2 R1 and R1
//   stage 3: reads U2, R1; base S2 = U2 + cs R1 (cs = x2); writes U3 = S2 + g3 dt L2
// Every stage's state is formed from the rounded increments, so the nonlinear and
// linear terms of a stage see the same state.  The FP64 state accumulation is
// unchanged; with FP32 bulk the increments carry FP32 rounding, as the products
// they come from already do.
struct FusedConst {
    EpiConst e;
    int project;    // apply CleanDivergence to the updated u and B (stage 3)
    int inv;        // write the next stage's inverse x pass of the updated fields to W
    double scale;   // 1/(NX NY NZ) of the next inverse pass
    double cu, cs;  // coefficients of the stored increment in the current state and the base
};

// vmcnt(0) and lgkmcnt(0) waits (gfx9 s_waitcnt encodings; other counters at
// their maxima)
#define SPK_WAIT_VMCNT0() __builtin_amdgcn_s_waitcnt(0x0F70)
#define SPK_WAIT_LGKMCNT0() __builtin_amdgcn_s_waitcnt(0xC07F)

// One field group of k_fwdx_epilogue: GRP 0 velocity (5 products T_ij), 1
// magnetic (3 products emf), 2 temperature (3 products u theta).  Fields
// updated in place may only be read by the group that updates them, and
// BoussinesqStrat (cuda_kernels.cu:310-325) couples u_s and theta.  STRAT
// (stratified runs): the temperature group and then the velocity group run in
// one block; the temperature group reads u_s's current state (not yet updated)
// and leaves theta's current state at its points in W's first velocity slot,
// which the velocity group reads back before its inverse x pass writes that
// slot.  Every lane runs the same
// straight-line code: a lane past the last column works on the last column and
// stores the same values as that column's own lane, so no memory operation sits
// behind a branch.  Element offsets inside a field fit 32 bits and are added to
// wave-uniform field bases.
template <class T, int N, int LY, int MODE, int EF, int GRP, bool STRAT = false>
__device__ __forceinline__ void fx_group(const cpx<T> *P, cpx<double> *__restrict__ U, cpx<T> *__restrict__ R,
                                         cpx<T> *__restrict__ W, const FusedConst &fc,
                                         const cpx<T> *__restrict__ tw, cpx<T> *twl, cpx<T> *ldsc, cpx<T> *pbuf,
                                         const TimeState *__restrict__ ts, int tid, int nt, int t, long long col) {
    using L = Layout<LY>;
    constexpr int E = fx_e(N, EF);
    constexpr int TPL = N / E;
    constexpr int NPROD = GRP == 0 ? 5 : 3;
    constexpr int NCOMP = GRP == 2 ? 1 : 3;
    constexpr int F0 = GRP == 0 ? L::VX : (GRP == 1 ? L::BX : L::TH);
    // FP64 bulk: product lines reach LDS by direct global->LDS loads
    // (global_load_lds_dwordx4), which take no registers.  Lane l of wave w puts
    // slot i at pbuf[(w E + i) 64 + l] and reads back only its own slots, so only
    // the wave's own waits order them.  FP32 bulk (8-byte elements) prefetches
    // into registers.
    constexpr bool DMA = sizeof(cpx<T>) == 16;
    const EpiConst &c = fc.e;
    const long long N0 = c.ntc;
    const unsigned ncols = (unsigned)((long long)c.kg.ny * c.kg.nzh);
    const unsigned ucol = (unsigned)col;
    const int j = (int)(col / c.kg.nzh);
    const int kk = (int)(col - (long long)j * c.kg.nzh);
    const int p0 = GRP == 0 ? c.oT : (GRP == 1 ? c.oE : c.oQ);
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    // ky, kz are constant along the x line
    double ky, kz;
    {
        double kx0, m0;
        kvec_at(c.kg, 0, j, kk, kx0, ky, kz, m0);
    }

    cpx<double> acc[NCOMP][E];
#pragma unroll
    for (int q = 0; q < NCOMP; q++)
#pragma unroll
        for (int i = 0; i < E; i++) acc[q][i] = cpx<double>{0.0, 0.0};

    cpx<T> vn[DMA ? 1 : E];
    auto fetch = [&](int q) {
        const cpx<T> *src = P + (long long)(p0 + q) * N0;
#pragma unroll
        for (int i = 0; i < E; i++) {
            const cpx<T> *g = src + (ucol + (unsigned)(t + i * TPL) * ncols);
            if constexpr (DMA)
                __builtin_amdgcn_global_load_lds((const void *)g,
                                                 (__attribute__((address_space(3))) void *)(pbuf + (wv * E + i) * 64),
                                                 16, 0, 0);
            else
                vn[i] = *g;
        }
    };
    fetch(0);
    // Twiddles are read from LDS: a global twiddle load would wait for every
    // earlier vector memory operation (they retire in issue order), i.e. for the
    // next product's prefetch and for the state stores.
    for (int e = tid; e < N; e += nt) twl[e] = tw[e];
    __syncthreads();

#pragma unroll 1
    for (int q = 0; q < NPROD; q++) {
        cpx<T> v[E];
        if constexpr (DMA) {
            SPK_WAIT_VMCNT0();
#pragma unroll
            for (int i = 0; i < E; i++) v[i] = pbuf[(wv * E + i) * 64 + lane];
            SPK_WAIT_LGKMCNT0();  // the slots are read before they are refilled
        } else {
#pragma unroll
            for (int i = 0; i < E; i++) v[i] = vn[DMA ? 0 : i];
        }
        // the next product's line is loaded while this one is transformed
        if (q + 1 < NPROD) fetch(q + 1);
        LineFFT<T, N, E, -1>::run(v, ldsc, t, twl);
        // coefficient of product q in each accumulated component, as a select on
        // kx (per point) and ky, kz (per line)
        const bool q0 = q == 0, q1 = q == 1, q2 = q == 2, q3 = q == 3, q4 = q == 4;
#pragma unroll
        for (int i = 0; i < E; i++) {
            const int n = t + i * TPL;
            const double kx = c.kg.kfx * (double)(n < c.kg.nx / 2 ? n : n - c.kg.nx);  // no shear here
            const cpx<double> a = cconv<double>(v[i]);
            if constexpr (GRP == 0) {
                // NonLinHydroAdv: k_j T_ij with T_zz = -T_xx - T_yy
                const double c0 = q0 ? kx : (q1 ? ky : (q2 ? kz : 0.0));
                const double c1 = q1 ? kx : (q3 ? ky : (q4 ? kz : 0.0));
                const double c2 = (q0 || q3) ? -kz : (q2 ? kx : (q4 ? ky : 0.0));
                acc[0][i] = cpx<double>{acc[0][i].x + c0 * a.x, acc[0][i].y + c0 * a.y};
                acc[1][i] = cpx<double>{acc[1][i].x + c1 * a.x, acc[1][i].y + c1 * a.y};
                acc[2][i] = cpx<double>{acc[2][i].x + c2 * a.x, acc[2][i].y + c2 * a.y};
            } else if constexpr (GRP == 1) {
                // MagneticShear: k x emf
                const double c0 = q1 ? -kz : (q2 ? ky : 0.0);
                const double c1 = q0 ? kz : (q2 ? -kx : 0.0);
                const double c2 = q0 ? -ky : (q1 ? kx : 0.0);
                acc[0][i] = cpx<double>{acc[0][i].x + c0 * a.x, acc[0][i].y + c0 * a.y};
                acc[1][i] = cpx<double>{acc[1][i].x + c1 * a.x, acc[1][i].y + c1 * a.y};
                acc[2][i] = cpx<double>{acc[2][i].x + c2 * a.x, acc[2][i].y + c2 * a.y};
            } else {
                // NonLinBoussinesqAdv: k.(u theta)
                const double c0 = q0 ? kx : (q1 ? ky : kz);
                acc[0][i] = cpx<double>{acc[0][i].x + c0 * a.x, acc[0][i].y + c0 * a.y};
            }
        }
    }

    // per point: sources, diffusion, projection, RK3 update.  The state of point
    // i + PD is loaded while point i is computed; the schedule barrier at the end
    // of each point keeps the compiler from hoisting all E points' loads at once.
    const double dt = ts->dt;
    const double gam = c.gam, ga = c.gam * dt;
    const double cu = fc.cu, cs = fc.cs;
    constexpr int PD = 1;
    // STRAT: the temperature group also loads u_s (component 1), the velocity
    // group theta's current state parked by the temperature group
    constexpr bool SX = STRAT && GRP == 2;
    constexpr bool SV = STRAT && GRP == 0;
    constexpr int NLD = NCOMP + (SX ? 1 : 0);
    cpx<double> su[E][NLD];
    cpx<T> sr[E][NLD];
    cpx<T> pk[SV ? E : 1];
    auto load_state = [&](int i) {
        const unsigned pi = ucol + (unsigned)(t + i * TPL) * ncols;
#pragma unroll
        for (int q = 0; q < NLD; q++) {
            const int f = q < NCOMP ? F0 + q : L::VX + c.sd;
            su[i][q] = (U + f * N0)[pi];
            if constexpr (MODE != 1) sr[i][q] = (R + f * N0)[pi];
        }
        if constexpr (SV) pk[i] = (W + L::VX * N0)[pi];
    };
#pragma unroll
    for (int i = 0; i < PD && i < E; i++) load_state(i);
#pragma unroll
    for (int i = 0; i < E; i++) {
        if (i + PD < E) load_state(i + PD);
        const int n = t + i * TPL;
        const unsigned pi = ucol + (unsigned)n * ncols;
        double kx, kyi, kzi, m;
        kvec_at(c.kg, n, j, kk, kx, kyi, kzi, m);
        const double k2 = kx * kx + ky * ky + kz * kz;
        // the stage's current state (diffusion) and the base of its update
        cpx<double> uc[NLD], base[NLD];
#pragma unroll
        for (int q = 0; q < NLD; q++) {
            if constexpr (MODE == 1) {
                uc[q] = su[i][q];
                base[q] = su[i][q];
            } else {
                const cpx<double> r = cconv<double>(sr[i][q]);
                uc[q] = MODE == 2 ? cpx<double>{su[i][q].x + cu * r.x, su[i][q].y + cu * r.y} : su[i][q];
                base[q] = cpx<double>{su[i][q].x + cs * r.x, su[i][q].y + cs * r.y};
            }
        }
        cpx<double> d[NCOMP];
        if constexpr (GRP == 0) {
#pragma unroll
            for (int q = 0; q < 3; q++) d[q] = mul_minus_im(m, acc[q][i]);
            if constexpr (SV) {
                // BoussinesqStrat: d u_s -= theta (current state)
                const cpx<double> th = cconv<double>(pk[i]);
#pragma unroll
                for (int q = 0; q < 3; q++)
                    if (q == c.sd) d[q] = cpx<double>{d[q].x - th.x, d[q].y - th.y};
            }
            const double f = -c.nu * k2;
#pragma unroll
            for (int q = 0; q < 3; q++) d[q] = diffuse(c.diff_mask, L::VX + q, d[q], uc[q], f);
            // GradPseudoPressure
            const cpx<double> sd3 = kdot3(kx, d[0], ky, d[1], kz, d[2]);
            const cpx<double> div = {-sd3.y, sd3.x};
            const double ik2 = k2 > 0.0 ? 1.0 / k2 : 1.0;
            const double ax = kx * ik2, ay = ky * ik2, az = kz * ik2;
            d[0] = cadd(d[0], cpx<double>{-ax * div.y, ax * div.x});
            d[1] = cadd(d[1], cpx<double>{-ay * div.y, ay * div.x});
            d[2] = cadd(d[2], cpx<double>{-az * div.y, az * div.x});
        } else if constexpr (GRP == 1) {
#pragma unroll
            for (int q = 0; q < 3; q++) d[q] = mul_plus_im(m, acc[q][i]);
            const double f = -c.nu_m * k2;
#pragma unroll
            for (int q = 0; q < 3; q++) d[q] = diffuse(c.diff_mask, L::BX + q, d[q], uc[q], f);
        } else {
            d[0] = mul_minus_im(m, acc[0][i]);
            if constexpr (SX) {
                // BoussinesqStrat: d theta += N2 u_s (current state); theta's current
                // state parked for the velocity group
                d[0] = cadd(d[0], cscale(uc[1], c.N2));
                (W + L::VX * N0)[pi] = cconv<T>(uc[0]);
            }
            d[0] = diffuse(c.diff_mask, L::TH, d[0], uc[0], -c.nu_th * k2);
        }
        cpx<double> un[NCOMP];
#pragma unroll
        for (int q = 0; q < NCOMP; q++) {
            if constexpr (MODE < 3) {
                // the increment in the bulk precision, and the update from that rounded value
                const cpx<T> rq = cconv<T>(cscale(d[q], dt));
                (R + (F0 + q) * N0)[pi] = rq;
                const cpx<double> r = cconv<double>(rq);
                un[q] = cpx<double>{base[q].x + gam * r.x, base[q].y + gam * r.y};
            } else {
                un[q] = cpx<double>{base[q].x + ga * d[q].x, base[q].y + ga * d[q].y};
            }
        }
        if constexpr (GRP < 2) {
            if (fc.project) {
                // CleanDivergence of the new u / B (the next step's CleanFieldDivergence)
                const cpx<double> sd3 = kdot3(kx, un[0], ky, un[1], kz, un[2]);
                const cpx<double> q0 = {-sd3.y, sd3.x};
                const double ik2 = k2 > 0.0 ? 1.0 / k2 : 1.0;
                const double kq[3] = {kx, ky, kz};
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const cpx<double> ikq = {-kq[q] * q0.y, kq[q] * q0.x};
                    un[q] = cpx<double>{un[q].x + ikq.x * ik2, un[q].y + ikq.y * ik2};
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NCOMP; q++) {
            if constexpr (MODE != 1) (U + (F0 + q) * N0)[pi] = un[q];  // stage 1 keeps U0
            acc[q][i] = un[q];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if (fc.inv) {
        // the next stage's inverse x pass of the updated fields (k_inv_x without derivatives)
#pragma unroll
        for (int q = 0; q < NCOMP; q++) {
            cpx<T> v[E];
#pragma unroll
            for (int i = 0; i < E; i++) v[i] = cconv<T>(cscale(acc[q][i], fc.scale));
            LineFFT<T, N, E, +1>::run(v, ldsc, t, twl);
            cpx<T> *dst = W + (long long)(F0 + q) * N0;
#pragma unroll
            for (int i = 0; i < E; i++) dst[ucol + (unsigned)(t + i * TPL) * ncols] = v[i];
        }
    }
}

// STRAT (Boussinesq layouts with stratification): the temperature group and
// then the velocity group in the first grid row's blocks, grid y = 1 + MHD
template <class T, int N, int LY, int MODE, int EF, bool BAT = false, bool STRAT = false>
__global__ __launch_bounds__(fx_tcols(N, EF) * fx_tpl(N, EF)) void k_fwdx_epilogue(
    const cpx<T> *P, cpx<double> *__restrict__ U, cpx<T> *__restrict__ S, cpx<T> *__restrict__ W,
    FusedConst fc, const cpx<T> *__restrict__ tw, const TimeState *__restrict__ ts, Batch bt) {
    using L = Layout<LY>;
    if constexpr (BAT) {
        P = rebase(bt, 0, P);
        U = rebase(bt, 1, U);
        S = rebase(bt, 2, S);
        W = rebase(bt, 3, W);
        ts = rebase(bt, 4, ts);
        const BoxRow *br = batch_row(bt);
        fc.e.nu = br->nu;
        fc.e.nu_m = br->nu_m;
        fc.e.nu_th = br->nu_th;
    }
    constexpr int E = fx_e(N, EF);
    constexpr int TPL = N / E;
    constexpr int TC = fx_tcols(N, EF);
    constexpr int LL = lds_len(N);
    constexpr int NT = TC * TPL;
    __shared__ cpx<T> lds[E != N ? TC * LL : 1];
    __shared__ cpx<T> twl[N];
    __shared__ cpx<T> pbuf[sizeof(cpx<T>) == 16 ? NT * E : 1];
    const long long ncols = (long long)fc.e.kg.ny * fc.e.kg.nzh;
    const int tid = threadIdx.x;
    const int cl = tid % TC, t = tid / TC;
    long long col = (long long)xcd_remap(blockIdx.x, gridDim.x) * TC + cl;
    if (col >= ncols) col = ncols - 1;
    cpx<T> *ldsc = lds + cl * LL;
    // group: 0 velocity (+ temperature with STRAT), then magnetic (MHD), then
    // temperature (boussinesq without STRAT)
    const int g = blockIdx.y;
    if (g == 0) {
        if constexpr (L::BOUSS && STRAT) {
            fx_group<T, N, LY, MODE, EF, 2, true>(P, U, S, W, fc, tw, twl, ldsc, pbuf, ts, tid, NT, t, col);
            // the velocity group's per-point values are computed anew, not kept
            // live across the temperature group
            int t2 = t;
            long long col2 = col;
            asm volatile("" : "+v"(t2));
            asm volatile("" : "+v"(col2));
            fx_group<T, N, LY, MODE, EF, 0, true>(P, U, S, W, fc, tw, twl, ldsc, pbuf, ts, tid, NT, t2, col2);
        } else {
            fx_group<T, N, LY, MODE, EF, 0>(P, U, S, W, fc, tw, twl, ldsc, pbuf, ts, tid, NT, t, col);
        }
    } else if (L::MHD && g == 1) {
        fx_group<T, N, LY, MODE, EF, 1>(P, U, S, W, fc, tw, twl, ldsc, pbuf, ts, tid, NT, t, col);
    } else if constexpr (L::BOUSS && !STRAT) {
        fx_group<T, N, LY, MODE, EF, 2>(P, U, S, W, fc, tw, twl, ldsc, pbuf, ts, tid, NT, t, col);
    }
}

}  // namespace spk
