// spk_rklfused.hip -- one RKL2 stage of anisotropic conduction as a single
// sweep: forward x pass of the heat-flux products + flux divergence + the RKL2
// stage update + the next stage's inverse x pass of i k theta.
//
// An RKL2 stage of the anisotropic conduction of theta (rkl.cu:496-585 with
// ParabolicTermsSTS, physics.cu:233-330, and AnisotropicHeatFlux,
// anisotropic_conduction.cu:10-62) is, per stage: grad theta to real space, the
// flux chi b (b . grad theta + OmegaT2 b_s) / B^2 there, its forward transform,
// i mask k . flux, and the rkl2_stage update (cuda_kernels_generic.cu:160-172).
//
// Derivatives are applied in the pass where their wavenumber is the line's own
// or a constant, so fewer arrays cross HBM than the 3 + 3 components:
//   inverse: this kernel's x pass makes A = IFFT_x(i k_x theta) and
//            B = IFFT_x(theta); the y pass (k_pass_y_rkl) makes IFFT_y(A),
//            IFFT_y(i k_y B) and i k_z IFFT_y(B), the three gradient components
//            before their z pass (which the z kernel does with the flux);
//   forward: after the z kernel's R2C of the flux F, the y pass makes
//            Y = FFT_y(F_x) and H = k_y FFT_y(F_y) + k_z FFT_y(F_z); this
//            kernel's x pass gives k . F = k_x FFT_x(Y) + FFT_x(H).
// One block owns TC columns of whole x lines: it transforms Y and H along x in
// registers and LDS, applies the divergence and the update point by point, and
// transforms A and B of the new theta back along x.  The next stage starts at
// its inverse y pass.
#include "spk_internal.hpp"
#include "spk_kernels.hpp"

namespace spk {

// points per thread of the x line (16 in FP32 would widen the column tiles, but
// its registers spill)
constexpr int RKLF_E = 8;

template <class T, int N>
__global__ __launch_bounds__(fx_tcols(N, RKLF_E) * fx_tpl(N, RKLF_E)) void k_fwdx_rkl2(
    const cpx<T> *Q, cpx<double> *__restrict__ U, cpx<double> *__restrict__ Uc0, cpx<double> *__restrict__ Uc1,
    cpx<double> *__restrict__ dU0, cpx<T> *__restrict__ W, RklConst c, int stage, double mu, double nu,
    double dtmut, double gdt, int inv, double scale, const cpx<T> *__restrict__ tw,
    const RklPlan *__restrict__ plan, const TimeState *__restrict__ ts, Batch bt) {
    Q = rebase(bt, 0, Q);
    U = rebase(bt, 1, U);
    Uc0 = rebase(bt, 2, Uc0);
    Uc1 = rebase(bt, 3, Uc1);
    dU0 = rebase(bt, 4, dU0);
    W = rebase(bt, 5, W);
    plan = rebase(bt, 6, plan);
    ts = rebase(bt, 7, ts);
    if (plan != nullptr) {
        // device-planned cycle: stages past this box's count return at once; the
        // last active stage writes no next-stage input
        if (!plan->active[stage]) return;
        const double *cf = plan->coef + 4 * (stage - 1);
        const double dt = ts->dt;
        mu = cf[0];
        nu = cf[1];
        dtmut = cf[2] * dt;
        gdt = cf[3] * dt;
        inv = inv && plan->active[stage + 1];
    }
    constexpr int E = fx_e(N, RKLF_E);
    constexpr int TPL = N / E;
    constexpr int TC = fx_tcols(N, RKLF_E);
    constexpr int LL = lds_len(N);
    constexpr int NT = TC * TPL;
    __shared__ cpx<T> lds[E != N ? TC * LL : 1];
    __shared__ cpx<T> twl[N];
    const long long ncols = (long long)c.kg.ny * c.kg.nzh;
    const int tid = threadIdx.x;
    const int cl = tid % TC, t = tid / TC;
    const long long col0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * TC + cl;
    // a lane past the last column transforms the last column (every lane takes
    // part in the transforms' barriers) and stores nothing
    const bool live = col0 < ncols;
    const long long col = live ? col0 : ncols - 1;
    cpx<T> *ldsc = lds + cl * LL;
    const long long N0 = c.ntc;
    const int j = (int)(col / c.kg.nzh);
    const int kk = (int)(col - (long long)j * c.kg.nzh);
    const long long F = (long long)c.field[0] * N0;

    for (int e = tid; e < N; e += NT) twl[e] = tw[e];
    __syncthreads();

    // k . flux = k_x FFT_x(Y) + FFT_x(H), accumulated after each line FFT (the
    // second line is loaded while the first is transformed)
    cpx<double> acc[E];
#pragma unroll
    for (int i = 0; i < E; i++) acc[i] = cpx<double>{0.0, 0.0};
    cpx<T> vn[E];
#pragma unroll
    for (int i = 0; i < E; i++) vn[i] = Q[col + (long long)(t + i * TPL) * ncols];
#pragma unroll 1
    for (int q = 0; q < 2; q++) {
        cpx<T> v[E];
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = vn[i];
        if (q + 1 < 2) {
#pragma unroll
            for (int i = 0; i < E; i++) vn[i] = Q[(q + 1) * N0 + col + (long long)(t + i * TPL) * ncols];
        }
        LineFFT<T, N, E, -1>::run(v, ldsc, t, twl);
#pragma unroll
        for (int i = 0; i < E; i++) {
            const int n = t + i * TPL;
            const double kq = q == 0 ? c.kg.kfx * (double)(n < c.kg.nx / 2 ? n : n - c.kg.nx) : 1.0;
            const cpx<double> a = cconv<double>(v[i]);
            acc[i] = cpx<double>{acc[i].x + kq * a.x, acc[i].y + kq * a.y};
        }
    }

    // per point: dtheta = i mask k . flux (DivergenceMask), then rkl2_stage; the
    // new theta replaces the accumulator
    const double w = 1.0 - mu - nu;
#pragma unroll
    for (int i = 0; i < E; i++) {
        const int n = t + i * TPL;
        const long long pi = col + (long long)n * ncols;
        double kx, kyi, kzi, m;
        kvec_at(c.kg, n, j, kk, kx, kyi, kzi, m);
        const cpx<double> du = mul_plus_im(m, acc[i]);
        const cpx<double> uc = U[F + pi];
        cpx<double> un;
        if (stage == 1) {
            if (live) {
                Uc0[pi] = uc;
                dU0[pi] = du;
                Uc1[pi] = uc;
            }
            un = cpx<double>{uc.x + dtmut * du.x, uc.y + dtmut * du.y};
        } else {
            const cpx<double> u1 = Uc1[pi], u0 = Uc0[pi], d0 = dU0[pi];
            const cpx<double> Y = {mu * uc.x + nu * u1.x, mu * uc.y + nu * u1.y};
            if (live) Uc1[pi] = uc;
            un = cpx<double>{Y.x + w * u0.x + dtmut * du.x + gdt * d0.x, Y.y + w * u0.y + dtmut * du.y + gdt * d0.y};
        }
        if (live) U[F + pi] = un;
        acc[i] = un;
        // one point's state loads at a time: hoisting all E points' loads would
        // spill registers
        __builtin_amdgcn_sched_barrier(0);
    }

    if (inv) {
        // A = i k_x theta and B = theta of the new theta, 1/N scaled, inverse x
        // pass into the next stage's first two slots
#pragma unroll 1
        for (int d = 0; d < 2; d++) {
            cpx<T> v[E];
#pragma unroll
            for (int i = 0; i < E; i++) {
                const int n = t + i * TPL;
                const double kd = c.kg.kfx * (double)(n < c.kg.nx / 2 ? n : n - c.kg.nx);
                const cpx<double> a = d == 0 ? cpx<double>{-(kd * acc[i].y), kd * acc[i].x} : acc[i];
                v[i] = cconv<T>(cscale(a, scale));
            }
            LineFFT<T, N, E, +1>::run(v, ldsc, t, twl);
            if (live) {
                cpx<T> *dst = W + d * N0 + col;
#pragma unroll
                for (int i = 0; i < E; i++) dst[(long long)(t + i * TPL) * ncols] = v[i];
            }
        }
    }
}

bool rkl_fused_supported_nx(int nx) { return nx >= 32 && nx <= 1024 && is_pow2(nx); }

// The y pass of the fused RKL2 stage over slots buf + q fs, q = 0, 1, 2 (slot
// layout of a y pass: NX planes of NY lines by NZH columns; a column's index is
// its z wavenumber index).  Grid row 0 transforms slot 0 in place; grid row 1:
//   DIR = +1: slot 1 (B) -> slot 1 = IFFT_y(i k_y B), slot 2 = i k_z IFFT_y(B);
//   DIR = -1: slots 1, 2 -> slot 1 = k_y FFT_y(slot 1) + k_z FFT_y(slot 2).
// 8 points per thread: a block row 1 holds two lines per thread
__host__ __device__ constexpr int yrkl_e(int n) { return line_e(n, 8); }
__host__ __device__ constexpr int yrkl_tpl(int n) { return n / yrkl_e(n); }
// 256-byte segments per line point at n = 256: FP64 in 512-thread blocks (two
// lines of 8 complex doubles per thread need ~180 VGPRs), FP32 in 1024
template <class T>
__host__ __device__ constexpr int yrkl_tcols(int n) {
    return yrkl_tpl(n) >= 32 ? (sizeof(T) == 8 ? 512 : 1024) / yrkl_tpl(n) : 32;
}

template <class T, int N, int DIR, int TC>
__global__ __launch_bounds__(TC * yrkl_tpl(N)) void k_pass_y_rkl(cpx<T> *buf, long long fs, long long ncols,
                                                                   long long ostride, int nouter, KGeom kg,
                                                                   const cpx<T> *__restrict__ tw,
                                                                   const int *__restrict__ gate, Batch bt) {
    buf = rebase(bt, 0, buf);
    gate = rebase(bt, 1, gate);
    if (gate != nullptr && *gate == 0) return;
    constexpr int E = yrkl_e(N);
    constexpr int TPL = N / E;
    constexpr int LL = lds_len(N);
    __shared__ cpx<T> lds[E != N ? TC * LL : 1];

    const int tid = threadIdx.x;
    const int c = tid % TC, t = tid / TC;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const long long flat = (long long)b * TC + c;
    const long long o = flat / ncols;
    const long long col = flat - o * ncols;
    const bool active = o < nouter;
    cpx<T> *s0 = buf + (long long)blockIdx.y * fs + o * ostride + col;
    cpx<T> *ld = lds + c * LL;

    cpx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = active ? s0[(long long)(t + i * TPL) * ncols] : cpx<T>{0, 0};
    if (blockIdx.y == 0) {
        LineFFT<T, N, E, DIR>::run(v, ld, t, tw);
        if (active) {
#pragma unroll
            for (int i = 0; i < E; i++) s0[(long long)(t + i * TPL) * ncols] = v[i];
        }
        return;
    }
    const T kz = (T)(kg.kfz * (double)col);
    auto ky_at = [&](int n) { return (T)(kg.kfy * (double)(n < kg.ny / 2 ? n : n - kg.ny)); };
    cpx<T> w[E];
    if constexpr (DIR > 0) {
#pragma unroll
        for (int i = 0; i < E; i++) {
            const T ky = ky_at(t + i * TPL);
            w[i] = cpx<T>{-(ky * v[i].y), ky * v[i].x};
        }
        LineFFT<T, N, E, DIR>::run(w, ld, t, tw);
        LineFFT<T, N, E, DIR>::run(v, ld, t, tw);
        if (active) {
#pragma unroll
            for (int i = 0; i < E; i++) {
                s0[(long long)(t + i * TPL) * ncols] = w[i];
                s0[fs + (long long)(t + i * TPL) * ncols] = cpx<T>{-(kz * v[i].y), kz * v[i].x};
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < E; i++) w[i] = active ? s0[fs + (long long)(t + i * TPL) * ncols] : cpx<T>{0, 0};
        LineFFT<T, N, E, DIR>::run(v, ld, t, tw);
        LineFFT<T, N, E, DIR>::run(w, ld, t, tw);
        if (active) {
#pragma unroll
            for (int i = 0; i < E; i++) {
                const T ky = ky_at(t + i * TPL);
                s0[(long long)(t + i * TPL) * ncols] =
                    cpx<T>{ky * v[i].x + kz * w[i].x, ky * v[i].y + kz * w[i].y};
            }
        }
    }
}

template <class T, int DIR>
static int launch_y_rkl_t(spk_ctx ctx, const Grid3 &g, const cpx<T> *tw, void *buf, const KGeom &kg) {
    Batch bt;
    SPK_TRY(make_batch(ctx, {buf, ctx->gate}, &bt));
    const unsigned nbz = nboxes(ctx);
    const long long ncols = g.nzh, ostride = (long long)g.ny * g.nzh, fs = g.ntc;
    cpx<T> *p = (cpx<T> *)buf;
#define SPK_YRKL_CASE(NN)                                                                                  \
    case NN: {                                                                                             \
        constexpr int TC = yrkl_tcols<T>(NN);                                                              \
        if constexpr (TC > 2) {                                                                            \
            if (ncols == 2) {                                                                              \
                k_pass_y_rkl<T, NN, DIR, 2><<<dim3((unsigned)g.nx, 2, nbz), 2 * yrkl_tpl(NN), 0,          \
                                              ctx->stream>>>(p, fs, ncols, ostride, g.nx, kg, tw, ctx->gate, bt); \
                break;                                                                                     \
            }                                                                                              \
        }                                                                                                  \
        const long long ntiles = (ncols * g.nx + TC - 1) / TC;                                             \
        k_pass_y_rkl<T, NN, DIR, TC><<<dim3((unsigned)ntiles, 2, nbz), TC * yrkl_tpl(NN), 0, ctx->stream>>>( \
            p, fs, ncols, ostride, g.nx, kg, tw, ctx->gate, bt);                                           \
        break;                                                                                             \
    }
    switch (g.ny) {
        SPK_YRKL_CASE(2) SPK_YRKL_CASE(4) SPK_YRKL_CASE(8) SPK_YRKL_CASE(16) SPK_YRKL_CASE(32) SPK_YRKL_CASE(64)
        SPK_YRKL_CASE(128) SPK_YRKL_CASE(256) SPK_YRKL_CASE(512) SPK_YRKL_CASE(1024) SPK_YRKL_CASE(2048)
        default: set_error("fused RKL2 y pass: NY=%d not supported (power of two 2..2048)", g.ny); return SPK_ERR_ARG;
    }
#undef SPK_YRKL_CASE
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int launch_pass_y_rkl(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int dir, void *buf, const KGeom &kg) {
    SPK_REQUIRE(kg.kxshift == 0.0, SPK_ERR_STATE, "fused RKL2 y pass: sheared wavevectors are not supported");
    if (tw.prec == SPK_F64) {
        const cpx<double> *t = reinterpret_cast<const cpx<double> *>(tw.d) + tw.off_y;
        return dir > 0 ? launch_y_rkl_t<double, +1>(ctx, g, t, buf, kg) : launch_y_rkl_t<double, -1>(ctx, g, t, buf, kg);
    }
    const cpx<float> *t = reinterpret_cast<const cpx<float> *>(tw.d) + tw.off_y;
    return dir > 0 ? launch_y_rkl_t<float, +1>(ctx, g, t, buf, kg) : launch_y_rkl_t<float, -1>(ctx, g, t, buf, kg);
}

template <class T>
static int launch_t(spk_ctx ctx, int nx, const void *Q, void *U, void *Uc0, void *Uc1, void *dU0, void *W,
                    const RklConst &c, int stage, double mu, double nu, double dtmut, double gdt, int inv, double scale,
                    const void *tw, const RklPlan *plan, const TimeState *ts) {
    Batch bt;
    SPK_TRY(make_batch(ctx, {Q, U, Uc0, Uc1, dU0, W, plan, ts}, &bt));
    const long long ncols = (long long)c.kg.ny * c.kg.nzh;
    const cpx<T> *q = (const cpx<T> *)Q;
    cpx<double> *u = (cpx<double> *)U, *u0 = (cpx<double> *)Uc0, *u1 = (cpx<double> *)Uc1, *d0 = (cpx<double> *)dU0;
    cpx<T> *w = (cpx<T> *)W;
    const cpx<T> *t = (const cpx<T> *)tw;
#define SPK_RKLF_CASE(NN)                                                                                      \
    case NN: {                                                                                                 \
        constexpr int TC = fx_tcols(NN, RKLF_E);                                                          \
        const dim3 grid((unsigned)((ncols + TC - 1) / TC), 1, nboxes(ctx));                                    \
        k_fwdx_rkl2<T, NN><<<grid, TC * fx_tpl(NN, RKLF_E), 0, ctx->stream>>>(q, u, u0, u1, d0, w, c, stage, mu, nu, \
                                                                          dtmut, gdt, inv, scale, t, plan, ts, bt); \
        break;                                                                                                 \
    }
    switch (nx) {
        SPK_RKLF_CASE(32) SPK_RKLF_CASE(64) SPK_RKLF_CASE(128) SPK_RKLF_CASE(256) SPK_RKLF_CASE(512)
        SPK_RKLF_CASE(1024)
        default: set_error("fused RKL2 stage: NX=%d not instantiated", nx); return SPK_ERR_ARG;
    }
#undef SPK_RKLF_CASE
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int launch_fwdx_rkl2(spk_ctx ctx, int bulk, int nx, const void *Q, void *U, void *Uc0, void *Uc1, void *dU0, void *W,
                     const RklConst &c, int stage, double mu, double nu, double dtmut, double gdt, int inv,
                     double scale, const void *tw, const RklPlan *plan, const TimeState *ts) {
    if (bulk == SPK_F64)
        return launch_t<double>(ctx, nx, Q, U, Uc0, Uc1, dU0, W, c, stage, mu, nu, dtmut, gdt, inv, scale, tw, plan, ts);
    return launch_t<float>(ctx, nx, Q, U, Uc0, Uc1, dU0, W, c, stage, mu, nu, dtmut, gdt, inv, scale, tw, plan, ts);
}

}  // namespace spk
