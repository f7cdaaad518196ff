// spk_plane.hip -- the middle of an RHS evaluation on a quasi-2D grid (NZ = 2),
// one workgroup per x plane.
//
// With NZ = 2 a z row holds two complex values, the k_z = 0 and Nyquist modes,
// and its C2R keeps only their real parts:
//   f(z=0) = Re Z0 + Re Z1,   f(z=1) = Re Z0 - Re Z1,
// while the R2C of two real values p0, p1 is (p0 + p1, 0), (p0 - p1, 0).  These
// are the z rows of the general z kernel (spk_rows.hpp) at M = 1, operation for
// operation.  One x plane of one field is NY x 2 complex values (8 KB at
// NY = 512 in FP32), so a workgroup holds the plane of every input field in
// LDS: it makes the inverse y pass, the real-space products of the z kernel and
// the forward y pass of the products in one kernel, and neither the
// y-transformed fields nor the products before their y pass reach HBM.  Per RK3
// stage of a Boussinesq MHD box (7 fields, 11 products) this moves 18 field
// sweeps instead of 54 (inverse y 14, z kernel 18, forward y 22).
//
// Layout: wave w transforms one field's two z columns (lanes 2t + c: column c,
// points t + 32 i of the y line, so one load or store instruction covers 64
// consecutive complex values of the plane).  Region w of LDS holds that
// field's plane after the transform; the products at point y are made by thread
// y from every region and written back in place (point y of every region
// belongs to thread y), then wave q transforms product q.
//
// RKL2 variant (k_plane_rkl): the anisotropic flux of theta for one RKL2 stage,
// from A = IFFT_x(i k_x theta) and B = IFFT_x(theta) (two slots) to
// Y = FFT_y(F_x) and H = k_y FFT_y(F_y) + k_z FFT_y(F_z) (two slots), with B's
// real-space rows cached for the cycle; the same arithmetic as the fused RKL2
// y pass and z kernel (spk_rklfused.hip, k_zmid with ZP_B | ZP_G | ZP_BR).
#include "spk_internal.hpp"
#include "spk_kernels.hpp"

namespace spk {

// threads per y line: the two z columns of a field share one wavefront
constexpr int PL_TPL = 32;

// waves per workgroup of the RK3 kernel: one per input field, and up to 8 for
// the products (a program with more than 8 products transforms them in two rounds)
template <int PROG>
__host__ __device__ constexpr int plane_nw() {
    using Z = ZProg<PROG>;
    return Z::NIN > (Z::NOUT < 8 ? Z::NOUT : 8) ? Z::NIN : (Z::NOUT < 8 ? Z::NOUT : 8);
}

// wave-local transform of one line pair (WAVE ordering: the 64 lanes of the
// two lines are one wavefront), twiddles from the workgroup's LDS copy
template <class T, int NY, int DIR>
using PlaneFFT = LineFFT<T, NY, NY / PL_TPL, DIR, true, 0, PadSkew>;

template <class T, int NY, int PROG, bool BAT>
__global__ __launch_bounds__(64 * plane_nw<PROG>()) void k_plane(const cpx<T> *__restrict__ W, cpx<T> *__restrict__ P,
                                                               long long fs, const cpx<T> *__restrict__ twy,
                                                               ZConst zc, double *__restrict__ pmax,
                                                               const int *__restrict__ gate, Batch bt) {
    if constexpr (BAT) {
        W = rebase(bt, 0, W);
        P = rebase(bt, 1, P);
        pmax = rebase(bt, 2, pmax);
        gate = rebase(bt, 3, gate);
        zc.chi = 1.0 / batch_row(bt)->reynolds_ani;
    }
    if (gate != nullptr && *gate == 0) return;
    using Z = ZProg<PROG>;
    constexpr int NW = plane_nw<PROG>();
    constexpr int NT = 64 * NW;
    constexpr int E = NY / PL_TPL;
    constexpr int LL = lds_len(NY);
    constexpr int NKEEP = Z::NOUT > NW ? Z::NOUT - NW : 0;
    static_assert(Z::NIN <= NW && Z::NOUT <= 2 * NW, "one round of inputs, at most two of products");
    static_assert(NKEEP == 0 || NT >= NY, "products kept for a second round need one point per thread");
    static_assert(E >= 2, "NY >= 64");
    __shared__ cpx<T> reg[NW * 2 * LL];
    __shared__ cpx<T> twl[NY];
    __shared__ double smax[6][NW];

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int c = lane & 1, t = lane >> 1;
    const long long plane = (long long)blockIdx.x * NY * 2;
    cpx<T> *R = reg + (2 * w + c) * LL;

    // inverse y pass: wave w transforms input field w
    cpx<T> v[E];
    if (w < Z::NIN) {
        const cpx<T> *src = W + (long long)w * fs + plane + c;
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = src[2 * (t + i * PL_TPL)];
    }
    for (int e = tid; e < NY; e += NT) twl[e] = twy[e];
    __syncthreads();
    if (w < Z::NIN) {
        PlaneFFT<T, NY, +1>::run(v, R, t, twl);
        lsync<true>();
#pragma unroll
        for (int i = 0; i < E; i++) R[PadSkew::at(t + i * PL_TPL)] = v[i];
    }
    __syncthreads();

    // real-space products at the plane's points (thread y: both z points of y)
    T keep[NKEEP > 0 ? NKEEP : 1][2];
    double mx[6];
#pragma unroll
    for (int k = 0; k < 6; k++) mx[k] = 0.0;
    for (int y = tid; y < NY; y += NT) {
        const int ay = PadSkew::at(y);
        T a0[Z::NIN], a1[Z::NIN];
#pragma unroll
        for (int f = 0; f < Z::NIN; f++) {
            const T z0 = reg[(2 * f) * LL + ay].x, z1 = reg[(2 * f + 1) * LL + ay].x;
            a0[f] = z0 + z1;
            a1[f] = z0 - z1;
        }
        if constexpr (Z::U) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                mx[k] = fmax(mx[k], fmax(fabs((double)a0[Z::IU + k]), fabs((double)a1[Z::IU + k])));
                if constexpr (Z::B)
                    mx[3 + k] = fmax(mx[3 + k], fmax(fabs((double)a0[Z::IB + k]), fabs((double)a1[Z::IB + k])));
            }
        }
#pragma unroll
        for (int q = 0; q < Z::NOUT; q++) {
            const T p0 = zproduct<PROG>(q, a0, zc), p1 = zproduct<PROG>(q, a1, zc);
            const T s = p0 + p1, d = p0 - p1;
            if (q < NW) {
                reg[(2 * q) * LL + ay] = cpx<T>{s, (T)0};
                reg[(2 * q + 1) * LL + ay] = cpx<T>{d, (T)0};
            } else {
                keep[q - NW][0] = s;
                keep[q - NW][1] = d;
            }
        }
    }
    if constexpr (Z::U) {
        if (pmax != nullptr) {
            // max |u_i|, |B_i| over the plane (compute_dt.cu:90-106)
#pragma unroll
            for (int k = 0; k < 6; k++) {
                double m = mx[k];
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
                if (lane == 0) smax[k][w] = m;
            }
        }
    }
    __syncthreads();
    if constexpr (Z::U) {
        if (pmax != nullptr && tid < 6) {
            double m = 0.0;
#pragma unroll
            for (int k = 0; k < NW; k++) m = fmax(m, smax[tid][k]);
            pmax[(long long)blockIdx.x * 6 + tid] = m;
        }
    }

    // forward y pass: wave w transforms product w, then (second round) NW + w
#pragma unroll
    for (int r = 0; r < (NKEEP > 0 ? 2 : 1); r++) {
        if (r == 1) {
            __syncthreads();  // every wave is done with its region
            if (tid < NY) {
                const int ay = PadSkew::at(tid);
#pragma unroll
                for (int k = 0; k < NKEEP; k++) {
                    reg[(2 * k) * LL + ay] = cpx<T>{keep[k][0], (T)0};
                    reg[(2 * k + 1) * LL + ay] = cpx<T>{keep[k][1], (T)0};
                }
            }
            __syncthreads();
        }
        const int q = r * NW + w;
        if (q < Z::NOUT) {
#pragma unroll
            for (int i = 0; i < E; i++) v[i] = R[PadSkew::at(t + i * PL_TPL)];
            lsync<true>();
            PlaneFFT<T, NY, -1>::run(v, R, t, twl);
            cpx<T> *dst = P + (long long)q * fs + plane + c;
#pragma unroll
            for (int i = 0; i < E; i++) dst[2 * (t + i * PL_TPL)] = v[i];
        }
    }
}

// One RKL2 stage's flux, three waves: inverse y of A, of i k_y B and of B
// (times i k_z after it); products; forward y of F_x, F_y and F_z, the last two
// combined into H.  G: slots A, B; Br: B's real-space rows (3 slots); Q: Y, H.
template <class T, int NY, bool BAT>
__global__ __launch_bounds__(192) void k_plane_rkl(const cpx<T> *__restrict__ G, const cpx<T> *__restrict__ Br,
                                                  cpx<T> *__restrict__ Q, long long fs, KGeom kg,
                                                  const cpx<T> *__restrict__ twy, ZConst zc,
                                                  const int *__restrict__ gate, Batch bt) {
    if constexpr (BAT) {
        G = rebase(bt, 0, G);
        Br = rebase(bt, 1, Br);
        Q = rebase(bt, 2, Q);
        gate = rebase(bt, 3, gate);
        zc.chi = 1.0 / batch_row(bt)->reynolds_ani;
    }
    if (gate != nullptr && *gate == 0) return;
    constexpr int PROG = ZP_B | ZP_G | ZP_BR;
    using Z = ZProg<PROG>;
    static_assert(Z::NIN == 6 && Z::NOUT == 3 && Z::IB == 0 && Z::IG == 3, "flux program layout");
    constexpr int NT = 192;
    constexpr int E = NY / PL_TPL;
    constexpr int LL = lds_len(NY);
    static_assert(E >= 2, "NY >= 64");
    __shared__ cpx<T> reg[3 * 2 * LL];
    __shared__ cpx<T> twl[NY];

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int c = lane & 1, t = lane >> 1;
    const long long plane = (long long)blockIdx.x * NY * 2;
    cpx<T> *R = reg + (2 * w + c) * LL;
    const T kz = (T)(kg.kfz * (double)c);
    auto ky_at = [&](int n) { return (T)(kg.kfy * (double)(n < kg.ny / 2 ? n : n - kg.ny)); };

    // inverse y: wave 0 A, wave 1 i k_y B, wave 2 B (then i k_z)
    cpx<T> v[E];
    {
        const cpx<T> *src = G + (w == 0 ? 0 : fs) + plane + c;
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = src[2 * (t + i * PL_TPL)];
    }
    for (int e = tid; e < NY; e += NT) twl[e] = twy[e];
    __syncthreads();
    if (w == 1) {
#pragma unroll
        for (int i = 0; i < E; i++) {
            const T ky = ky_at(t + i * PL_TPL);
            v[i] = cpx<T>{-(ky * v[i].y), ky * v[i].x};
        }
    }
    PlaneFFT<T, NY, +1>::run(v, R, t, twl);
    if (w == 2) {
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = cpx<T>{-(kz * v[i].y), kz * v[i].x};
    }
    lsync<true>();
#pragma unroll
    for (int i = 0; i < E; i++) R[PadSkew::at(t + i * PL_TPL)] = v[i];
    __syncthreads();

    // flux at the plane's points: B from the cached rows, grad theta from LDS
    for (int y = tid; y < NY; y += NT) {
        const int ay = PadSkew::at(y);
        const long long row = (long long)blockIdx.x * NY + y;
        T a0[Z::NIN], a1[Z::NIN];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const cpx<T> b = Br[(long long)k * fs + 2 * row];
            a0[Z::IB + k] = b.x;
            a1[Z::IB + k] = b.y;
            const T z0 = reg[(2 * k) * LL + ay].x, z1 = reg[(2 * k + 1) * LL + ay].x;
            a0[Z::IG + k] = z0 + z1;
            a1[Z::IG + k] = z0 - z1;
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const T p0 = zproduct<PROG>(q, a0, zc), p1 = zproduct<PROG>(q, a1, zc);
            reg[(2 * q) * LL + ay] = cpx<T>{p0 + p1, (T)0};
            reg[(2 * q + 1) * LL + ay] = cpx<T>{p0 - p1, (T)0};
        }
    }
    __syncthreads();

    // forward y: wave 0 Y = FFT_y F_x; waves 1 and 2 transform F_y and F_z, wave 2
    // leaves k_z FFT_y F_z in its region and wave 1 adds k_y FFT_y F_y to it: H
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = R[PadSkew::at(t + i * PL_TPL)];
    lsync<true>();
    PlaneFFT<T, NY, -1>::run(v, R, t, twl);
    if (w == 0) {
        cpx<T> *dst = Q + plane + c;
#pragma unroll
        for (int i = 0; i < E; i++) dst[2 * (t + i * PL_TPL)] = v[i];
    } else if (w == 2) {
        lsync<true>();
#pragma unroll
        for (int i = 0; i < E; i++) R[PadSkew::at(t + i * PL_TPL)] = cpx<T>{kz * v[i].x, kz * v[i].y};
    }
    __syncthreads();
    if (w == 1) {
        const cpx<T> *Rz = reg + (4 + c) * LL;
        cpx<T> *dst = Q + fs + plane + c;
#pragma unroll
        for (int i = 0; i < E; i++) {
            const T ky = ky_at(t + i * PL_TPL);
            const cpx<T> z = Rz[PadSkew::at(t + i * PL_TPL)];
            dst[2 * (t + i * PL_TPL)] = cpx<T>{ky * v[i].x + z.x, ky * v[i].y + z.y};
        }
    }
}

// ---- launchers ----

bool plane_supported(const Grid3 &g, int prog) {
    if (g.nz != 2 || !(g.ny == 64 || g.ny == 128 || g.ny == 256 || g.ny == 512)) return false;
    return prog == ZP_U || prog == (ZP_U | ZP_B) || prog == (ZP_U | ZP_TH) || prog == (ZP_U | ZP_B | ZP_TH);
}

bool plane_rkl_supported(const Grid3 &g) {
    return g.nz == 2 && (g.ny == 64 || g.ny == 128 || g.ny == 256 || g.ny == 512);
}

template <class T, int NY, bool BAT>
static int plane_prog(int prog, spk_ctx ctx, const Grid3 &g, const cpx<T> *tw, const cpx<T> *W, cpx<T> *P,
                      const ZConst &zc, double *pmax, const Batch &bt) {
    const dim3 grid((unsigned)g.nx, 1, nboxes(ctx));
#define SPK_PL(PR)                                                                                      \
    case PR:                                                                                            \
        k_plane<T, NY, PR, BAT><<<grid, 64 * plane_nw<PR>(), 0, ctx->stream>>>(W, P, g.ntc, tw, zc, pmax, \
                                                                                ctx->gate, bt);         \
        break;
    switch (prog) {
        SPK_PL(ZP_U)
        SPK_PL(ZP_U | ZP_B)
        SPK_PL(ZP_U | ZP_TH)
        SPK_PL(ZP_U | ZP_B | ZP_TH)
        default: set_error("plane kernel: program %d not instantiated", prog); return SPK_ERR_STATE;
    }
#undef SPK_PL
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

template <class T, bool BAT>
static int plane_t(int prog, spk_ctx ctx, const Grid3 &g, const cpx<T> *tw, const cpx<T> *W, cpx<T> *P,
                   const ZConst &zc, double *pmax, const Batch &bt) {
    switch (g.ny) {
        case 64: return plane_prog<T, 64, BAT>(prog, ctx, g, tw, W, P, zc, pmax, bt);
        case 128: return plane_prog<T, 128, BAT>(prog, ctx, g, tw, W, P, zc, pmax, bt);
        case 256: return plane_prog<T, 256, BAT>(prog, ctx, g, tw, W, P, zc, pmax, bt);
        case 512: return plane_prog<T, 512, BAT>(prog, ctx, g, tw, W, P, zc, pmax, bt);
        default: set_error("plane kernel: NY=%d not supported", g.ny); return SPK_ERR_ARG;
    }
}

int launch_plane(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int prog, const void *W, void *P,
                 const ZConst &zc, double *pmax) {
    SPK_REQUIRE(plane_supported(g, prog), SPK_ERR_STATE, "plane kernel: grid %dx%dx%d, program %d not supported",
                g.nx, g.ny, g.nz, prog);
    Batch bt;
    SPK_TRY(make_batch(ctx, {W, P, pmax, ctx->gate}, &bt));
    if (tw.prec == SPK_F64) {
        const cpx<double> *t = reinterpret_cast<const cpx<double> *>(tw.d) + tw.off_y;
        const cpx<double> *w = (const cpx<double> *)W;
        cpx<double> *p = (cpx<double> *)P;
        return bt.rows ? plane_t<double, true>(prog, ctx, g, t, w, p, zc, pmax, bt)
                       : plane_t<double, false>(prog, ctx, g, t, w, p, zc, pmax, bt);
    }
    const cpx<float> *t = reinterpret_cast<const cpx<float> *>(tw.d) + tw.off_y;
    const cpx<float> *w = (const cpx<float> *)W;
    cpx<float> *p = (cpx<float> *)P;
    return bt.rows ? plane_t<float, true>(prog, ctx, g, t, w, p, zc, pmax, bt)
                   : plane_t<float, false>(prog, ctx, g, t, w, p, zc, pmax, bt);
}

template <class T, bool BAT>
static int plane_rkl_t(spk_ctx ctx, const Grid3 &g, const cpx<T> *tw, const cpx<T> *G, const cpx<T> *Br,
                       cpx<T> *Q, const KGeom &kg, const ZConst &zc, const Batch &bt) {
    const dim3 grid((unsigned)g.nx, 1, nboxes(ctx));
    switch (g.ny) {
        case 64: k_plane_rkl<T, 64, BAT><<<grid, 192, 0, ctx->stream>>>(G, Br, Q, g.ntc, kg, tw, zc, ctx->gate, bt); break;
        case 128: k_plane_rkl<T, 128, BAT><<<grid, 192, 0, ctx->stream>>>(G, Br, Q, g.ntc, kg, tw, zc, ctx->gate, bt); break;
        case 256: k_plane_rkl<T, 256, BAT><<<grid, 192, 0, ctx->stream>>>(G, Br, Q, g.ntc, kg, tw, zc, ctx->gate, bt); break;
        case 512: k_plane_rkl<T, 512, BAT><<<grid, 192, 0, ctx->stream>>>(G, Br, Q, g.ntc, kg, tw, zc, ctx->gate, bt); break;
        default: set_error("plane RKL2 kernel: NY=%d not supported", g.ny); return SPK_ERR_ARG;
    }
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int launch_plane_rkl(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, const void *G, const void *Breal, void *Q,
                     const KGeom &kg, const ZConst &zc) {
    SPK_REQUIRE(plane_rkl_supported(g), SPK_ERR_STATE, "plane RKL2 kernel: grid %dx%dx%d not supported", g.nx, g.ny,
                g.nz);
    SPK_REQUIRE(kg.kxshift == 0.0, SPK_ERR_STATE, "plane RKL2 kernel: sheared wavevectors are not supported");
    Batch bt;
    SPK_TRY(make_batch(ctx, {G, Breal, Q, ctx->gate}, &bt));
    if (tw.prec == SPK_F64) {
        const cpx<double> *t = reinterpret_cast<const cpx<double> *>(tw.d) + tw.off_y;
        const cpx<double> *gg = (const cpx<double> *)G, *br = (const cpx<double> *)Breal;
        cpx<double> *q = (cpx<double> *)Q;
        return bt.rows ? plane_rkl_t<double, true>(ctx, g, t, gg, br, q, kg, zc, bt)
                       : plane_rkl_t<double, false>(ctx, g, t, gg, br, q, kg, zc, bt);
    }
    const cpx<float> *t = reinterpret_cast<const cpx<float> *>(tw.d) + tw.off_y;
    const cpx<float> *gg = (const cpx<float> *)G, *br = (const cpx<float> *)Breal;
    cpx<float> *q = (cpx<float> *)Q;
    return bt.rows ? plane_rkl_t<float, true>(ctx, g, t, gg, br, q, kg, zc, bt)
                   : plane_rkl_t<float, false>(ctx, g, t, gg, br, q, kg, zc, bt);
}

}  // namespace spk
