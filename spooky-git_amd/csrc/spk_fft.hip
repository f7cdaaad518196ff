// spk_fft.hip -- 3D R2C/C2R FFT for gfx950: one kernel per direction pass.
//
// Forward transform = z pass (real rows -> half-complex rows, via a complex FFT
// of length NZ/2 and the split-radix real post-processing step) -> y pass ->
// x pass.  Inverse = x pass (with the 1/(NX NY NZ) scale applied to the input,
// as scaleKernel does in cufft_routines.cu:30) -> y pass -> z pass.
//
// x and y passes: a workgroup owns TCOLS adjacent columns (contiguous in
// memory) and the whole line along the transformed axis, so every global load
// and store moves TCOLS*sizeof(complex) contiguous bytes per line point.
// z passes: a workgroup owns RW whole rows.
#include "spk_internal.hpp"
#include "spk_rows.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace spk {

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------


// Line FFT over columns: element (f, o, n, c) at f*fstride + o*ostride + n*lstride + c.
// AXIS (0 = x, 1 = y) only names the instantiation so profiles tell the passes apart.
// TC columns per block, taken from the flattened (o, c) index, so a tile may
// span two planes: the y pass of a 256^3 grid has 129 columns per x-plane, and
// tiles per plane would leave one block in five with a single live column.
// The y pass of a quasi-2D grid (NZ = 2: two columns of adjacent complex values
// per x) uses TC = 2 with one plane per block, so that every lane has a column and
// a wave reads 32 consecutive lines' two values as one contiguous run.
template <class T, int N, int DIR, int AXIS, int TC = strided_tcols(N)>
__global__ __launch_bounds__(TC * strided_tpl(N)) void k_pass_strided(
    const cpx<T> *in, cpx<T> *out, long long ncols, long long lstride, long long ostride,
    int nouter, long long fs_in, long long fs_out, const cpx<T> *__restrict__ tw, T scale,
    const int *__restrict__ gate, Batch bt) {
    in = rebase(bt, 0, in);
    out = rebase(bt, 1, out);
    gate = rebase(bt, 2, gate);
    if (gate != nullptr && *gate == 0) return;
    constexpr int E = strided_e(N);
    constexpr int TPL = N / E;
    constexpr int LL = lds_len(N);
    __shared__ cpx<T> lds[line_needs_lds<N>() ? TC * LL : 1];

    const int tid = threadIdx.x;
    const int c = tid % TC, t = tid / TC;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const long long flat = (long long)b * TC + c;
    const long long o = flat / ncols;
    const long long col = flat - o * ncols;
    const bool active = o < nouter;
    const long long base_in = (long long)blockIdx.y * fs_in + o * ostride + col;
    const long long base_out = (long long)blockIdx.y * fs_out + o * ostride + col;

    cpx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
        cpx<T> a = active ? in[base_in + (long long)(t + i * TPL) * lstride] : cpx<T>{0, 0};
        v[i] = cscale(a, scale);
    }
    LineFFT<T, N, E, DIR>::run(v, lds + c * LL, t, tw);
    if (active) {
#pragma unroll
        for (int i = 0; i < E; i++) out[base_out + (long long)(t + i * TPL) * lstride] = v[i];
    }
}


// Real rows -> half-complex rows.  Row r of field f: real input at
// in + f*fs_in*2 + r*2*nzh (reals), complex output at out + f*fs_out + r*nzh.
template <class T, int NZ>
__global__ __launch_bounds__(zpass_rw(NZ / 2) * zpass_tpl(NZ / 2)) void k_pass_z_r2c(
    const T *in, cpx<T> *out, long long nrows, long long fs_in, long long fs_out,
    const cpx<T> *__restrict__ tw_nz, const cpx<T> *__restrict__ tw_m, Batch bt) {
    constexpr int M = NZ / 2;
    constexpr int NZH = M + 1;
    constexpr int E = zpass_e(M);
    using Row = ZRow<T, M, E>;
    constexpr int RW = zpass_rw(M);
    __shared__ cpx<T> lds[RW * Row::LL];
    in = rebase(bt, 0, in);
    out = rebase(bt, 1, out);

    const int tid = threadIdx.x;
    const int rl = tid / Row::TPL, t = tid % Row::TPL;
    const long long row = (long long)xcd_remap(blockIdx.x, gridDim.x) * RW + rl;
    const bool active = row < nrows;
    const cpx<T> *src = reinterpret_cast<const cpx<T> *>(in + (long long)blockIdx.y * fs_in * 2) + row * NZH;
    cpx<T> *dst = out + (long long)blockIdx.y * fs_out + row * NZH;

    cpx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; i++) v[i] = active ? src[t + i * Row::TPL] : cpx<T>{0, 0};
    // all loads of the block precede the first barrier inside r2c: in == out is safe
    Row::r2c(v, active, dst, lds + rl * Row::LL, t, tw_nz, tw_m);
}

// Half-complex rows -> real rows (padding reals written as 0).
template <class T, int NZ>
__global__ __launch_bounds__(zpass_rw(NZ / 2) * zpass_tpl(NZ / 2)) void k_pass_z_c2r(
    const cpx<T> *in, T *out, long long nrows, long long fs_in, long long fs_out,
    const cpx<T> *__restrict__ tw_nz, const cpx<T> *__restrict__ tw_m, Batch bt) {
    constexpr int M = NZ / 2;
    constexpr int NZH = M + 1;
    constexpr int E = zpass_e(M);
    using Row = ZRow<T, M, E>;
    constexpr int RW = zpass_rw(M);
    __shared__ cpx<T> lds[RW * Row::LL];
    in = rebase(bt, 0, in);
    out = rebase(bt, 1, out);

    const int tid = threadIdx.x;
    const int rl = tid / Row::TPL, t = tid % Row::TPL;
    const long long row = (long long)xcd_remap(blockIdx.x, gridDim.x) * RW + rl;
    const bool active = row < nrows;
    const cpx<T> *src = in + (long long)blockIdx.y * fs_in + row * NZH;
    cpx<T> *dstc = reinterpret_cast<cpx<T> *>(out + (long long)blockIdx.y * fs_out * 2) + row * NZH;

    cpx<T> v[E];
    // the whole row is staged in LDS before any store: in == out is safe
    Row::c2r(src, active, lds + rl * Row::LL, t, tw_nz, tw_m, (T)1, v);
    if (active) {
#pragma unroll
        for (int i = 0; i < E; i++) dstc[t + i * Row::TPL] = v[i];
        if (t == 0) dstc[M] = cpx<T>{(T)0, (T)0};
    }
}

// ------------------------------------------------------------------------
// twiddles
// ------------------------------------------------------------------------

static void fill_tw(std::vector<double> &h, size_t off, int n) {
    for (int e = 0; e < n; e++) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)e / (long double)n;
        h[2 * (off + e)] = (double)cosl(a);
        h[2 * (off + e) + 1] = (double)sinl(a);
    }
}

int twiddles_create(spk_ctx ctx, const Grid3 &g, int prec, Twiddles *tw) {
    // tables: nx | ny | nz | nz/2 (the half-length complex FFT of the z pass)
    const int mh = g.nz / 2 > 0 ? g.nz / 2 : 1;
    size_t n = (size_t)g.nx + g.ny + g.nz + mh;
    std::vector<double> h(2 * n);
    tw->off_x = 0;
    tw->off_y = g.nx;
    tw->off_z = g.nx + g.ny;
    fill_tw(h, 0, g.nx);
    fill_tw(h, g.nx, g.ny);
    fill_tw(h, g.nx + g.ny, g.nz);
    fill_tw(h, g.nx + g.ny + g.nz, mh);
    tw->prec = prec;
    size_t bytes = n * (prec == SPK_F64 ? 16 : 8);
    SPK_HIP(hipMalloc(&tw->d, bytes));
    if (prec == SPK_F64) {
        SPK_HIP(hipMemcpyAsync(tw->d, h.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    } else {
        std::vector<float> f(2 * n);
        for (size_t i = 0; i < 2 * n; i++) f[i] = (float)h[i];
        SPK_HIP(hipMemcpyAsync(tw->d, f.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    SPK_HIP(hipStreamSynchronize(ctx->stream));
    return SPK_OK;
}

void twiddles_destroy(Twiddles *tw) {
    if (tw->d) (void)hipFree(tw->d);
    tw->d = nullptr;
}

template <class T>
static const cpx<T> *tw_ptr(const Twiddles &tw, long long off) {
    return reinterpret_cast<const cpx<T> *>(tw.d) + off;
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------

#define SPK_POW2_CASES(X) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)

template <class T, int DIR, int AXIS>
static int launch_strided(spk_ctx ctx, int n, long long ncols, long long lstride, long long ostride,
                          int nouter, const cpx<T> *twp, const PassArgs &a) {
    Batch bt;
    SPK_TRY(make_batch(ctx, {a.in, a.out, ctx->gate}, &bt));
    const unsigned nbz = nboxes(ctx);
#define SPK_STRIDED_CASE(NN)                                                                       \
    case NN: {                                                                                     \
        constexpr int TC = strided_tcols_t<T>(NN);                                                 \
        if constexpr (AXIS == 1 && TC > 2) {                                                       \
            if (ncols == 2) {                                                                      \
                dim3 grid((unsigned)nouter, a.nfields, nbz);                                      \
                k_pass_strided<T, NN, DIR, AXIS, 2><<<grid, 2 * strided_tpl(NN), 0, ctx->stream>>>(   \
                    (const cpx<T> *)a.in, (cpx<T> *)a.out, ncols, lstride, ostride, nouter,         \
                    a.fstride_in, a.fstride_out, twp, (T)a.scale, ctx->gate, bt);                  \
                break;                                                                             \
            }                                                                                      \
        }                                                                                          \
        const long long ntiles = (ncols * nouter + TC - 1) / TC;                                   \
        dim3 grid((unsigned)ntiles, a.nfields, nbz);                                               \
        k_pass_strided<T, NN, DIR, AXIS, TC><<<grid, TC * strided_tpl(NN), 0, ctx->stream>>>(         \
            (const cpx<T> *)a.in, (cpx<T> *)a.out, ncols, lstride, ostride, nouter, a.fstride_in, \
            a.fstride_out, twp, (T)a.scale, ctx->gate, bt);                                        \
        break;                                                                                     \
    }
    switch (n) {
        SPK_POW2_CASES(SPK_STRIDED_CASE)
        default:
            set_error("line length %d not supported (power of two 2..2048)", n);
            return SPK_ERR_ARG;
    }
#undef SPK_STRIDED_CASE
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

template <int DIR, int AXIS>
static int pass_dispatch(spk_ctx ctx, int n, long long ncols, long long lstride, long long ostride,
                         int nouter, const Twiddles &tw, long long off, int tin, int tout,
                         const PassArgs &a) {
    SPK_REQUIRE(tin == tout && tin == tw.prec, SPK_ERR_ARG, "pass precision mismatch");
    if (tw.prec == SPK_F64)
        return launch_strided<double, DIR, AXIS>(ctx, n, ncols, lstride, ostride, nouter, tw_ptr<double>(tw, off), a);
    return launch_strided<float, DIR, AXIS>(ctx, n, ncols, lstride, ostride, nouter, tw_ptr<float>(tw, off), a);
}

int launch_pass_x(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int dir, int tin, int tout,
                  const PassArgs &a) {
    const long long cols = (long long)g.ny * g.nzh;
    if (dir < 0) return pass_dispatch<-1, 0>(ctx, g.nx, cols, cols, 0, 1, tw, tw.off_x, tin, tout, a);
    return pass_dispatch<+1, 0>(ctx, g.nx, cols, cols, 0, 1, tw, tw.off_x, tin, tout, a);
}

int launch_pass_y(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int dir, int tin, int tout,
                  const PassArgs &a) {
    const long long plane = (long long)g.ny * g.nzh;
    if (dir < 0) return pass_dispatch<-1, 1>(ctx, g.ny, g.nzh, g.nzh, plane, g.nx, tw, tw.off_y, tin, tout, a);
    return pass_dispatch<+1, 1>(ctx, g.ny, g.nzh, g.nzh, plane, g.nx, tw, tw.off_y, tin, tout, a);
}

#define SPK_Z_CASES(X) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

template <class T>
static int launch_z(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, bool r2c, const PassArgs &a) {
    const long long nrows = (long long)g.nx * g.ny;
    const cpx<T> *tz = tw_ptr<T>(tw, tw.off_z);
    const cpx<T> *tm = tw_ptr<T>(tw, tw.off_z + g.nz);
    Batch bt;
    SPK_TRY(make_batch(ctx, {a.in, a.out}, &bt));
    const unsigned nbz = nboxes(ctx);
#define SPK_Z_CASE(NN)                                                                         \
    case NN: {                                                                                 \
        constexpr int RW = zpass_rw(NN / 2);                                                   \
        dim3 grid((unsigned)((nrows + RW - 1) / RW), a.nfields, nbz);                          \
        const int bs = RW * zpass_tpl(NN / 2);                                                 \
        if (r2c)                                                                               \
            k_pass_z_r2c<T, NN><<<grid, bs, 0, ctx->stream>>>((const T *)a.in, (cpx<T> *)a.out, \
                                                              nrows, a.fstride_in,             \
                                                              a.fstride_out, tz, tm, bt);      \
        else                                                                                   \
            k_pass_z_c2r<T, NN><<<grid, bs, 0, ctx->stream>>>((const cpx<T> *)a.in, (T *)a.out, \
                                                              nrows, a.fstride_in,             \
                                                              a.fstride_out, tz, tm, bt);      \
        break;                                                                                 \
    }
    switch (g.nz) {
        SPK_Z_CASES(SPK_Z_CASE)
        default:
            set_error("NZ=%d not supported (power of two 2..1024)", g.nz);
            return SPK_ERR_ARG;
    }
#undef SPK_Z_CASE
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

int launch_pass_z_r2c(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int tin, int tout, const PassArgs &a) {
    SPK_REQUIRE(tin == tout && tin == tw.prec, SPK_ERR_ARG, "pass precision mismatch");
    return tw.prec == SPK_F64 ? launch_z<double>(ctx, g, tw, true, a) : launch_z<float>(ctx, g, tw, true, a);
}

int launch_pass_z_c2r(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int tin, int tout, const PassArgs &a) {
    SPK_REQUIRE(tin == tout && tin == tw.prec, SPK_ERR_ARG, "pass precision mismatch");
    return tw.prec == SPK_F64 ? launch_z<double>(ctx, g, tw, false, a) : launch_z<float>(ctx, g, tw, false, a);
}

}  // namespace spk

// ------------------------------------------------------------------------
// C ABI: plans
// ------------------------------------------------------------------------

struct spk_fft_plan_s {
    spk_ctx ctx;
    spk::Grid3 g;
    int prec;
    int batch;
    spk::Twiddles tw;
};

extern "C" int spk_fft_plan_create(spk_ctx ctx, int nx, int ny, int nz, int precision, int batch,
                                   spk_fft_plan *out) {
    using namespace spk;
    SPK_REQUIRE(ctx && out, SPK_ERR_ARG, "spk_fft_plan_create: null argument");
    SPK_REQUIRE(is_pow2(nx) && is_pow2(ny) && is_pow2(nz) && nx >= 2 && ny >= 2 && nz >= 2 &&
                    nx <= 2048 && ny <= 2048 && nz <= 1024,
                SPK_ERR_ARG, "spk_fft_plan_create: unsupported grid %dx%dx%d", nx, ny, nz);
    SPK_REQUIRE(precision == SPK_F64 || precision == SPK_F32, SPK_ERR_ARG, "bad precision %d", precision);
    SPK_REQUIRE(batch >= 1 && batch <= 65535, SPK_ERR_ARG, "bad batch %d", batch);
    SPK_HIP(hipSetDevice(ctx->device));
    spk_fft_plan p = new spk_fft_plan_s();
    p->ctx = ctx;
    p->g = make_grid(nx, ny, nz);
    p->prec = precision;
    p->batch = batch;
    int r = twiddles_create(ctx, p->g, precision, &p->tw);
    if (r != SPK_OK) {
        delete p;
        return r;
    }
    ctx_retain(ctx);
    *out = p;
    return SPK_OK;
}

extern "C" int spk_fft_plan_destroy(spk_fft_plan p) {
    if (!p) return SPK_OK;
    spk::twiddles_destroy(&p->tw);
    spk_ctx c = p->ctx;
    delete p;
    spk::ctx_release(c);
    return SPK_OK;
}

extern "C" int spk_fft_r2c(spk_fft_plan p, const void *in, void *out) {
    using namespace spk;
    SPK_REQUIRE(p && in && out, SPK_ERR_ARG, "spk_fft_r2c: null argument");
    PassArgs a{in, out, p->batch, p->g.ntc, p->g.ntc, 1.0};
    SPK_TRY(launch_pass_z_r2c(p->ctx, p->g, p->tw, p->prec, p->prec, a));
    PassArgs b{out, out, p->batch, p->g.ntc, p->g.ntc, 1.0};
    SPK_TRY(launch_pass_y(p->ctx, p->g, p->tw, -1, p->prec, p->prec, b));
    SPK_TRY(launch_pass_x(p->ctx, p->g, p->tw, -1, p->prec, p->prec, b));
    return SPK_OK;
}

extern "C" int spk_fft_c2r(spk_fft_plan p, const void *in, void *out) {
    using namespace spk;
    SPK_REQUIRE(p && in && out, SPK_ERR_ARG, "spk_fft_c2r: null argument");
    const double scale = 1.0 / ((double)p->g.nx * p->g.ny * p->g.nz);
    PassArgs a{in, out, p->batch, p->g.ntc, p->g.ntc, scale};
    SPK_TRY(launch_pass_x(p->ctx, p->g, p->tw, +1, p->prec, p->prec, a));
    PassArgs b{out, out, p->batch, p->g.ntc, p->g.ntc, 1.0};
    SPK_TRY(launch_pass_y(p->ctx, p->g, p->tw, +1, p->prec, p->prec, b));
    SPK_TRY(launch_pass_z_c2r(p->ctx, p->g, p->tw, p->prec, p->prec, b));
    return SPK_OK;
}

extern "C" int spk_fft_pass(spk_fft_plan p, int axis, int dir, const void *in, void *out) {
    using namespace spk;
    SPK_REQUIRE(p && in && out, SPK_ERR_ARG, "spk_fft_pass: null argument");
    SPK_REQUIRE(axis >= 0 && axis <= 2 && (dir == -1 || dir == 1), SPK_ERR_ARG, "spk_fft_pass: axis %d dir %d", axis, dir);
    PassArgs a{in, out, p->batch, p->g.ntc, p->g.ntc, 1.0};
    if (axis == 0) return launch_pass_x(p->ctx, p->g, p->tw, dir, p->prec, p->prec, a);
    if (axis == 1) return launch_pass_y(p->ctx, p->g, p->tw, dir, p->prec, p->prec, a);
    if (dir < 0) return launch_pass_z_r2c(p->ctx, p->g, p->tw, p->prec, p->prec, a);
    return launch_pass_z_c2r(p->ctx, p->g, p->tw, p->prec, p->prec, a);
}
