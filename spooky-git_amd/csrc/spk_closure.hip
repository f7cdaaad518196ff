// spk_closure.hip -- learned subgrid-stress closure: a 4-layer 3D convolution
// network on the periodic n^3 grid of a filtered velocity field, BF16 weights
// and activations on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16) with
// FP32 accumulation; forward, backward (input and weight gradients) and Adam
// written here.
//
//   layer 1: 3 (padded to 16) -> 32 channels, 3^3 taps, GELU
//   layers 2, 3: 32 -> 32, GELU
//   layer 4: 32 -> 6 (padded to 32), linear: the normalized stress
//            tau_xx, tau_xy, tau_xz, tau_yy, tau_yz, tau_zz
// Activations are channels-last [point][channel] BF16, point = (x n + y) n + z.
//
// Convolution as a GEMM per 4x4x32 brick of points: D[co][point] +=
// W[tap][co][ci-chunk] x X[point + tap][ci-chunk], 27 taps x (CIN / 16)
// K-chunks of 16 channels.  A block (8 waves: wave w takes x plane w % 4 and
// the two y columns 2 (w / 4), 2 (w / 4) + 1 of 32 z points) keeps all 27 taps'
// weights in LDS for its lifetime and stages each brick's input with a
// one-point halo; blocks loop over bricks.  16-byte LDS reads are
// conflict-free: the weights through an XOR swizzle of their 8-channel chunks
// by output channel, the halo at a padded point pitch.  The weight gradient
// contracts over points instead:
// dW[tap][co][ci] = sum_p dZ[p][co] X[p + tap][ci], with both operands staged
// channel-major so that a lane's 8 K values are 8 consecutive z points; the
// +-1 z shift of a tap is made with 16-bit funnel shifts of aligned loads.
#include "spk_internal.hpp"
#include "spk_kernels.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace spk {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int HX = 6, HY = 6, HZ = 34;  // brick 4 x 4 x 32 with a one-point halo
constexpr int CH = 32;                   // hidden channels
constexpr int NTAP = 27;
constexpr int NBLK = 256;                // persistent blocks (one per CU)

__host__ __device__ constexpr int swz_shift(int cin) { return cin == 32 ? 2 : 3; }
__device__ __forceinline__ int swz(int q, int p, int cin) { return q ^ ((p >> swz_shift(cin)) & (cin / 8 - 1)); }

// erfc(x) for x >= 0 by the Chebyshev fit of Press, Teukolsky, Vetterling and
// Flannery, Numerical Recipes (2nd ed., sec. 6.2, erfcc): fractional error below
// 1.2e-7 in exact arithmetic, below 1e-5 evaluated in FP32 up to x = 9; branch
// free, one reciprocal and one exponential.  The activations it feeds are
// rounded to BF16 (relative step 2^-8).
__device__ __forceinline__ float erfc_pos(float x) {
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.5f * x);
    float p = 0.17087277f;
    p = fmaf(p, t, -0.82215223f);
    p = fmaf(p, t, 1.48851587f);
    p = fmaf(p, t, -1.13520398f);
    p = fmaf(p, t, 0.27886807f);
    p = fmaf(p, t, -0.18628806f);
    p = fmaf(p, t, 0.09678418f);
    p = fmaf(p, t, 0.37409196f);
    p = fmaf(p, t, 1.00002368f);
    p = fmaf(p, t, -1.26551223f);
    return t * __expf(fmaf(-x, x, p));
}
// Phi(z) = erfc(-z / sqrt 2) / 2, the normal distribution function
__device__ __forceinline__ float norm_cdf(float z) {
    const float e = 0.5f * erfc_pos(fabsf(z) * 0.70710678118654752f);
    return z > 0.0f ? 1.0f - e : e;
}
// GELU(z) = z Phi(z), GELU'(z) = Phi(z) + z phi(z)
__device__ __forceinline__ float gelu(float z) { return z * norm_cdf(z); }
__device__ __forceinline__ float gelu_d(float z) {
    return norm_cdf(z) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Halo points in LDS at a pitch of CIN + 8 channels (80 or 48 bytes): the 16
// lanes of one 16-byte read cycle, on consecutive points, then fall on disjoint
// groups of 4 banks (64 banks of 4 bytes), and every tap's address is the lane's
// base address plus a constant, so the transform needs no per-tap address registers.
__host__ __device__ constexpr int xpitch(int cin) { return cin + 8; }
constexpr size_t conv_lds(int cin) { return (size_t)(NTAP * CH * cin + HX * HY * HZ * xpitch(cin)) * sizeof(bf16); }

constexpr int CONV_T = 512;  // 8 waves: two per SIMD, so one wave's LDS reads overlap the other's MFMAs

// One brick's input with its one-point halo (HX x HY x HZ points of CIN
// channels, channels-last, periodic wrap), held in registers between its global
// loads (fetch) and its LDS stores (stash: channels-last at the xpitch pitch;
// stash_t: channel-major rows of the weight-gradient layout), so that a
// block loads brick b + 1 while it computes brick b.
template <int CIN>
struct HaloPrefetch {
    static constexpr int NQ = CIN / 8;
    static constexpr int NCH = HX * HY * HZ * NQ;  // 16-byte chunks
    static constexpr int NPF = (NCH + CONV_T - 1) / CONV_T;
    bf16x8 v[NPF];

    __device__ __forceinline__ void fetch(const bf16 *__restrict__ X, int b, int n, int tid) {
        const int nbz = n / 32, nby = n / 4;
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * 4, y0 = by * 4, z0 = bz * 32;
#pragma unroll
        for (int k = 0; k < NPF; k++) {
            const int i = tid + k * CONV_T;
            if (i < NCH) {
                const int q = i % NQ, hp = i / NQ;
                const int hz = hp % HZ, hy = (hp / HZ) % HY, hx = hp / (HZ * HY);
                const int gx = wrap(x0 + hx - 1, n), gy = wrap(y0 + hy - 1, n), gz = wrap(z0 + hz - 1, n);
                v[k] = *reinterpret_cast<const bf16x8 *>(X + (((size_t)gx * n + gy) * n + gz) * CIN + 8 * q);
            }
        }
    }
    __device__ __forceinline__ void stash(bf16 *Xs, int tid) const {
#pragma unroll
        for (int k = 0; k < NPF; k++) {
            const int i = tid + k * CONV_T;
            if (i < NCH) {
                const int q = i % NQ, hp = i / NQ;
                *reinterpret_cast<bf16x8 *>(Xs + hp * xpitch(CIN) + 8 * q) = v[k];
            }
        }
    }
    // Xs[ci][XPL]: halo point (hx, hy, hz) at (hx HY + hy) 48 + 7 + hz
    template <int XPLEN>
    __device__ __forceinline__ void stash_t_impl(bf16 *Xs, int tid) const {
#pragma unroll
        for (int k = 0; k < NPF; k++) {
            const int i = tid + k * CONV_T;
            if (i < NCH) {
                const int q = i % NQ, hp = i / NQ;
                const int hz = hp % HZ, hy = (hp / HZ) % HY, hx = hp / (HZ * HY);
                const int e = (hx * HY + hy) * 48 + 7 + hz;
#pragma unroll
                for (int j = 0; j < 8; j++) Xs[(8 * q + j) * XPLEN + e] = v[k][j];
            }
        }
    }
    __device__ __forceinline__ void stash_t(bf16 *Xs, int tid) const;
};

// MODE 0: hidden layer, Y = GELU(conv + b); 1: also Z = conv + b (training);
// 2: output layer, out[c][p] = conv + b for c < 6 (FP32, channel-major);
// 3: input gradient, Y = conv(dZ, flipped transposed W) * GELU'(Zp)

template <int CIN, int MODE>
__global__ __launch_bounds__(CONV_T) void k_conv(const bf16 *__restrict__ X, const bf16 *__restrict__ W,
                                                 const float *__restrict__ bias, bf16 *__restrict__ Y,
                                                 bf16 *__restrict__ Z, const bf16 *__restrict__ Zp,
                                                 float *__restrict__ out, int n) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    bf16 *Ws = reinterpret_cast<bf16 *>(smem);  // [27][32][CIN], chunk-swizzled by co
    bf16 *Xs = Ws + NTAP * CH * CIN;            // [HX*HY*HZ][xpitch(CIN)]
    constexpr int NQ = CIN / 8;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 31, half = lane >> 5;
    const int xw = w & 3, yg = 2 * (w >> 2);  // the wave's x plane and first of its two y columns

    for (int i = tid; i < NTAP * CH * NQ; i += CONV_T) {
        const int q = i % NQ, row = i / NQ;
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(W + (size_t)row * CIN + 8 * q);
        *reinterpret_cast<bf16x8 *>(Ws + row * CIN + 8 * swz(q, row & 31, CIN)) = v;
    }
    const int nbz = n / 32, nby = n / 4;
    const int nb = (n / 4) * nby * nbz;
    const size_t P = (size_t)n * n * n;
    HaloPrefetch<CIN> hf;
    if (blockIdx.x < nb) hf.fetch(X, blockIdx.x, n, tid);
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * 4, y0 = by * 4, z0 = bz * 32;
        __syncthreads();
        hf.stash(Xs, tid);
        __syncthreads();
        // the next brick's loads are in flight during this brick's matrix work (the
        // scheduling barrier keeps them ahead of it)
        if (b + (int)gridDim.x < nb) hf.fetch(X, b + gridDim.x, n, tid);
        __builtin_amdgcn_sched_barrier(0);
        f32x16 acc[2];
#pragma unroll
        for (int ys = 0; ys < 2; ys++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[ys][r] = 0.0f;
#pragma unroll
        for (int t = 0; t < NTAP; t++) {
            const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
#pragma unroll
            for (int kc = 0; kc < CIN / 16; kc++) {
                const int q = 2 * kc + half;
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(Ws + (t * CH + li) * CIN + 8 * swz(q, li, CIN));
#pragma unroll
                for (int ys = 0; ys < 2; ys++) {
                    const int hp = ((xw + 1 + dx) * HY + (yg + ys + 1 + dy)) * HZ + li + 1 + dz;
                    const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(Xs + hp * xpitch(CIN) + 8 * q);
                    acc[ys] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc[ys], 0, 0, 0);
                }
            }
        }
        // D[co][point]: lane holds point li, co = 8 (r / 4) + 4 half + r % 4
#pragma unroll
        for (int ys = 0; ys < 2; ys++) {
            const size_t p = ((size_t)(x0 + xw) * n + (y0 + yg + ys)) * n + z0 + li;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int c0 = 8 * g + 4 * half;
                if (MODE == 2) {
                    if (c0 < 6) {
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            if (c0 + j < 6) out[(size_t)(c0 + j) * P + p] = acc[ys][4 * g + j] + bias[c0 + j];
                    }
                } else if (MODE == 3) {
                    const bf16x4 zp = *reinterpret_cast<const bf16x4 *>(Zp + p * CH + c0);
                    bf16x4 y;
#pragma unroll
                    for (int j = 0; j < 4; j++) y[j] = (bf16)(acc[ys][4 * g + j] * gelu_d((float)zp[j]));
                    *reinterpret_cast<bf16x4 *>(Y + p * CH + c0) = y;
                } else {
                    bf16x4 y, zz;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const float zf = acc[ys][4 * g + j] + bias[c0 + j];
                        zz[j] = (bf16)zf;
                        y[j] = (bf16)gelu(zf);
                    }
                    *reinterpret_cast<bf16x4 *>(Y + p * CH + c0) = y;
                    if (MODE == 1) *reinterpret_cast<bf16x4 *>(Z + p * CH + c0) = zz;
                }
            }
        }
    }
}

// Output layer (32 -> 6 channels) on the 16-row matrix instruction
// v_mfma_f32_16x16x32_bf16: D[co < 16][16 points] += W[tap][co][32 ci] X[point + tap][32 ci],
// one instruction per tap and 16-point tile (the 6 stresses in rows 0..5), half
// the work of a 32-row tile.  out[c][p] = D + b for c < 6.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr size_t OUT_LDS = (size_t)(NTAP * 16 * CH + HX * HY * HZ * xpitch(CH)) * sizeof(bf16);

__global__ __launch_bounds__(CONV_T) void k_conv_out(const bf16 *__restrict__ X, const bf16 *__restrict__ W,
                                                     const float *__restrict__ bias, float *__restrict__ out, int n) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    bf16 *Ws = reinterpret_cast<bf16 *>(smem);  // [27][16][32], chunk-swizzled by co
    bf16 *Xs = Ws + NTAP * 16 * CH;             // [HX*HY*HZ][xpitch(32)]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int l16 = lane & 15, q = lane >> 4;   // 16x16x32: lane holds row/column l16, K chunk q
    const int xw = w & 3, yg = 2 * (w >> 2);

    for (int i = tid; i < NTAP * 16 * 4; i += CONV_T) {
        const int qq = i & 3, row = i >> 2;  // row = t * 16 + co
        const int t = row >> 4, co = row & 15;
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(W + ((size_t)t * CH + co) * CH + 8 * qq);
        *reinterpret_cast<bf16x8 *>(Ws + row * CH + 8 * swz(qq, co, CH)) = v;
    }
    const int nbz = n / 32, nby = n / 4;
    const int nb = (n / 4) * nby * nbz;
    const size_t P = (size_t)n * n * n;
    HaloPrefetch<CH> hf;
    if (blockIdx.x < nb) hf.fetch(X, blockIdx.x, n, tid);
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * 4, y0 = by * 4, z0 = bz * 32;
        __syncthreads();
        hf.stash(Xs, tid);
        __syncthreads();
        if (b + (int)gridDim.x < nb) hf.fetch(X, b + gridDim.x, n, tid);
        __builtin_amdgcn_sched_barrier(0);
        f32x4 acc[2][2];  // [column][16-point half]
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int r = 0; r < 4; r++) acc[c][h][r] = 0.0f;
#pragma unroll
        for (int t = 0; t < NTAP; t++) {
            const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(Ws + (t * 16 + l16) * CH + 8 * swz(q, l16, CH));
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int hp = ((xw + 1 + dx) * HY + (yg + c + 1 + dy)) * HZ + 16 * h + l16 + 1 + dz;
                    const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(Xs + hp * xpitch(CH) + 8 * q);
                    acc[c][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[c][h], 0, 0, 0);
                }
        }
        // D[co][point]: lane holds point l16 of its half, co = 4 q + r
        if (q < 2) {
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const size_t p = ((size_t)(x0 + xw) * n + (y0 + yg + c)) * n + z0 + 16 * h + l16;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int co = 4 * q + r;
                        if (co < 6) out[(size_t)co * P + p] = acc[c][h][r] + bias[co];
                    }
                }
        }
    }
}

// weight gradient, per-block partial sums part[block][27][32 co][32 ci]
constexpr int DZS = 512 + 8;                 // dZ^T row (one co): 512 brick points + pad (65 16-byte slots)
constexpr int XPL = HX * HY * 48 + 8;        // X^T plane (one ci): z rows of 48 + pad (217 slots)
constexpr size_t DW_LDS = (size_t)(CH * DZS + CH * XPL) * sizeof(bf16);

template <int CIN>
__device__ __forceinline__ void HaloPrefetch<CIN>::stash_t(bf16 *Xs, int tid) const {
    stash_t_impl<XPL>(Xs, tid);
}

template <int CIN>
__global__ __launch_bounds__(CONV_T) void k_conv_dw(const bf16 *__restrict__ X, const bf16 *__restrict__ dZ,
                                                 float *__restrict__ part, float *__restrict__ bpart, int n) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    bf16 *Ds = reinterpret_cast<bf16 *>(smem);  // [32 co][DZS]: point (bx, by, z) at (bx 4 + by) 32 + z
    bf16 *Xs = Ds + CH * DZS;                   // [32 ci][XPL]: (hx, hy, z0 - 8 + e) at (hx 6 + hy) 48 + e
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 31, half = lane >> 5;
    for (int i = tid; i < CH * XPL; i += CONV_T) Xs[i] = (bf16)0.0f;  // padding and channels >= CIN stay 0
    f32x16 acc[4];  // taps w, w + 8, w + 16, w + 24 of wave w
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[k][r] = 0.0f;
    // bias gradient: the block's sum of dZ per output channel, read from the
    // bricks already in LDS; thread tid sums channel tid / 16, 8-point chunks
    // (tid % 16) + 16 r of each brick (consecutive lanes, consecutive 16 bytes)
    const int bco = tid >> 4, bseg = tid & 15;
    float bacc = 0.0f;
    const int nbz = n / 32, nby = n / 4;
    const int nb = (n / 4) * nby * nbz;
    HaloPrefetch<CIN> hf;
    constexpr int NPD = 512 * 4 / CONV_T;  // dZ chunks of 8 channels per thread
    bf16x8 pd[NPD];
    auto fetch_dz = [&](int b) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * 4, y0 = by * 4, z0 = bz * 32;
#pragma unroll
        for (int k = 0; k < NPD; k++) {
            const int i = tid + k * CONV_T;
            const int q = i & 3, pt = i >> 2;
            const size_t p = ((size_t)(x0 + (pt >> 7)) * n + y0 + ((pt >> 5) & 3)) * n + z0 + (pt & 31);
            pd[k] = *reinterpret_cast<const bf16x8 *>(dZ + p * CH + 8 * q);
        }
    };
    if (blockIdx.x < nb) {
        fetch_dz(blockIdx.x);
        hf.fetch(X, blockIdx.x, n, tid);
    }
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NPD; k++) {
            const int i = tid + k * CONV_T;
            const int q = i & 3, pt = i >> 2;
#pragma unroll
            for (int j = 0; j < 8; j++) Ds[(8 * q + j) * DZS + pt] = pd[k][j];
        }
        hf.stash_t(Xs, tid);
        __syncthreads();
        // the next brick's loads are in flight during this brick's matrix work
        if (b + (int)gridDim.x < nb) {
            fetch_dz(b + gridDim.x);
            hf.fetch(X, b + gridDim.x, n, tid);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const bf16x8 v = *reinterpret_cast<const bf16x8 *>(Ds + bco * DZS + (bseg + 16 * r) * 8);
#pragma unroll
            for (int j = 0; j < 8; j++) bacc += (float)v[j];
        }
        for (int s = 0; s < 32; s++) {
            const int run = 2 * s + half, col = run >> 2, zr = (run & 3) * 8;
            const int bxl = col >> 2, byl = col & 3;
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(Ds + li * DZS + col * 32 + zr);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int t = w + 8 * k;
                if (t < NTAP) {
                    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
                    const int base = li * XPL + ((bxl + 1 + dx) * HY + (byl + 1 + dy)) * 48 + 8 + zr;
                    const u32x4 d = *reinterpret_cast<const u32x4 *>(Xs + base);
                    u32x4 sv = d;
                    if (dz > 0) {
                        const unsigned ex = *reinterpret_cast<const unsigned *>(Xs + base + 8);
                        sv[0] = __builtin_amdgcn_alignbit(d[1], d[0], 16);
                        sv[1] = __builtin_amdgcn_alignbit(d[2], d[1], 16);
                        sv[2] = __builtin_amdgcn_alignbit(d[3], d[2], 16);
                        sv[3] = __builtin_amdgcn_alignbit(ex, d[3], 16);
                    } else if (dz < 0) {
                        const unsigned ex = *reinterpret_cast<const unsigned *>(Xs + base - 2);
                        sv[0] = __builtin_amdgcn_alignbit(d[0], ex, 16);
                        sv[1] = __builtin_amdgcn_alignbit(d[1], d[0], 16);
                        sv[2] = __builtin_amdgcn_alignbit(d[2], d[1], 16);
                        sv[3] = __builtin_amdgcn_alignbit(d[3], d[2], 16);
                    }
                    bf16x8 bv;
                    __builtin_memcpy(&bv, &sv, 16);
                    acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc[k], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) bacc += __shfl_xor(bacc, o, 64);
    if (bseg == 0) bpart[(size_t)blockIdx.x * CH + bco] = bacc;
    // D[co][ci]: lane holds ci = li, co = 8 (r / 4) + 4 half + r % 4
    float *pb = part + (size_t)blockIdx.x * NTAP * CH * CH;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int t = w + 8 * k;
        if (t < NTAP) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int co = 8 * (r / 4) + 4 * half + (r & 3);
                pb[((size_t)t * CH + co) * CH + li] = acc[k][r];
            }
        }
    }
}

__global__ void k_fill0(float *__restrict__ p, int count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) p[k] = 0.0f;
}

// grad[k] += sum over slice blockIdx.y of the parts: part[b][k] (grad zeroed before)
constexpr int SUM_SLICES = 16;
__global__ void k_sum_parts(const float *__restrict__ part, int nparts, int count, float *__restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int per = (nparts + SUM_SLICES - 1) / SUM_SLICES;
    const int b0 = blockIdx.y * per, b1 = b0 + per < nparts ? b0 + per : nparts;
    float s = 0.0f;
    for (int b = b0; b < b1; b++) s += part[(size_t)b * count + k];
    atomicAdd(&grad[k], s);
}

// padded real FP64 u [3][n][n][n + 2] -> BF16 [P][16] (channels 3..15 zero), times in_scale
__global__ void k_pack_input(const double *__restrict__ u, bf16 *__restrict__ xin, int n, float in_scale) {
    const size_t P = (size_t)n * n * n;
    const size_t fs = (size_t)n * n * (n + 2);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        const size_t r = p / n, z = p - r * n;
        const size_t o = r * (n + 2) + z;
        bf16x8 v0, v1;
        for (int j = 0; j < 8; j++) v0[j] = v1[j] = (bf16)0.0f;
        v0[0] = (bf16)((float)u[o] * in_scale);
        v0[1] = (bf16)((float)u[fs + o] * in_scale);
        v0[2] = (bf16)((float)u[2 * fs + o] * in_scale);
        *reinterpret_cast<bf16x8 *>(xin + p * 16) = v0;
        *reinterpret_cast<bf16x8 *>(xin + p * 16 + 8) = v1;
    }
}

// normalized FP32 [6][P] <-> padded real FP64 [6][n][n][n + 2] with per-component scales
__global__ void k_unpack_output(const float *__restrict__ out, double *__restrict__ tau, int n, double s0, double s1,
                                double s2, double s3, double s4, double s5) {
    const size_t P = (size_t)n * n * n;
    const size_t fs = (size_t)n * n * (n + 2);
    const double sc[6] = {s0, s1, s2, s3, s4, s5};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < 6 * fs; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i / fs);
        const size_t e = i - (size_t)c * fs;
        const size_t r = e / (n + 2), z = e - r * (n + 2);
        tau[i] = z < (size_t)n ? (double)out[(size_t)c * P + r * n + z] * sc[c] : 0.0;
    }
}
__global__ void k_prep_target(const double *__restrict__ tau, float *__restrict__ tgt, int n, double s0, double s1,
                              double s2, double s3, double s4, double s5) {
    const size_t P = (size_t)n * n * n;
    const size_t fs = (size_t)n * n * (n + 2);
    const double sc[6] = {s0, s1, s2, s3, s4, s5};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < 6 * P; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i / P);
        const size_t p = i - (size_t)c * P;
        const size_t r = p / n, z = p - r * n;
        tgt[i] = (float)(tau[(size_t)c * fs + r * (n + 2) + z] / sc[c]);
    }
}

// loss = mean over 6 P of (pred - tgt)^2; dZ4[p][c] = 2 (pred - tgt) / (6 P) (c < 6, else 0)
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ pred, const float *__restrict__ tgt,
                                              bf16 *__restrict__ dz, size_t P, float *__restrict__ part) {
    __shared__ float sm[256];
    float acc = 0.0f;
    const float g = 2.0f / (6.0f * (float)P);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        bf16x8 v0, v1, v2, v3;
        for (int j = 0; j < 8; j++) v0[j] = v1[j] = v2[j] = v3[j] = (bf16)0.0f;
        for (int c = 0; c < 6; c++) {
            const float d = pred[(size_t)c * P + p] - tgt[(size_t)c * P + p];
            acc += d * d;
            v0[c] = (bf16)(g * d);
        }
        bf16 *o = dz + p * CH;
        *reinterpret_cast<bf16x8 *>(o) = v0;
        *reinterpret_cast<bf16x8 *>(o + 8) = v1;
        *reinterpret_cast<bf16x8 *>(o + 16) = v2;
        *reinterpret_cast<bf16x8 *>(o + 24) = v3;
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Adam (Kingma & Ba) on the FP32 master parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int count, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// BF16 weights of one layer from its FP32 master weights W[27][32][cin]; and the
// flipped transpose WT[26 - t][ci][co] = W[t][co][ci] (cin = 32) for the input gradient
__global__ void k_pack_weights(const float *__restrict__ W, bf16 *__restrict__ Wb, bf16 *__restrict__ WT, int cin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NTAP * CH * cin) return;
    const int ci = k % cin, co = (k / cin) % CH, t = k / (cin * CH);
    const bf16 v = (bf16)W[k];
    Wb[k] = v;
    if (WT) WT[((NTAP - 1 - t) * CH + ci) * CH + co] = v;
}

// ---- filtered DNS data: truncation to a coarse grid, products, SGS divergence ----

// coarse-grid index -> fine-grid index of the same wavenumber along one axis;
// -1 where the coarse mode is cut (beyond |k| <= kc, or the coarse Nyquist)
__device__ __forceinline__ long long fine_of(int i, int n, int N, int kc) {
    const int k = i < n / 2 ? i : i - n;
    if (i == n / 2 || (k < 0 ? -k : k) > kc) return -1;
    return k >= 0 ? k : k + N;
}

// lo[f][i][j][k] = scale * hi[f][fine(i)][fine(j)][k] on the modes kept, 0 elsewhere
// (complex, [NX][NY][NZ/2+1] layouts; hi at stride fs_hi, lo at stride fs_lo)
__global__ void k_truncate(const cpx<double> *__restrict__ hi, int N, long long fs_hi, cpx<double> *__restrict__ lo,
                           int n, long long fs_lo, int nf, int kc, double scale) {
    const long long ntc = (long long)n * n * (n / 2 + 1);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ntc * nf;
         e += (long long)gridDim.x * blockDim.x) {
        const int f = (int)(e / ntc);
        const long long idx = e - (long long)f * ntc;
        const int k = (int)(idx % (n / 2 + 1));
        const long long r = idx / (n / 2 + 1);
        const int j = (int)(r % n), i = (int)(r / n);
        const long long fi = fine_of(i, n, N, kc), fj = fine_of(j, n, N, kc);
        const bool keep = fi >= 0 && fj >= 0 && k < n / 2 && k <= kc;
        cpx<double> v = {0.0, 0.0};
        if (keep) {
            const cpx<double> h = hi[(long long)f * fs_hi + (fi * N + fj) * (N / 2 + 1) + k];
            v = cpx<double>{h.x * scale, h.y * scale};
        }
        lo[(long long)f * fs_lo + idx] = v;
    }
}

// the same cut applied on the fine grid itself (modes beyond kc set to 0)
__global__ void k_mask_fine(const cpx<double> *__restrict__ in, cpx<double> *__restrict__ out, int N, long long fs,
                            int nf, int kc) {
    const long long ntc = (long long)N * N * (N / 2 + 1);
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ntc * nf;
         e += (long long)gridDim.x * blockDim.x) {
        const int f = (int)(e / ntc);
        const long long idx = e - (long long)f * ntc;
        const int k = (int)(idx % (N / 2 + 1));
        const long long r = idx / (N / 2 + 1);
        const int j = (int)(r % N), i = (int)(r / N);
        const int kx = i < N / 2 ? i : i - N, ky = j < N / 2 ? j : j - N;
        const bool keep = (kx < 0 ? -kx : kx) <= kc && (ky < 0 ? -ky : ky) <= kc && k <= kc;
        out[(long long)f * fs + idx] = keep ? in[(long long)f * fs + idx] : cpx<double>{0.0, 0.0};
    }
}

// padded real rows: out = a_i a_j - b_i b_j (fields i, j of a and b at stride fs)
__global__ void k_stress_prod(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ out,
                              long long fs, int i, int j) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < fs; e += (long long)gridDim.x * blockDim.x)
        out[e] = a[i * fs + e] * a[j * fs + e] - b[i * fs + e] * b[j * fs + e];
}

// U_i += dt P(-i mask k_j tau_ij) for the velocity of a solver (tau: 6 spectral
// fields xx xy xz yy yz zz at stride ntc); P the projection on k-perpendicular
__global__ __launch_bounds__(256) void k_sgs_update(cpx<double> *__restrict__ U, const cpx<double> *__restrict__ T,
                                                    long long ntc, KGeom kg, const TimeState *__restrict__ ts) {
    const double dt = ts->dt;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < ntc;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long rowi = idx / kg.nzh;
        const int k = (int)(idx - rowi * kg.nzh);
        const int j = (int)(rowi % kg.ny), i = (int)(rowi / kg.ny);
        double kx, ky, kz, m;
        kvec_at(kg, i, j, k, kx, ky, kz, m);
        const cpx<double> t0 = T[idx], t1 = T[ntc + idx], t2 = T[2 * ntc + idx], t3 = T[3 * ntc + idx],
                          t4 = T[4 * ntc + idx], t5 = T[5 * ntc + idx];
        const cpx<double> d0 = kdot3(kx, t0, ky, t1, kz, t2);
        const cpx<double> d1 = kdot3(kx, t1, ky, t3, kz, t4);
        const cpx<double> d2 = kdot3(kx, t2, ky, t4, kz, t5);
        // -i m d
        cpx<double> g[3] = {{m * d0.y, -m * d0.x}, {m * d1.y, -m * d1.x}, {m * d2.y, -m * d2.x}};
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 > 0.0) {
            const cpx<double> kg_ = kdot3(kx, g[0], ky, g[1], kz, g[2]);
            const double kk[3] = {kx, ky, kz};
#pragma unroll
            for (int q = 0; q < 3; q++) g[q] = cpx<double>{g[q].x - kk[q] * kg_.x / k2, g[q].y - kk[q] * kg_.y / k2};
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
            cpx<double> &u = U[q * ntc + idx];
            u = cpx<double>{u.x + dt * g[q].x, u.y + dt * g[q].y};
        }
    }
}

}  // namespace
}  // namespace spk

using namespace spk;

// parameter layout: for layer l = 1..4: W_l[27][32][cin_l] then b_l[32]; cin = 16, 32, 32, 32
static const int CIN_L[4] = {16, 32, 32, 32};
static size_t w_count(int l) { return (size_t)NTAP * CH * CIN_L[l]; }
static size_t param_offset(int l) {
    size_t o = 0;
    for (int i = 0; i < l; i++) o += w_count(i) + CH;
    return o;
}
static size_t param_total() { return param_offset(4); }

struct spk_closure_s {
    spk_ctx ctx = nullptr;
    int n = 0;
    size_t P = 0;
    double in_scale = 1.0, out_scale[6] = {1, 1, 1, 1, 1, 1};
    float *prm = nullptr, *grd = nullptr, *am = nullptr, *av = nullptr;  // FP32 master, gradient, Adam moments
    bf16 *Wb[4] = {}, *WT[4] = {};
    bf16 *xin = nullptr, *Z[3] = {}, *A[3] = {}, *dzA = nullptr, *dzB = nullptr;
    float *out = nullptr, *tgt = nullptr, *part = nullptr, *bpart = nullptr, *lpart = nullptr;
    int adam_t = 0;
    double *tau = nullptr;          // coupling: padded real stresses [6][n][n][n + 2], transformed in place
    spk_fft_plan plan6 = nullptr;   // FP64 n^3 transforms of 6 fields
};

namespace {

int closure_alloc(spk_closure c) {
    auto al = [&](void **p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            set_error("spk_closure: hipMalloc(%zu) failed", bytes);
            return SPK_ERR_ALLOC;
        }
        return SPK_OK;
    };
    const size_t NP = param_total();
    SPK_TRY(al((void **)&c->prm, NP * 4));
    SPK_TRY(al((void **)&c->grd, NP * 4));
    SPK_TRY(al((void **)&c->am, NP * 4));
    SPK_TRY(al((void **)&c->av, NP * 4));
    for (int l = 0; l < 4; l++) {
        SPK_TRY(al((void **)&c->Wb[l], w_count(l) * 2));
        if (l > 0) SPK_TRY(al((void **)&c->WT[l], w_count(l) * 2));
    }
    SPK_TRY(al((void **)&c->xin, c->P * 16 * 2));
    for (int l = 0; l < 3; l++) {
        SPK_TRY(al((void **)&c->Z[l], c->P * CH * 2));
        SPK_TRY(al((void **)&c->A[l], c->P * CH * 2));
    }
    SPK_TRY(al((void **)&c->dzA, c->P * CH * 2));
    SPK_TRY(al((void **)&c->dzB, c->P * CH * 2));
    SPK_TRY(al((void **)&c->out, c->P * 6 * 4));
    SPK_TRY(al((void **)&c->tgt, c->P * 6 * 4));
    SPK_TRY(al((void **)&c->part, (size_t)NBLK * NTAP * CH * CH * 4));
    SPK_TRY(al((void **)&c->bpart, (size_t)NBLK * CH * 4));
    SPK_TRY(al((void **)&c->lpart, (size_t)NBLK * 4));
    SPK_HIP(hipMemsetAsync(c->prm, 0, NP * 4, c->ctx->stream));
    SPK_HIP(hipMemsetAsync(c->am, 0, NP * 4, c->ctx->stream));
    SPK_HIP(hipMemsetAsync(c->av, 0, NP * 4, c->ctx->stream));
    return SPK_OK;
}

void closure_free(spk_closure c) {
    if (c->plan6) spk_fft_plan_destroy(c->plan6);
    void *b[] = {c->prm, c->grd, c->am, c->av, c->xin, c->dzA, c->dzB, c->out, c->tgt, c->part, c->bpart, c->lpart,
                 c->tau};
    for (void *p : b)
        if (p) (void)hipFree(p);
    for (int l = 0; l < 4; l++) {
        if (c->Wb[l]) (void)hipFree(c->Wb[l]);
        if (c->WT[l]) (void)hipFree(c->WT[l]);
    }
    for (int l = 0; l < 3; l++) {
        if (c->Z[l]) (void)hipFree(c->Z[l]);
        if (c->A[l]) (void)hipFree(c->A[l]);
    }
}

int set_lds_attributes() {
    static bool done = false;
    if (done) return SPK_OK;
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv<16, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)conv_lds(16)));
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv<16, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)conv_lds(16)));
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv<32, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)conv_lds(32)));
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv<32, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)conv_lds(32)));
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv_out, hipFuncAttributeMaxDynamicSharedMemorySize, (int)OUT_LDS));
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv<32, 3>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)conv_lds(32)));
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv_dw<16>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)DW_LDS));
    SPK_HIP(hipFuncSetAttribute((const void *)k_conv_dw<32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)DW_LDS));
    done = true;
    return SPK_OK;
}

int pack_weights(spk_closure c) {
    for (int l = 0; l < 4; l++) {
        const int cnt = (int)w_count(l);
        k_pack_weights<<<(cnt + 255) / 256, 256, 0, c->ctx->stream>>>(c->prm + param_offset(l), c->Wb[l], c->WT[l],
                                                                     CIN_L[l]);
    }
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

unsigned conv_grid(spk_closure c) {
    const int nb = (c->n / 4) * (c->n / 4) * (c->n / 32);
    return (unsigned)(nb < NBLK ? nb : NBLK);
}

int forward(spk_closure c, const void *u, bool train) {
    hipStream_t st = c->ctx->stream;
    const unsigned gp = (unsigned)((c->P + 255) / 256 < 4096 ? (c->P + 255) / 256 : 4096);
    k_pack_input<<<gp, 256, 0, st>>>((const double *)u, c->xin, c->n, (float)c->in_scale);
    const unsigned g = conv_grid(c);
    const float *b[4];
    for (int l = 0; l < 4; l++) b[l] = c->prm + param_offset(l) + w_count(l);
    if (train) {
        k_conv<16, 1><<<g, CONV_T, conv_lds(16), st>>>(c->xin, c->Wb[0], b[0], c->A[0], c->Z[0], nullptr, nullptr, c->n);
        k_conv<32, 1><<<g, CONV_T, conv_lds(32), st>>>(c->A[0], c->Wb[1], b[1], c->A[1], c->Z[1], nullptr, nullptr, c->n);
        k_conv<32, 1><<<g, CONV_T, conv_lds(32), st>>>(c->A[1], c->Wb[2], b[2], c->A[2], c->Z[2], nullptr, nullptr, c->n);
    } else {
        k_conv<16, 0><<<g, CONV_T, conv_lds(16), st>>>(c->xin, c->Wb[0], b[0], c->A[0], nullptr, nullptr, nullptr, c->n);
        k_conv<32, 0><<<g, CONV_T, conv_lds(32), st>>>(c->A[0], c->Wb[1], b[1], c->A[1], nullptr, nullptr, nullptr, c->n);
        k_conv<32, 0><<<g, CONV_T, conv_lds(32), st>>>(c->A[1], c->Wb[2], b[2], c->A[2], nullptr, nullptr, nullptr, c->n);
    }
    k_conv_out<<<g, CONV_T, OUT_LDS, st>>>(c->A[2], c->Wb[3], b[3], c->out, c->n);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

// weight and bias gradients of layer l from its input X and its dZ, into grd
int layer_grads(spk_closure c, int l, const bf16 *X, const bf16 *dz) {
    hipStream_t st = c->ctx->stream;
    const unsigned g = conv_grid(c);
    if (CIN_L[l] == 16) k_conv_dw<16><<<g, CONV_T, DW_LDS, st>>>(X, dz, c->part, c->bpart, c->n);
    else k_conv_dw<32><<<g, CONV_T, DW_LDS, st>>>(X, dz, c->part, c->bpart, c->n);
    const int cnt = NTAP * CH * CH;
    // the [27][32][32] sums land in grd for cin = 32; for cin = 16 take ci < 16
    float *gw = c->grd + param_offset(l);
    const dim3 sg((cnt + 255) / 256, SUM_SLICES);
    float *sum_to = CIN_L[l] == 32 ? gw : c->out;  // out is free here
    k_fill0<<<(cnt + 255) / 256, 256, 0, st>>>(sum_to, cnt);
    k_sum_parts<<<sg, 256, 0, st>>>(c->part, (int)g, cnt, sum_to);
    if (CIN_L[l] != 32)
        SPK_HIP(hipMemcpy2DAsync(gw, 16 * 4, c->out, 32 * 4, 16 * 4, NTAP * CH, hipMemcpyDeviceToDevice, st));
    // the bias partials came from the same kernel (k_conv_dw), one per block
    k_fill0<<<1, CH, 0, st>>>(gw + w_count(l), CH);
    k_sum_parts<<<dim3(1, SUM_SLICES), CH, 0, st>>>(c->bpart, (int)g, CH, gw + w_count(l));
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

}  // namespace

extern "C" size_t spk_closure_num_params(void) { return param_total(); }

extern "C" int spk_closure_create(spk_ctx ctx, int n, spk_closure *out) {
    SPK_REQUIRE(ctx && out, SPK_ERR_ARG, "spk_closure_create: null argument");
    SPK_REQUIRE(n >= 32 && n % 32 == 0 && n <= 1024, SPK_ERR_ARG,
                "spk_closure_create: n = %d (a multiple of 32, 32..1024)", n);
    SPK_TRY(set_lds_attributes());
    spk_closure c = new spk_closure_s();
    c->ctx = ctx;
    c->n = n;
    c->P = (size_t)n * n * n;
    const int rc = closure_alloc(c);
    if (rc != SPK_OK) {
        closure_free(c);
        delete c;
        return rc;
    }
    ctx_retain(ctx);
    *out = c;
    return pack_weights(c);
}

extern "C" int spk_closure_destroy(spk_closure c) {
    if (!c) return SPK_OK;
    (void)hipStreamSynchronize(c->ctx->stream);
    closure_free(c);
    spk_ctx cx = c->ctx;
    delete c;
    ctx_release(cx);
    return SPK_OK;
}

extern "C" int spk_closure_set_params(spk_closure c, const float *host) {
    SPK_REQUIRE(c && host, SPK_ERR_ARG, "spk_closure_set_params: null argument");
    SPK_HIP(hipMemcpyAsync(c->prm, host, param_total() * 4, hipMemcpyHostToDevice, c->ctx->stream));
    SPK_HIP(hipMemsetAsync(c->am, 0, param_total() * 4, c->ctx->stream));
    SPK_HIP(hipMemsetAsync(c->av, 0, param_total() * 4, c->ctx->stream));
    c->adam_t = 0;
    SPK_TRY(pack_weights(c));
    SPK_HIP(hipStreamSynchronize(c->ctx->stream));
    return SPK_OK;
}

extern "C" int spk_closure_get_params(spk_closure c, float *host) {
    SPK_REQUIRE(c && host, SPK_ERR_ARG, "spk_closure_get_params: null argument");
    SPK_HIP(hipMemcpyAsync(host, c->prm, param_total() * 4, hipMemcpyDeviceToHost, c->ctx->stream));
    SPK_HIP(hipStreamSynchronize(c->ctx->stream));
    return SPK_OK;
}

extern "C" int spk_closure_get_grads(spk_closure c, float *host) {
    SPK_REQUIRE(c && host, SPK_ERR_ARG, "spk_closure_get_grads: null argument");
    SPK_HIP(hipMemcpyAsync(host, c->grd, param_total() * 4, hipMemcpyDeviceToHost, c->ctx->stream));
    SPK_HIP(hipStreamSynchronize(c->ctx->stream));
    return SPK_OK;
}

extern "C" int spk_closure_set_scales(spk_closure c, double in_scale, const double *out_scale) {
    SPK_REQUIRE(c && out_scale && in_scale > 0, SPK_ERR_ARG, "spk_closure_set_scales: bad arguments");
    c->in_scale = in_scale;
    for (int i = 0; i < 6; i++) {
        SPK_REQUIRE(out_scale[i] > 0, SPK_ERR_ARG, "spk_closure_set_scales: out_scale[%d] <= 0", i);
        c->out_scale[i] = out_scale[i];
    }
    return SPK_OK;
}

extern "C" int spk_closure_forward(spk_closure c, const void *u, void *tau) {
    SPK_REQUIRE(c && u && tau, SPK_ERR_ARG, "spk_closure_forward: null argument");
    SPK_TRY(forward(c, u, false));
    const double *s = c->out_scale;
    k_unpack_output<<<4096, 256, 0, c->ctx->stream>>>(c->out, (double *)tau, c->n, s[0], s[1], s[2], s[3], s[4], s[5]);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

extern "C" int spk_closure_grads(spk_closure c, const void *u, const void *tau_true, double *loss) {
    SPK_REQUIRE(c && u && tau_true, SPK_ERR_ARG, "spk_closure_grads: null argument");
    hipStream_t st = c->ctx->stream;
    SPK_TRY(forward(c, u, true));
    const double *s = c->out_scale;
    k_prep_target<<<4096, 256, 0, st>>>((const double *)tau_true, c->tgt, c->n, s[0], s[1], s[2], s[3], s[4], s[5]);
    k_loss<<<NBLK, 256, 0, st>>>(c->out, c->tgt, c->dzA, c->P, c->lpart);
    SPK_LAUNCH_CHECK();
    const unsigned g = conv_grid(c);
    // layer 4, then back through layers 3, 2, 1
    SPK_TRY(layer_grads(c, 3, c->A[2], c->dzA));
    k_conv<32, 3><<<g, CONV_T, conv_lds(32), st>>>(c->dzA, c->WT[3], nullptr, c->dzB, nullptr, c->Z[2], nullptr, c->n);
    SPK_TRY(layer_grads(c, 2, c->A[1], c->dzB));
    k_conv<32, 3><<<g, CONV_T, conv_lds(32), st>>>(c->dzB, c->WT[2], nullptr, c->dzA, nullptr, c->Z[1], nullptr, c->n);
    SPK_TRY(layer_grads(c, 1, c->A[0], c->dzA));
    k_conv<32, 3><<<g, CONV_T, conv_lds(32), st>>>(c->dzA, c->WT[1], nullptr, c->dzB, nullptr, c->Z[0], nullptr, c->n);
    SPK_TRY(layer_grads(c, 0, c->xin, c->dzB));
    SPK_LAUNCH_CHECK();
    if (loss) {
        std::vector<float> h(NBLK);
        SPK_HIP(hipMemcpyAsync(h.data(), c->lpart, NBLK * 4, hipMemcpyDeviceToHost, st));
        SPK_HIP(hipStreamSynchronize(st));
        double sum = 0.0;
        for (float v : h) sum += v;
        *loss = sum / (6.0 * (double)c->P);
    }
    return SPK_OK;
}

extern "C" int spk_closure_train_step(spk_closure c, const void *u, const void *tau_true, double lr, double *loss) {
    SPK_TRY(spk_closure_grads(c, u, tau_true, loss));
    const int NP = (int)param_total();
    c->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)c->adam_t), c2 = 1.0f - powf(b2, (float)c->adam_t);
    k_adam<<<(NP + 255) / 256, 256, 0, c->ctx->stream>>>(c->prm, c->grd, c->am, c->av, NP, (float)lr, b1, b2, 1e-8f,
                                                         c1, c2);
    SPK_LAUNCH_CHECK();
    return pack_weights(c);
}

// SGS term of a solver's state after its step: U += dt P(-i m k_j tau_ij(u))
extern "C" int spk_closure_apply(spk_closure c, spk_solver s) {
    SPK_REQUIRE(c && s, SPK_ERR_ARG, "spk_closure_apply: null argument");
    SolverViews v;
    solver_views(s, &v);
    SPK_REQUIRE(v.incompressible && v.nx == c->n && v.ny == c->n && v.nz == c->n, SPK_ERR_ARG,
                "spk_closure_apply: needs an incompressible run on the closure's %d^3 grid", c->n);
    SPK_REQUIRE(v.ctx->stream == c->ctx->stream, SPK_ERR_ARG,
                "spk_closure_apply: the closure and the solver must share a stream");
    const int n = c->n;
    const size_t fs = (size_t)n * n * (n + 2);
    if (!c->tau) {
        if (hipMalloc((void **)&c->tau, fs * 6 * sizeof(double)) != hipSuccess) {
            c->tau = nullptr;
            set_error("spk_closure_apply: hipMalloc of the stress buffer failed");
            return SPK_ERR_ALLOC;
        }
        SPK_TRY(spk_fft_plan_create(c->ctx, n, n, n, SPK_F64, 6, &c->plan6));
    }
    SPK_TRY(spk_complex2real(s));
    SPK_TRY(forward(c, v.realbuf + (size_t)v.VX * fs, false));
    const double *sc = c->out_scale;
    k_unpack_output<<<4096, 256, 0, c->ctx->stream>>>(c->out, c->tau, n, sc[0], sc[1], sc[2], sc[3], sc[4], sc[5]);
    SPK_LAUNCH_CHECK();
    SPK_TRY(spk_fft_r2c(c->plan6, c->tau, c->tau));
    k_sgs_update<<<4096, 256, 0, c->ctx->stream>>>(v.U + (long long)v.VX * v.ntc,
                                                   reinterpret_cast<const cpx<double> *>(c->tau), v.ntc, v.kg, v.ts);
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

// Filtered velocity and subgrid stress on an n^3 grid from a fine N^3 state:
// with F the sharp cut to |k_i| <= kc, kc = (2/3)(n/2 - 1) (the dealiasing
// cut of the coarse grid, wavevector.cu:70-96), ubar = F u and
// tau_ij = F(u_i u_j) - F(ubar_i ubar_j), both products formed on the fine grid.
extern "C" int spk_sgs_stress(spk_ctx ctx, const void *U_hi, int N, void *ubar_lo, void *tau_lo, int n) {
    SPK_REQUIRE(ctx && U_hi && ubar_lo && tau_lo, SPK_ERR_ARG, "spk_sgs_stress: null argument");
    SPK_REQUIRE(n >= 4 && N >= n && N % n == 0, SPK_ERR_ARG, "spk_sgs_stress: N = %d, n = %d", N, n);
    hipStream_t st = ctx->stream;
    const long long ntc_hi = (long long)N * N * (N / 2 + 1), ntc_lo = (long long)n * n * (n / 2 + 1);
    const int kc = (int)(2.0 / 3.0 * (n / 2 - 1));
    const double scale = std::pow((double)n / N, 3);
    double *ur = nullptr, *br = nullptr, *pr = nullptr;
    spk_fft_plan hi3 = nullptr, hi1 = nullptr, lo3 = nullptr, lo1 = nullptr;
    int rc = SPK_OK;
    auto run = [&]() -> int {
        SPK_HIP(hipMalloc((void **)&ur, sizeof(double) * 2 * ntc_hi * 3));
        SPK_HIP(hipMalloc((void **)&br, sizeof(double) * 2 * ntc_hi * 3));
        SPK_HIP(hipMalloc((void **)&pr, sizeof(double) * 2 * ntc_hi));
        SPK_TRY(spk_fft_plan_create(ctx, N, N, N, SPK_F64, 3, &hi3));
        SPK_TRY(spk_fft_plan_create(ctx, N, N, N, SPK_F64, 1, &hi1));
        SPK_TRY(spk_fft_plan_create(ctx, n, n, n, SPK_F64, 3, &lo3));
        SPK_TRY(spk_fft_plan_create(ctx, n, n, n, SPK_F64, 1, &lo1));
        const cpx<double> *U = reinterpret_cast<const cpx<double> *>(U_hi);
        // fine real u, and fine real F u
        SPK_TRY(spk_fft_c2r(hi3, U, ur));
        k_mask_fine<<<4096, 256, 0, st>>>(U, reinterpret_cast<cpx<double> *>(br), N, ntc_hi, 3, kc);
        SPK_LAUNCH_CHECK();
        SPK_TRY(spk_fft_c2r(hi3, br, br));
        // coarse F u
        k_truncate<<<4096, 256, 0, st>>>(U, N, ntc_hi, reinterpret_cast<cpx<double> *>(ubar_lo), n, ntc_lo, 3, kc,
                                         scale);
        SPK_LAUNCH_CHECK();
        SPK_TRY(spk_fft_c2r(lo3, ubar_lo, ubar_lo));
        const int I[6] = {0, 0, 0, 1, 1, 2}, J[6] = {0, 1, 2, 1, 2, 2};
        for (int c = 0; c < 6; c++) {
            k_stress_prod<<<4096, 256, 0, st>>>(ur, br, pr, 2 * ntc_hi, I[c], J[c]);
            SPK_LAUNCH_CHECK();
            SPK_TRY(spk_fft_r2c(hi1, pr, pr));
            double *t = (double *)tau_lo + (size_t)c * 2 * ntc_lo;
            k_truncate<<<4096, 256, 0, st>>>(reinterpret_cast<const cpx<double> *>(pr), N, ntc_hi,
                                             reinterpret_cast<cpx<double> *>(t), n, ntc_lo, 1, kc, scale);
            SPK_LAUNCH_CHECK();
            SPK_TRY(spk_fft_c2r(lo1, t, t));
        }
        SPK_HIP(hipStreamSynchronize(st));
        return SPK_OK;
    };
    rc = run();
    (void)hipStreamSynchronize(st);
    for (spk_fft_plan pl : {hi3, hi1, lo3, lo1})
        if (pl) spk_fft_plan_destroy(pl);
    for (double *b : {ur, br, pr})
        if (b) (void)hipFree(b);
    return rc;
}
