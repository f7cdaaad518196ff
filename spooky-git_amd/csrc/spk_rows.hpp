// spk_rows.hpp -- real<->half-complex transforms of one z row held by TPL threads.
//
// A real row x[0..NZ-1] (NZ = 2M) is handled as the complex sequence
// z[n] = x[2n] + i x[2n+1], n < M, transformed by a complex FFT of length M.
// Forward post-processing (k = 0..M, Z[M] == Z[0]):
//   X[k] = E[k] + W^k O[k],  E = (Z[k] + conj Z[M-k]) / 2,  O = (Z[k] - conj Z[M-k]) / 2i,
//   W = exp(-2 pi i / NZ).
// Inverse pre-processing (unnormalized C2R, imaginary parts of X[0], X[M] dropped):
//   Z'[k] = (X[k] + conj X[M-k]) + i W^{-k} (X[k] - conj X[M-k]),  z = IDFT_M(Z').
// Thread t of a row owns z[t + i*TPL], i < E, in both directions, i.e. the real
// points 2n and 2n+1 for those n: a C2R followed by an R2C on the same row
// needs no data movement between them.
#pragma once

#include "spk_device.hpp"

namespace spk {

// REG: the kernel loads a thread's twiddles once (load_tw) and hands them to
// every transform of the row (c2r_regs / r2c with a RowTw); otherwise each
// twiddle is read from the tables where it is used.
template <class T, int M, int E, bool REG = false>
struct ZRow {
    static constexpr int TPL = M / E;
    using PAD = PadXor;
    static constexpr int LL = PAD::len(M + 1);  // LDS elements per row
    // the threads of a row share one wavefront: wavefront-level ordering only
    static constexpr bool WAVE = TPL <= 64;
    using FFTI = LineFFT<T, M, E, +1, WAVE, REG ? 2 : 0, PAD>;
    using FFTF = LineFFT<T, M, E, -1, WAVE, REG ? 2 : 0, PAD>;

    // a thread's twiddles: the length-M FFT's and W^k = exp(-2 pi i k / NZ) of
    // its points k = t + i*TPL for the split pre/post-processing
    struct RowTw {
        LineTw<T, M, E> f;
        cpx<T> wz[E];
    };
    __device__ static __forceinline__ void load_tw(RowTw &r, int t, const cpx<T> *__restrict__ tw_nz,
                                                   const cpx<T> *__restrict__ tw_m) {
        FFTF::load_tw(r.f, t, tw_m);
#pragma unroll
        for (int i = 0; i < E; i++) r.wz[i] = tw_nz[t + i * TPL];
    }

    // Global loads of a complex row src[0..M] into registers: v[i] = src[t + i*TPL],
    // nyq = src[M] (thread 0 only).  Kept apart from the transform so a kernel can
    // issue the loads of several rows before the first LDS fence.
    template <class TI>
    __device__ static __forceinline__ void load(const cpx<TI> *src, bool active, int t, cpx<T> (&v)[E], cpx<T> &nyq) {
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = active ? cconv<T>(src[t + i * TPL]) : cpx<T>{(T)0, (T)0};
        nyq = (active && t == 0) ? cconv<T>(src[M]) : cpx<T>{(T)0, (T)0};
    }

    // complex row src[0..M] -> z values v (natural layout).  All threads of the
    // block call this; inactive rows pass active=false.
    template <class TI>
    __device__ static __forceinline__ void c2r(const cpx<TI> *src, bool active, cpx<T> *L, int t,
                                               const cpx<T> *__restrict__ tw_nz,
                                               const cpx<T> *__restrict__ tw_m, T scale, cpx<T> (&v)[E]) {
        cpx<T> nyq;
        load(src, active, t, v, nyq);
        c2r_regs(v, nyq, active, L, t, tw_nz, tw_m, scale);
    }

    // as c2r, from a row already loaded by load(): v in = row values, out = z values
    __device__ static __forceinline__ void c2r_regs(cpx<T> (&v)[E], cpx<T> nyq, bool active, cpx<T> *L, int t,
                                                    const cpx<T> *__restrict__ tw_nz,
                                                    const cpx<T> *__restrict__ tw_m, T scale) {
        static_assert(!REG, "register twiddles: use c2r_tw");
        c2r_impl(v, nyq, active, L, t, tw_nz, tw_m, nullptr, scale);
    }
    __device__ static __forceinline__ void c2r_tw(cpx<T> (&v)[E], cpx<T> nyq, bool active, cpx<T> *L, int t,
                                                  const RowTw &rt, T scale) {
        static_assert(REG, "c2r_tw needs REG");
        c2r_impl(v, nyq, active, L, t, nullptr, nullptr, &rt, scale);
    }
    __device__ static __forceinline__ void c2r_impl(cpx<T> (&v)[E], cpx<T> nyq, bool active, cpx<T> *L, int t,
                                                    const cpx<T> *__restrict__ tw_nz,
                                                    const cpx<T> *__restrict__ tw_m, const RowTw *rt, T scale) {
        if (active) {
#pragma unroll
            for (int i = 0; i < E; i++) {
                const int k = t + i * TPL;
                cpx<T> a = cscale(v[i], scale);
                if (k == 0) a.y = (T)0;
                L[PAD::at(k)] = a;
                v[i] = a;  // X[k] stays in registers; only X[M-k] comes back through LDS
            }
            if (t == 0) {
                cpx<T> a = cscale(nyq, scale);
                a.y = (T)0;
                L[PAD::at(M)] = a;
            }
        }
        lsync<WAVE>();
#pragma unroll
        for (int i = 0; i < E; i++) {
            const int k = t + i * TPL;
            const cpx<T> xk = v[i], xm = L[PAD::at(M - k)];
            const cpx<T> s = {xk.x + xm.x, xk.y - xm.y};
            const cpx<T> d = {xk.x - xm.x, xk.y + xm.y};
            cpx<T> w;
            if constexpr (REG) w = rt->wz[i];
            else w = tw_nz[k];
            const cpx<T> wd = cmulc(d, w);  // W^{-k} d
            v[i] = cpx<T>{s.x - wd.y, s.y + wd.x};  // s + i W^{-k} d
        }
        lsync<WAVE>();
        if constexpr (REG) FFTI::run_tw(v, L, t, rt->f);
        else FFTI::run(v, L, t, tw_m);
    }

    // z values v (natural layout) -> complex row dst[0..M].  Ends with a barrier
    // so the caller may reuse L immediately.
    template <class TO>
    __device__ static __forceinline__ void r2c(cpx<T> (&v)[E], bool active, cpx<TO> *dst, cpx<T> *L, int t,
                                               const cpx<T> *__restrict__ tw_nz,
                                               const cpx<T> *__restrict__ tw_m) {
        static_assert(!REG, "register twiddles: use r2c_tw");
        r2c_impl(v, active, dst, L, t, tw_nz, tw_m, nullptr);
    }
    template <class TO>
    __device__ static __forceinline__ void r2c_tw(cpx<T> (&v)[E], bool active, cpx<TO> *dst, cpx<T> *L, int t,
                                                  const RowTw &rt) {
        static_assert(REG, "r2c_tw needs REG");
        r2c_impl(v, active, dst, L, t, nullptr, nullptr, &rt);
    }
    template <class TO>
    __device__ static __forceinline__ void r2c_impl(cpx<T> (&v)[E], bool active, cpx<TO> *dst, cpx<T> *L, int t,
                                                    const cpx<T> *__restrict__ tw_nz,
                                                    const cpx<T> *__restrict__ tw_m, const RowTw *rt) {
        if constexpr (REG) FFTF::run_tw(v, L, t, rt->f);
        else FFTF::run(v, L, t, tw_m);
#pragma unroll
        for (int i = 0; i < E; i++) L[PAD::at(t + i * TPL)] = v[i];
        lsync<WAVE>();
        if (active) {
            const T h = (T)0.5;
#pragma unroll
            for (int i = 0; i < E; i++) {
                const int k = t + i * TPL;
                const cpx<T> zk = v[i], zm = L[PAD::at((M - k) & (M - 1))];
                const cpx<T> ev = {h * (zk.x + zm.x), h * (zk.y - zm.y)};
                const cpx<T> od = {h * (zk.y + zm.y), -h * (zk.x - zm.x)};
                cpx<T> w;
                if constexpr (REG) w = rt->wz[i];
                else w = tw_nz[k];
                dst[k] = cconv<TO>(ev + cmul(od, w));
            }
            if (t == 0) {
                const cpx<T> z0 = L[PAD::at(0)];
                dst[M] = cconv<TO>(cpx<T>{z0.x - z0.y, (T)0});
            }
        }
        lsync<WAVE>();
    }
};

}  // namespace spk
