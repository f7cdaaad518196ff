// spk_device.hpp -- device building blocks for the gfx950 pseudospectral solver.
//
// Complex type, constant twiddles, in-register DFTs of radix 2..16, and the
// LDS-exchanged Stockham line FFT used by every transform pass.
//
// Line FFT layout ("natural layout"): a line of N points is held by T = N/E
// threads, thread t owning points t + i*T for i = 0..E-1 in registers v[i].
// Radix-E Stockham passes (Govindaraju et al., SC08, "High performance
// discrete Fourier transforms on graphics processors") run with one LDS
// exchange between passes; a final radix-b pass (b = N / E^a < E) needs no
// exchange because its butterflies fall on each thread's own registers.
// Input and output are both in natural layout, so the first pass reads global
// memory directly and the last pass writes it directly, coalesced across t.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace spk {

template <class T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <class T>
__device__ __forceinline__ cpx<T> mk(T a, T b) { return cpx<T>{a, b}; }
template <class T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <class T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <class T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }

template <class TO, class TI>
__device__ __forceinline__ cpx<TO> cconv(cpx<TI> a) { return {(TO)a.x, (TO)a.y}; }

// cos(2*pi*p/32), p = 0..8, to 20 digits
__host__ __device__ constexpr double cos32(int p) {
    return p == 0 ? 1.0
         : p == 1 ? 0.98078528040323044913
         : p == 2 ? 0.92387953251128675613
         : p == 3 ? 0.83146961230254523708
         : p == 4 ? 0.70710678118654752440
         : p == 5 ? 0.55557023301960222474
         : p == 6 ? 0.38268343236508977173
         : p == 7 ? 0.19509032201612826785
                  : 0.0;
}
// cos and sin of 2*pi*p/32 for any p via octant symmetry
__host__ __device__ constexpr double ccos32(int p) {
    p = ((p % 32) + 32) % 32;
    return p <= 8 ? cos32(p) : p <= 16 ? -cos32(16 - p) : p <= 24 ? -cos32(p - 16) : cos32(32 - p);
}
__host__ __device__ constexpr double csin32(int p) { return ccos32(8 - p); }

// multiply by exp(DIR * 2*pi*i * P/Q) with Q | 32, constant-folded
template <int P, int Q, int DIR, class T>
__device__ __forceinline__ cpx<T> wconst(cpx<T> a) {
    static_assert(32 % Q == 0, "Q must divide 32");
    constexpr int p = (((P * (32 / Q)) % 32) + 32) % 32;  // in units of 2pi/32
    if constexpr (p == 0) {
        return a;
    } else if constexpr (p == 8) {  // exp(DIR i pi/2) = DIR*i
        return DIR > 0 ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
    } else if constexpr (p == 16) {
        return cpx<T>{-a.x, -a.y};
    } else if constexpr (p == 24) {  // exp(DIR i 3pi/2) = -DIR*i
        return DIR > 0 ? cpx<T>{a.y, -a.x} : cpx<T>{-a.y, a.x};
    } else {
        constexpr T c = (T)ccos32(p);
        constexpr T s = (T)(DIR * csin32(p));
        return cpx<T>{a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

// ---- in-register DFTs, natural order in and out: X[k] = sum_n x[n] w^{nk}, w = exp(DIR 2 pi i / R)

template <int DIR, class T>
__device__ __forceinline__ void dft2(cpx<T>& a, cpx<T>& b) {
    cpx<T> t = a;
    a = t + b;
    b = t - b;
}

template <int DIR, class T>
__device__ __forceinline__ void dft4(cpx<T>& x0, cpx<T>& x1, cpx<T>& x2, cpx<T>& x3) {
    cpx<T> a = x0 + x2, b = x0 - x2, c = x1 + x3, d = wconst<1, 4, DIR>(x1 - x3);
    x0 = a + c;
    x2 = a - c;
    x1 = b + d;
    x3 = b - d;
}

template <int DIR, class T>
__device__ __forceinline__ void dft8(cpx<T> (&v)[8]) {
    // n = n1 + 2 n2 (n1<2, n2<4), k = k2 + 4 k1
    dft4<DIR>(v[0], v[2], v[4], v[6]);
    dft4<DIR>(v[1], v[3], v[5], v[7]);
    v[3] = wconst<1, 8, DIR>(v[3]);
    v[5] = wconst<2, 8, DIR>(v[5]);
    v[7] = wconst<3, 8, DIR>(v[7]);
    cpx<T> o[8];
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) {
        o[k2] = v[2 * k2] + v[2 * k2 + 1];
        o[k2 + 4] = v[2 * k2] - v[2 * k2 + 1];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = o[i];
}

template <int DIR, class T>
__device__ __forceinline__ void dft16(cpx<T> (&v)[16]) {
    // n = n1 + 4 n2, k = k2 + 4 k1
#pragma unroll
    for (int n1 = 0; n1 < 4; n1++) dft4<DIR>(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]);
    // now v[n1 + 4 k2] = Y[n1][k2]; twiddle W16^{n1 k2}
    v[5] = wconst<1, 16, DIR>(v[5]);
    v[6] = wconst<2, 16, DIR>(v[6]);
    v[7] = wconst<3, 16, DIR>(v[7]);
    v[9] = wconst<2, 16, DIR>(v[9]);
    v[10] = wconst<4, 16, DIR>(v[10]);
    v[11] = wconst<6, 16, DIR>(v[11]);
    v[13] = wconst<3, 16, DIR>(v[13]);
    v[14] = wconst<6, 16, DIR>(v[14]);
    v[15] = wconst<9, 16, DIR>(v[15]);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) dft4<DIR>(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
    // v[k1 + 4 k2] = X[k2 + 4 k1]: transpose 4x4
    cpx<T> o[16];
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++)
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) o[k2 + 4 * k1] = v[k1 + 4 * k2];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = o[i];
}

// DFT of R values taken from v at indices OFF + r*STRIDE (compile-time), in place
template <int R, int DIR, int OFF, int STRIDE, int E, class T>
__device__ __forceinline__ void dft_group(cpx<T> (&v)[E]) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        dft2<DIR>(v[OFF], v[OFF + STRIDE]);
    } else if constexpr (R == 4) {
        dft4<DIR>(v[OFF], v[OFF + STRIDE], v[OFF + 2 * STRIDE], v[OFF + 3 * STRIDE]);
    } else {
        cpx<T> g[R];
#pragma unroll
        for (int r = 0; r < R; r++) g[r] = v[OFF + r * STRIDE];
        if constexpr (R == 8) dft8<DIR>(g);
        else dft16<DIR>(g);
#pragma unroll
        for (int r = 0; r < R; r++) v[OFF + r * STRIDE] = g[r];
    }
}

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// LDS index with one pad slot every 16 elements (breaks power-of-two strides)
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_len(int n) { return n + n / 16 + 1; }

// LDS placement policies of a line transform: at(i) is the slot of element i,
// len(n) the slots a line of n elements takes.
struct PadSkew {  // one pad slot every 16 elements
    __device__ static __forceinline__ int at(int i) { return lpad(i); }
    __host__ __device__ static constexpr int len(int n) { return lds_len(n); }
};
// No pad: the 16 elements of group g = i / 16 are permuted within the group,
// slot i ^ s(g) with s(g) = 3 (g + (g & 1)) mod 16.  For the access patterns
// of the z rows (several rows of 16-128 points per wavefront, natural-order
// stores and loads, the mirrored loads of the real<->complex split and the
// radix-4 exchanges) this leaves about a quarter of the bank-conflict cycles
// of the skewed layout (64-bank model of 16-byte and 8-byte accesses).
struct PadXor {
    __device__ static __forceinline__ int at(int i) {
        const int g = i >> 4;
        return i ^ ((3 * (g + (g & 1))) & 15);
    }
    __host__ __device__ static constexpr int len(int n) { return (n + 15) & ~15; }
};

// Twiddle table convention: tw[e] = exp(-2 pi i e / N), e = 0..N-1 (forward).
template <int DIR, class T>
__device__ __forceinline__ cpx<T> tw_load(const cpx<T>* __restrict__ tw, int e) {
    cpx<T> w = tw[e];
    if constexpr (DIR > 0) w.y = -w.y;
    return w;
}

// Orders LDS traffic between the threads that share a line.  WAVE: all of
// them sit in one wavefront, whose LDS instructions execute in issue order, so
// a wavefront-scope fence plus a code-motion barrier suffices; otherwise a
// workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void lsync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Stockham line FFT, natural layout in and out.  lds: per-line region of
// lds_len(N) elements.  Caller guarantees every thread of the block (WAVE:
// of the wavefront) executes this (barriers inside).
// Twiddle modes (TWM):
//   0  each twiddle is loaded from the table where it is used;
//   1  each pass loads the next pass's twiddles before its LDS exchange, so
//      their latency hides behind the exchange (E-1 extra complex registers);
//   2  the caller loads all of a thread's twiddles once (load_tw) and passes
//      them to run_tw: a kernel that runs many transforms of one length on
//      the same thread layout reads the table once.
// All twiddles one thread of a line FFT of length N (E points per thread)
// uses, forward convention: w[P-1][r-1] for full pass P >= 1, wt[] for the tail.
template <class T, int N, int E>
struct LineTw {
    static constexpr int NPASS = (E > 1) ? ilog2c(N) / ilog2c(E) : 0;
    static constexpr int NW = E > 1 ? E - 1 : 1;
    cpx<T> w[NPASS > 1 ? NPASS - 1 : 1][NW];
    cpx<T> wt[NW];
};

template <class T, int N, int E, int DIR, bool WAVE = false, int TWM = 0, class PAD = PadSkew>
struct LineFFT {
    static_assert(N % E == 0, "E must divide N");
    static constexpr int TPL = N / E;                  // threads per line
    static constexpr int NPASS = (E > 1) ? ilog2c(N) / ilog2c(E) : 0;  // full radix-E passes
    static constexpr int B = (E > 1) ? N / ipow(E, NPASS) : 1;  // final radix (< E)
    static constexpr int G = (B > 1) ? E / B : 1;       // radix-B groups per thread in the tail
    static constexpr int NW = E > 1 ? E - 1 : 1;        // twiddles per thread per pass

    using Tw = LineTw<T, N, E>;  // conjugated at use for DIR > 0

    __device__ static __forceinline__ cpx<T> dirc(cpx<T> w) {
        if constexpr (DIR > 0) w.y = -w.y;
        return w;
    }
    // twiddle r (1 <= r < E) of full pass P (P >= 1)
    template <int P>
    __device__ static __forceinline__ cpx<T> tw_full(int t, int r, const cpx<T>* __restrict__ tw) {
        constexpr int Ns = ipow(E, P);
        return tw_load<DIR>(tw, (t % Ns) * r * (N / (Ns * E)));
    }
    // twiddle r (1 <= r < B) of tail group M
    __device__ static __forceinline__ cpx<T> tw_tail(int t, int M, int r, const cpx<T>* __restrict__ tw) {
        return tw_load<DIR>(tw, (t + M * TPL) * r);
    }

    template <int P>
    __device__ static __forceinline__ void load_tw_pass(Tw& tr, int t, const cpx<T>* __restrict__ tw) {
        if constexpr (P < NPASS) {
            constexpr int Ns = ipow(E, P);
#pragma unroll
            for (int r = 1; r < E; r++) tr.w[P - 1][r - 1] = tw[(t % Ns) * r * (N / (Ns * E))];
            load_tw_pass<P + 1>(tr, t, tw);
        }
    }
    __device__ static __forceinline__ void load_tw(Tw& tr, int t, const cpx<T>* __restrict__ tw) {
        if constexpr (NPASS > 1) load_tw_pass<1>(tr, t, tw);
        if constexpr (B > 1) {
#pragma unroll
            for (int m = 0; m < G; m++)
#pragma unroll
                for (int r = 1; r < B; r++) tr.wt[m * (B - 1) + r - 1] = tw[(t + m * TPL) * r];
        }
    }

    template <int P>
    __device__ static __forceinline__ void pass_full(cpx<T> (&v)[E], cpx<T>* lds, int t,
                                                     const cpx<T>* __restrict__ tw, const cpx<T> (&w)[NW],
                                                     const Tw* tr) {
        constexpr int Ns = ipow(E, P);
        if constexpr (P > 0) {
#pragma unroll
            for (int r = 1; r < E; r++) {
                cpx<T> wr;
                if constexpr (TWM == 2) wr = dirc(tr->w[P - 1][r - 1]);
                else if constexpr (TWM == 1) wr = w[r - 1];
                else wr = tw_full<P>(t, r, tw);
                v[r] = cmul(v[r], wr);
            }
        }
        dft_group<E, DIR, 0, 1>(v);
        cpx<T> wn[NW];
        if constexpr (TWM == 1) {
            if constexpr (P + 1 < NPASS) {
#pragma unroll
                for (int r = 1; r < E; r++) wn[r - 1] = tw_full<P + 1>(t, r, tw);
            } else if constexpr (B > 1) {
#pragma unroll
                for (int m = 0; m < G; m++)
#pragma unroll
                    for (int r = 1; r < B; r++) wn[m * (B - 1) + r - 1] = tw_tail(t, m, r, tw);
            }
        }
        constexpr bool last = (P == NPASS - 1) && (B == 1);
        if constexpr (!last) {
            const int jm = t % Ns;
            const int idxD = (t / Ns) * Ns * E + jm;
#pragma unroll
            for (int r = 0; r < E; r++) lds[PAD::at(idxD + r * Ns)] = v[r];
            lsync<WAVE>();
#pragma unroll
            for (int i = 0; i < E; i++) v[i] = lds[PAD::at(t + i * TPL)];
            lsync<WAVE>();
        }
        if constexpr (P + 1 < NPASS) pass_full<P + 1>(v, lds, t, tw, wn, tr);
        else if constexpr (B > 1) pass_tail<0>(v, t, tw, wn, tr);
    }

    template <int M>
    __device__ static __forceinline__ void pass_tail(cpx<T> (&v)[E], int t, const cpx<T>* __restrict__ tw,
                                                     const cpx<T> (&w)[NW], const Tw* tr) {
        // radix-B butterflies j = t + M*TPL on registers v[M + r*G]
#pragma unroll
        for (int r = 1; r < B; r++) {
            cpx<T> wr;
            if constexpr (TWM == 2) wr = dirc(tr->wt[M * (B - 1) + r - 1]);
            else if constexpr (TWM == 1) wr = w[M * (B - 1) + r - 1];
            else wr = tw_tail(t, M, r, tw);
            v[M + r * G] = cmul(v[M + r * G], wr);
        }
        dft_group<B, DIR, M, G>(v);
        if constexpr (M + 1 < G) pass_tail<M + 1>(v, t, tw, w, tr);
    }

    __device__ static __forceinline__ void run(cpx<T> (&v)[E], cpx<T>* lds, int t,
                                               const cpx<T>* __restrict__ tw) {
        static_assert(TWM != 2, "TWM 2 transforms take their twiddles through run_tw");
        if constexpr (E > 1) {
            cpx<T> w0[NW];
            pass_full<0>(v, lds, t, tw, w0, nullptr);
        }
    }
    __device__ static __forceinline__ void run_tw(cpx<T> (&v)[E], cpx<T>* lds, int t, const Tw& tr) {
        static_assert(TWM == 2, "run_tw needs TWM 2");
        if constexpr (E > 1) {
            cpx<T> w0[NW];
            pass_full<0>(v, lds, t, nullptr, w0, &tr);
        }
    }
};

// elements per thread for a line of N points
__host__ __device__ constexpr int line_e(int n, int emax) { return n < emax ? n : emax; }

// ---- launch shapes shared by the transform kernels ----
// x/y passes: E=16 points per thread, TPL threads per line, TCOLS columns per block (256 threads)
__host__ __device__ constexpr int strided_e(int n) { return line_e(n, 16); }
__host__ __device__ constexpr int strided_tpl(int n) { return n / strided_e(n); }
__host__ __device__ constexpr int strided_tcols(int n) {
    // 512-thread blocks: 32 adjacent columns at n = 256 (512 B per line point
    // in FP64), wider HBM segments for the strided line accesses
    return strided_tpl(n) >= 512 ? 1 : 512 / strided_tpl(n);
}
// FP32 lines take twice the columns (1024-thread blocks at n = 256), so a line
// point is again a 512-byte contiguous segment
template <class T>
__host__ __device__ constexpr int strided_tcols_t(int n) {
    return sizeof(T) == 4 && strided_tpl(n) <= 16 ? 2 * strided_tcols(n) : strided_tcols(n);
}
template <int N>
__host__ __device__ constexpr bool line_needs_lds() {
    return strided_e(N) != N;
}
// z passes: rows of M = NZ/2 complex points, RW rows per block
__host__ __device__ constexpr int zpass_e(int m) { return line_e(m, 16); }
__host__ __device__ constexpr int zpass_tpl(int m) { return m / zpass_e(m); }
__host__ __device__ constexpr int zpass_rw(int m) { return zpass_tpl(m) >= 256 ? 1 : 256 / zpass_tpl(m); }


// ---- batched launches over boxes of one shape ----
// A batch is a set of boxes that share grid, physics and precision but own
// their arrays and may differ in their diffusion coefficients.  A batched
// launch runs one kernel over every box: blockIdx.z is the box, and each
// pointer argument the launch lists is re-based onto that box's array of the
// same kind (slot) at the same byte offset.  rows == nullptr is a plain
// single-box launch and leaves every argument as passed.
constexpr int BOX_SLOTS = 16;
struct BoxRow {
    char *p[BOX_SLOTS];                   // the box's arrays, by slot
    double nu, nu_m, nu_th, reynolds_ani; // the coefficients a box may set for itself
};
constexpr int BATCH_ARGS = 8;
struct Batch {
    const BoxRow *rows;
    signed char slot[BATCH_ARGS];  // -1: the argument is the same for every box
    long long off[BATCH_ARGS];
};

// argument i of a batched launch for this workgroup's box.  The address read
// from the table is marked as global memory: a pointer loaded from memory is
// otherwise generic, and accesses through it would be flat loads and stores.
template <class P>
__device__ __forceinline__ P rebase(const Batch &b, int i, P p) {
    if (b.rows == nullptr || b.slot[i] < 0) return p;
    using E = __attribute__((address_space(1))) std::remove_pointer_t<P>;
    using GC = __attribute__((address_space(1))) char;
    GC *g = (GC *)(b.rows[blockIdx.z].p[b.slot[i]]) + b.off[i];
    return (P)(E *)g;
}
__device__ __forceinline__ const BoxRow *batch_row(const Batch &b) {
    return b.rows == nullptr ? nullptr : b.rows + blockIdx.z;
}

}  // namespace spk
