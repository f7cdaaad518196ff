// spk_internal.hpp -- host-side internals shared by the libspk translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <initializer_list>
#include <string>

#include "../../include/spk.h"
#include "spk_device.hpp"

struct spk_ctx_s {
    int device;
    hipStream_t stream;
    // the caller's handle plus one per solver, plan and closure made on this
    // context: spk_ctx_destroy drops the caller's, and the context is deleted
    // when the last holder lets go (a solver destroyed after its context still
    // synchronizes on a live stream)
    std::atomic<int> refs{1};
    // When set, transform kernels launched on this context return at once unless
    // *gate is non-zero (device-sized loops: the stages of an RKL2 cycle beyond
    // the stage count the device computed for this step)
    const int *gate = nullptr;
    // Batched launches (spk_ensemble): while rows is set, every kernel launched on
    // this context runs over nbox boxes of one shape.  base/size are the arrays
    // of the box whose addresses the launch sites pass, slot by slot, so that a
    // pointer argument maps to its slot and byte offset (make_batch).
    const spk::BoxRow *rows = nullptr;
    int nbox = 1;
    const char *base[spk::BOX_SLOTS] = {};
    size_t size[spk::BOX_SLOTS] = {};
};

namespace spk {

// thread-local error text behind spk_last_error()
void set_error(const char *fmt, ...);

// context reference counting (spk_core.hip)
spk_ctx ctx_retain(spk_ctx ctx);
void ctx_release(spk_ctx ctx);

// The Batch of a launch on ctx whose pointer arguments are ptrs, in the order
// the kernel re-bases them: each non-null pointer is mapped to the slot of the
// array it lies in and its byte offset there.  Outside a batched run the result
// is a plain launch.  A pointer in none of the arrays is an error (the launch
// would touch one box only).
int make_batch(spk_ctx ctx, std::initializer_list<const void *> ptrs, Batch *out);
// blockIdx.z extent of a launch on ctx
inline unsigned nboxes(spk_ctx ctx) { return ctx->rows ? (unsigned)ctx->nbox : 1u; }

// Evaluates a HIP call; on failure records file:line and the HIP message and
// returns SPK_ERR_HIP from the enclosing function.
#define SPK_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            spk::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,                \
                           hipGetErrorString(e_));                                           \
            return SPK_ERR_HIP;                                                              \
        }                                                                                    \
    } while (0)

// Checks the launch that was just issued.
#define SPK_LAUNCH_CHECK() SPK_HIP(hipGetLastError())

#define SPK_REQUIRE(cond, code, ...)        \
    do {                                    \
        if (!(cond)) {                      \
            spk::set_error(__VA_ARGS__);    \
            return (code);                  \
        }                                   \
    } while (0)

#define SPK_TRY(expr)              \
    do {                           \
        int r_ = (expr);           \
        if (r_ != SPK_OK) return r_; \
    } while (0)

inline bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

// Maps a block id so that the blocks one XCD receives (round-robin dispatch,
// 8 XCDs) cover a contiguous range of logical tiles: neighbouring tiles share
// that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, i = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// ---- transform geometry ----

struct Grid3 {
    int nx, ny, nz, nzh;
    long long ntc;  // nx*ny*nzh
};

inline Grid3 make_grid(int nx, int ny, int nz) {
    Grid3 g;
    g.nx = nx; g.ny = ny; g.nz = nz; g.nzh = nz / 2 + 1;
    g.ntc = (long long)nx * ny * g.nzh;
    return g;
}

// Twiddle tables for one grid: exp(-2 pi i e / n) for n = nx, ny, nz, in the
// bulk precision, contiguous: [nx][ny][nz].
struct Twiddles {
    void *d = nullptr;  // device
    int prec = SPK_F64;
    long long off_x = 0, off_y = 0, off_z = 0;
};

int twiddles_create(spk_ctx ctx, const Grid3 &g, int prec, Twiddles *tw);
void twiddles_destroy(Twiddles *tw);

// ---- pass launchers (spk_fft.hip) ----
// One field per blockIdx.y; field strides are in complex elements.
struct PassArgs {
    const void *in;
    void *out;
    int nfields;
    long long fstride_in, fstride_out;  // complex elements (real rows: 2*NTC reals = NTC complex)
    double scale;                        // multiplies the input (x inverse pass: 1/N)
};

// x-direction (stride ny*nzh) C2C pass, dir -1 forward / +1 inverse; in/out
// precision tin/tout, compute precision = tcomp (must equal the twiddle precision)
int launch_pass_x(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int dir, int tin, int tout,
                  const PassArgs &a);
int launch_pass_y(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int dir, int tin, int tout,
                  const PassArgs &a);
// z R2C (real padded rows -> complex rows) and C2R (complex rows -> real padded rows)
int launch_pass_z_r2c(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int tin, int tout,
                      const PassArgs &a);
int launch_pass_z_c2r(spk_ctx ctx, const Grid3 &g, const Twiddles &tw, int tin, int tout,
                      const PassArgs &a);

// looks up, with `device` current, the address of that device's copy of the
// static scratch array behind spk_energy_noalloc (spk_solver.hip); once per device
int noalloc_setup(int device);

// solver internals the ensemble needs (spk_solver.hip)
spk_ctx solver_ctx(spk_solver s);
int solver_host_step(spk_solver s);
void solver_advance_host(spk_solver s, int n);
// batch slots of a solver's arrays
enum : int {
    SL_U = 0, SL_S, SL_DU, SL_REAL, SL_W, SL_P, SL_W2, SL_P2, SL_BREAL, SL_TS, SL_PMAX, SL_UC0, SL_UC1, SL_DU0,
    SL_PLAN, SL_COUNT
};
static_assert(SL_COUNT <= BOX_SLOTS, "batch slots");
// the row of s in a batch table (arrays by slot, coefficients) and the byte
// size of each array
void solver_box_row(spk_solver s, BoxRow *row, size_t *sizes);
// a and b can run in one batch: same grid, physics, precision, back end and
// every parameter but nu, nu_m, nu_th and reynolds_ani
bool solver_same_shape(spk_solver a, spk_solver b);
// s takes the back end (fused or not, points per thread) that lead runs
int solver_follow_back_end(spk_solver s, spk_solver lead);
// whether W holds the inverse x pass the last fused step of s left (the host's
// record of it), and setting that record (for steps replayed from a graph)
bool solver_w_ok(spk_solver s);
void solver_set_w_ok(spk_solver s, bool v);
// a fused step of s leaves W holding the next step's inverse x pass (fused back
// end, no forcing, no super-time-stepped u or B)
bool solver_continues(spk_solver s);

}  // namespace spk
