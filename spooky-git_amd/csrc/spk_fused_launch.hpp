// spk_fused_launch.hpp -- launcher templates of the fused forward x pass +
// spectral epilogue + RK3 update kernel (k_fwdx_epilogue).  The instantiations
// are spread over several translation units (precision x batched or not) so
// that they compile in parallel: spk_fused*.hip.
#pragma once

#include "spk_internal.hpp"
#include "spk_kernels.hpp"

namespace spk {

template <class T, int N, int LY, int EF, bool BAT, bool STRAT = false>
static void launch_ly_mode(int mode, dim3 grid, int bs, hipStream_t st, const cpx<T> *P, cpx<double> *U,
                           cpx<T> *S, cpx<T> *W, const FusedConst &fc, const cpx<T> *tw, const TimeState *ts,
                           const Batch &bt) {
    switch (mode) {
        case 1: k_fwdx_epilogue<T, N, LY, 1, EF, BAT, STRAT><<<grid, bs, 0, st>>>(P, U, S, W, fc, tw, ts, bt); break;
        case 2: k_fwdx_epilogue<T, N, LY, 2, EF, BAT, STRAT><<<grid, bs, 0, st>>>(P, U, S, W, fc, tw, ts, bt); break;
        default: k_fwdx_epilogue<T, N, LY, 3, EF, BAT, STRAT><<<grid, bs, 0, st>>>(P, U, S, W, fc, tw, ts, bt); break;
    }
}

// Boussinesq layouts: stratified runs couple the velocity and temperature
// groups in one block (k_fwdx_epilogue's STRAT instantiation)
template <class T, int N, int LY, int EF, bool BAT>
static int launch_bouss(int mode, dim3 grid, int bs, hipStream_t st, const cpx<T> *P, cpx<double> *U, cpx<T> *S,
                        cpx<T> *W, const FusedConst &fc, const cpx<T> *tw, const TimeState *ts, const Batch &bt) {
    if (!fc.e.stratification) {
        launch_ly_mode<T, N, LY, EF, BAT>(mode, grid, bs, st, P, U, S, W, fc, tw, ts, bt);
        return SPK_OK;
    }
    launch_ly_mode<T, N, LY, EF, BAT, true>(mode, grid, bs, st, P, U, S, W, fc, tw, ts, bt);
    return SPK_OK;
}

template <class T, int N, int EF, bool BAT>
static int launch_n(int ly, int mode, hipStream_t st, const void *P, void *U, void *S, void *W, const FusedConst &fc,
                    const void *tw, const TimeState *ts, int ngroups, const Batch &bt, unsigned nbz) {
    constexpr int TC = fx_tcols(N, EF);
    const long long ncols = (long long)fc.e.kg.ny * fc.e.kg.nzh;
    dim3 grid((unsigned)((ncols + TC - 1) / TC), ngroups, nbz);
    const int bs = TC * fx_tpl(N, EF);
    const cpx<T> *p = (const cpx<T> *)P;
    cpx<double> *u = (cpx<double> *)U;
    cpx<T> *s = (cpx<T> *)S;  // the RK3 increment register, bulk precision
    cpx<T> *w = (cpx<T> *)W;
    const cpx<T> *t = (const cpx<T> *)tw;
    switch (ly) {
        case LY_HD: launch_ly_mode<T, N, LY_HD, EF, BAT>(mode, grid, bs, st, p, u, s, w, fc, t, ts, bt); break;
        case LY_HD_TH: SPK_TRY((launch_bouss<T, N, LY_HD_TH, EF, BAT>(mode, grid, bs, st, p, u, s, w, fc, t, ts, bt))); break;
        case LY_MHD: launch_ly_mode<T, N, LY_MHD, EF, BAT>(mode, grid, bs, st, p, u, s, w, fc, t, ts, bt); break;
        case LY_MHD_TH: SPK_TRY((launch_bouss<T, N, LY_MHD_TH, EF, BAT>(mode, grid, bs, st, p, u, s, w, fc, t, ts, bt))); break;
        default: set_error("fused epilogue: layout %d has no FFT products", ly); return SPK_ERR_STATE;
    }
    SPK_LAUNCH_CHECK();
    return SPK_OK;
}

template <class T, int EF, bool BAT>
static int launch_t(int nx, int ly, int mode, hipStream_t st, const void *P, void *U, void *S, void *W,
                    const FusedConst &fc, const void *tw, const TimeState *ts, int ngroups, const Batch &bt,
                    unsigned nbz) {
    switch (nx) {
        case 32: return launch_n<T, 32, EF, BAT>(ly, mode, st, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
        case 64: return launch_n<T, 64, EF, BAT>(ly, mode, st, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
        case 128: return launch_n<T, 128, EF, BAT>(ly, mode, st, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
        case 256: return launch_n<T, 256, EF, BAT>(ly, mode, st, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
        case 512: return launch_n<T, 512, EF, BAT>(ly, mode, st, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
        case 1024: return launch_n<T, 1024, EF, BAT>(ly, mode, st, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
        default: set_error("fused epilogue: NX=%d not instantiated", nx); return SPK_ERR_ARG;
    }
}

// the instantiations of one precision, batched or not; ef 4 or 8
template <class T, bool BAT>
int launch_fwdx_epilogue_t(spk_ctx ctx, int nx, int ly, int mode, int ef, const void *P, void *U, void *S, void *W,
                           const FusedConst &fc, const void *tw, const TimeState *ts, int ngroups, const Batch &bt) {
    const unsigned nbz = nboxes(ctx);
    if (ef == 8) return launch_t<T, 8, BAT>(nx, ly, mode, ctx->stream, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
    return launch_t<T, 4, BAT>(nx, ly, mode, ctx->stream, P, U, S, W, fc, tw, ts, ngroups, bt, nbz);
}

#define SPK_FUSED_DECL(NAME)                                                                                    \
    int NAME(spk_ctx ctx, int nx, int ly, int mode, int ef, const void *P, void *U, void *S, void *W,           \
             const FusedConst &fc, const void *tw, const TimeState *ts, int ngroups, const Batch &bt)
SPK_FUSED_DECL(launch_fused_f64);
SPK_FUSED_DECL(launch_fused_f32);
SPK_FUSED_DECL(launch_fused_f64_batched);
SPK_FUSED_DECL(launch_fused_f32_batched);

}  // namespace spk
