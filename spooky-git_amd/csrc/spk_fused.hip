// spk_fused.hip -- the launcher of the fused forward x pass + spectral
// epilogue + RK3 update kernel (k_fwdx_epilogue): picks the instantiation set
// (precision; batched launch or not) compiled in spk_fused_*.hip.
#include "spk_fused_launch.hpp"

namespace spk {

bool fused_supported_nx(int nx) { return nx >= 32 && nx <= 1024 && is_pow2(nx); }

// ef: points per thread of the x line, 4 or 8 (picked by the solver's tile autotune)
int launch_fwdx_epilogue(spk_ctx ctx, int bulk, int nx, int ly, int mode, int ef, const void *P, void *U, void *S,
                         void *W, const FusedConst &fc, const void *tw, const TimeState *ts, int ngroups) {
    Batch bt;
    SPK_TRY(make_batch(ctx, {P, U, S, W, ts}, &bt));
    if (bulk == SPK_F64)
        return bt.rows ? launch_fused_f64_batched(ctx, nx, ly, mode, ef, P, U, S, W, fc, tw, ts, ngroups, bt)
                       : launch_fused_f64(ctx, nx, ly, mode, ef, P, U, S, W, fc, tw, ts, ngroups, bt);
    return bt.rows ? launch_fused_f32_batched(ctx, nx, ly, mode, ef, P, U, S, W, fc, tw, ts, ngroups, bt)
                   : launch_fused_f32(ctx, nx, ly, mode, ef, P, U, S, W, fc, tw, ts, ngroups, bt);
}

}  // namespace spk
