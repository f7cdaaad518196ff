// spk_fused_f32.hip -- k_fwdx_epilogue instantiations: FP32 bulk, single-box launches.
#include "spk_fused_launch.hpp"

namespace spk {

SPK_FUSED_DECL(launch_fused_f32) {
    return launch_fwdx_epilogue_t<float, false>(ctx, nx, ly, mode, ef, P, U, S, W, fc, tw, ts, ngroups, bt);
}

}  // namespace spk
