// spk_fused_f64.hip -- k_fwdx_epilogue instantiations: FP64 bulk, single-box launches.
#include "spk_fused_launch.hpp"

namespace spk {

SPK_FUSED_DECL(launch_fused_f64) {
    return launch_fwdx_epilogue_t<double, false>(ctx, nx, ly, mode, ef, P, U, S, W, fc, tw, ts, ngroups, bt);
}

}  // namespace spk
