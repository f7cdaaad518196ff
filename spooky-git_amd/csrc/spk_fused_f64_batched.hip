// spk_fused_f64_batched.hip -- k_fwdx_epilogue instantiations: FP64 bulk, batched launches.
#include "spk_fused_launch.hpp"

namespace spk {

SPK_FUSED_DECL(launch_fused_f64_batched) {
    return launch_fwdx_epilogue_t<double, true>(ctx, nx, ly, mode, ef, P, U, S, W, fc, tw, ts, ngroups, bt);
}

}  // namespace spk
