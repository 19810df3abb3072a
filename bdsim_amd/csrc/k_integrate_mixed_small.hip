// k_integrate_mixed_small.hip -- the mixed form's integrate kernels for chains of
// up to 64 beads (one wave per chain) and the Cholesky kernels with the factor
// in LDS (the register-factor kernels are in k_integrate_dreg.hip).
#include "bdsim_launch.hpp"

namespace bdgpu {

template <>
void launch_integrate_small<true>(int wpe, int nb, bool f32_factor, dim3 grid, hipStream_t stream,
                                   const KArgs& a, size_t lws_bytes, size_t ptab_bytes) {
    if (lws_bytes > 0 && nb == 0) {
        // ptab_bytes > 0: the table follows the workspace part in the same
        // dynamic LDS
        if (ptab_bytes > 0)
            hipLaunchKernelGGL((k_integrate_lws<true, true>), grid, dim3(64), lws_bytes + ptab_bytes, stream, a);
        else
            hipLaunchKernelGGL((k_integrate_lws<true>), grid, dim3(64), lws_bytes, stream, a);
        return;
    }
    if (ptab_bytes > 0 && nb == 0) {
        if (wpe == 1) hipLaunchKernelGGL((k_integrate_ptab<1>), grid, dim3(64), ptab_bytes, stream, a);
        else hipLaunchKernelGGL((k_integrate_ptab<2>), grid, dim3(64), ptab_bytes, stream, a);
        return;
    }
    if ((nb == 4 || nb == 6) && !f32_factor && dense_reg_enabled()) {
        launch_integrate_dreg<true>(nb, grid, stream, a);
        return;
    }
    if (nb == 4 && f32_factor) hipLaunchKernelGGL((k_integrate_dense<64, true, 4, float>), grid, dim3(64), 0, stream, a);
    else if (nb == 6 && f32_factor) hipLaunchKernelGGL((k_integrate_dense<64, true, 6, float>), grid, dim3(64), 0, stream, a);
    else if (nb == 4) hipLaunchKernelGGL((k_integrate_dense<64, true, 4>), grid, dim3(64), 0, stream, a);
    else if (nb == 6) hipLaunchKernelGGL((k_integrate_dense<64, true, 6>), grid, dim3(64), 0, stream, a);
    else if (wpe == 4) hipLaunchKernelGGL((k_integrate<64, true, 4>), grid, dim3(64), 0, stream, a);
    else if (wpe == 2) hipLaunchKernelGGL((k_integrate<64, true, 2>), grid, dim3(64), 0, stream, a);
    else hipLaunchKernelGGL((k_integrate<64, true, 1>), grid, dim3(64), 0, stream, a);
}

}  // namespace bdgpu
