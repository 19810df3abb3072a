// k_integrate_fp64_small.hip -- the fp64 form's integrate kernels for chains of
// up to 64 beads (one wave per chain) and the Cholesky kernels with the factor
// in LDS (the register-factor kernels are in k_integrate_dreg.hip).
#include "bdsim_launch.hpp"

namespace bdgpu {

template <>
void launch_integrate_small<false>(int wpe, int nb, bool, dim3 grid, hipStream_t stream,
                                    const KArgs& a, size_t lws_bytes, size_t) {
    if (lws_bytes > 0 && nb == 0) {
        hipLaunchKernelGGL((k_integrate_lws<false>), grid, dim3(64), lws_bytes, stream, a);
        return;
    }
    if ((nb == 4 || nb == 6) && dense_reg_enabled()) {
        launch_integrate_dreg<false>(nb, grid, stream, a);
        return;
    }
    if (nb == 4) hipLaunchKernelGGL((k_integrate_dense<64, false, 4>), grid, dim3(64), 0, stream, a);
    else if (nb == 6) hipLaunchKernelGGL((k_integrate_dense<64, false, 6>), grid, dim3(64), 0, stream, a);
    else if (wpe == 4) hipLaunchKernelGGL((k_integrate<64, false, 4>), grid, dim3(64), 0, stream, a);
    else if (wpe == 2) hipLaunchKernelGGL((k_integrate<64, false, 2>), grid, dim3(64), 0, stream, a);
    else hipLaunchKernelGGL((k_integrate<64, false, 1>), grid, dim3(64), 0, stream, a);
}

}  // namespace bdgpu
