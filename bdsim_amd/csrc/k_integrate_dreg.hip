// k_integrate_dreg.hip -- the Cholesky kernels with the FP64 factor held in
// registers (both precision forms), in a translation unit of their own.
#include "bdsim_launch.hpp"

namespace bdgpu {

// The register budget: two waves per SIMD (the default: 8 chains per CU;
// measured 1.9x (mixed) / 1.4x (FP64) the one-wave build on 8192 chains of
// N = 30, despite some spilling) or one (BDSIM_DREG_WPE=1 in the environment).
static int dreg_wpe() {
    static const int w = [] {
        const char* e = std::getenv("BDSIM_DREG_WPE");
        return (e && e[0] == '1') ? 1 : 2;
    }();
    return w;
}

template <bool MIXED>
void launch_integrate_dreg(int nb, dim3 grid, hipStream_t stream, const KArgs& a) {
    const bool two = dreg_wpe() == 2;
    if (nb == 4 && two) hipLaunchKernelGGL((k_integrate_dreg<MIXED, 4, 2>), grid, dim3(64), 0, stream, a);
    else if (nb == 4) hipLaunchKernelGGL((k_integrate_dreg<MIXED, 4, 1>), grid, dim3(64), 0, stream, a);
    else if (two) hipLaunchKernelGGL((k_integrate_dreg<MIXED, 6, 2>), grid, dim3(64), 0, stream, a);
    else hipLaunchKernelGGL((k_integrate_dreg<MIXED, 6, 1>), grid, dim3(64), 0, stream, a);
}

template void launch_integrate_dreg<true>(int, dim3, hipStream_t, const KArgs&);
template void launch_integrate_dreg<false>(int, dim3, hipStream_t, const KArgs&);

}  // namespace bdgpu
