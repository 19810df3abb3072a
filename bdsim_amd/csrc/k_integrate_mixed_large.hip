// k_integrate_mixed_large.hip -- the mixed form's integrate kernels for chains of
// 65..1024 beads (2..16 waves per chain) and for chains split over several
// workgroups.
#include "bdsim_launch.hpp"

namespace bdgpu {

template <>
void launch_integrate_large<true>(int block, int wpe, dim3 grid, hipStream_t stream, const KArgs& a) {
    const bool four = wpe == 4;
    switch (block) {
        case 128:
            if (four) hipLaunchKernelGGL((k_integrate<128, true, 4>), grid, dim3(128), 0, stream, a);
            else hipLaunchKernelGGL((k_integrate<128, true, 1>), grid, dim3(128), 0, stream, a);
            break;
        case 256:
            if (four) hipLaunchKernelGGL((k_integrate<256, true, 4>), grid, dim3(256), 0, stream, a);
            else hipLaunchKernelGGL((k_integrate<256, true, 1>), grid, dim3(256), 0, stream, a);
            break;
        case 512:
            if (four) hipLaunchKernelGGL((k_integrate<512, true, 4>), grid, dim3(512), 0, stream, a);
            else hipLaunchKernelGGL((k_integrate<512, true, 1>), grid, dim3(512), 0, stream, a);
            break;
        default: hipLaunchKernelGGL((k_integrate<1024, true, 1>), grid, dim3(1024), 0, stream, a); break;
    }
}

template <>
hipError_t launch_split_kernel<true>(dim3 grid, hipStream_t stream, KArgs& a) {
    void* args[] = {&a};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_integrate_split<true>), grid,
                                      dim3(1024), args, 0, stream);
}

}  // namespace bdgpu
