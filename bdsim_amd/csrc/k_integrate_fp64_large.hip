// k_integrate_fp64_large.hip -- the fp64 form's integrate kernels for chains of
// 65..1024 beads (2..16 waves per chain) and for chains split over several
// workgroups.
#include "bdsim_launch.hpp"

namespace bdgpu {

template <>
void launch_integrate_large<false>(int block, int wpe, dim3 grid, hipStream_t stream, const KArgs& a) {
    const bool four = wpe == 4;
    switch (block) {
        case 128:
            if (four) hipLaunchKernelGGL((k_integrate<128, false, 4>), grid, dim3(128), 0, stream, a);
            else hipLaunchKernelGGL((k_integrate<128, false, 1>), grid, dim3(128), 0, stream, a);
            break;
        case 256:
            if (four) hipLaunchKernelGGL((k_integrate<256, false, 4>), grid, dim3(256), 0, stream, a);
            else hipLaunchKernelGGL((k_integrate<256, false, 1>), grid, dim3(256), 0, stream, a);
            break;
        case 512:
            if (four) hipLaunchKernelGGL((k_integrate<512, false, 4>), grid, dim3(512), 0, stream, a);
            else hipLaunchKernelGGL((k_integrate<512, false, 1>), grid, dim3(512), 0, stream, a);
            break;
        default: hipLaunchKernelGGL((k_integrate<1024, false, 1>), grid, dim3(1024), 0, stream, a); break;
    }
}

template <>
hipError_t launch_split_kernel<false>(dim3 grid, hipStream_t stream, KArgs& a) {
    void* args[] = {&a};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_integrate_split<false>), grid,
                                      dim3(1024), args, 0, stream);
}

}  // namespace bdgpu
