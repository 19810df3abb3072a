// bdsim_launch.hpp -- launches of the integrate kernels, one definition per
// precision form and chain-size class in its own translation unit
// (k_integrate_{mixed,fp64}_{small,large}.hip), so that they compile in
// parallel.
#pragma once
#include "bdsim_chain.hpp"

#include <cstdlib>

namespace bdgpu {

// 64-thread chains: k_integrate<64, MIXED, wpe> (wpe = waves per SIMD the
// register allocation is held to: 1, 2 or 4), or with nb = 4 / 6 the Cholesky
// kernel with its factor in LDS, k_integrate_dense<64, MIXED, nb, FT>, FT
// float for the mixed form's FP32-factor variant (f32_factor), else double.
// lws_bytes > 0: k_integrate_lws<MIXED> with that much dynamic LDS for the
// hot part of the workspace (few chains; nb = 0). ptab_bytes > 0 (mixed form,
// nb = 0): that much more dynamic LDS for the step's pair-weight table, after
// the workspace part (k_integrate_lws<true, true>) or alone
// (k_integrate_ptab<wpe>).
template <bool MIXED>
void launch_integrate_small(int wpe, int nb, bool f32_factor, dim3 grid, hipStream_t stream,
                            const KArgs& a, size_t lws_bytes = 0, size_t ptab_bytes = 0);

// Cholesky chains with the FP64 factor in registers: k_integrate_dreg<MIXED,
// nb>, nb = 4 or 6 (k_integrate_dreg.hip).
template <bool MIXED>
void launch_integrate_dreg(int nb, dim3 grid, hipStream_t stream, const KArgs& a);

// The register factor is the default for FP64 factors; BDSIM_DENSE_REG=0 in
// the environment keeps them in LDS (k_integrate_dense).
inline bool dense_reg_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("BDSIM_DENSE_REG");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Multi-wave chains: k_integrate<block, MIXED, wpe>, block 128..1024, wpe 1 or
// 4 (1024: 1).
template <bool MIXED>
void launch_integrate_large(int block, int wpe, dim3 grid, hipStream_t stream, const KArgs& a);

// Chains split over a.split workgroups each: k_integrate_split<MIXED>,
// launched cooperatively (all parts resident together).
template <bool MIXED>
hipError_t launch_split_kernel(dim3 grid, hipStream_t stream, KArgs& a);

}  // namespace bdgpu
