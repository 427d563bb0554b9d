// The AHD walker kernel: walker blocks (renumbered per XCD) develop strips, the blocks after
// them the bilinear border pixels. Included by develop_ahd.hip (single output) and
// develop_ahd_dual.hip (RGBA8 + linear RGB16), which are built with different scheduler options.
#pragma once

#include "ahd_walk.h"

namespace frh {

// Three waves per SIMD (168 VGPRs); the transposed RGBA8 form (OUT 3) holds a row of pixels more
// and runs at two (256 VGPRs), where it does not spill.
template <int RP, bool FAST, bool TAME, int OUT = 0>
__global__ __launch_bounds__(AHD_MAX_WAVES * 64)
__attribute__((amdgpu_waves_per_eu(OUT == 3 ? 2 : 3, OUT == 3 ? 2 : 3))) void k_develop_ahd_walk(DevArgs args, WalkPlan P, uint32_t walk_blocks) {
    const DevArgs &a = kernel_args(args);
    __shared__ Lds s;
    __shared__ float wt[73 * 73];
    load_lds(a, s);
    // the walker blocks come first and are renumbered among themselves; border blocks follow
    const uint32_t blk = blockIdx.x < walk_blocks ? xcd_block(blockIdx.x, walk_blocks) : blockIdx.x;
    if (blk >= walk_blocks) {
        __syncthreads();
        const uint32_t i = (blk - walk_blocks) * blockDim.x + threadIdx.x;
        if (i < P.border.count) border_pixel<false>(a, s, P.border, i);
        return;
    }
    for (int i = threadIdx.x; i < 73 * 73; i += blockDim.x) wt[i] = a.wtab[i];
    __syncthreads();
    // the wave index is the same for every lane; readfirstlane lets the compiler keep all the strip
    // bookkeeping in scalar registers
    const int item = (int)uniform(blk * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (item >= P.strips * P.chunks) return;
    walk_item<RP, FAST, TAME, OUT>(a, P, s, wt, item);
}

}  // namespace frh
