// AHD develop launcher: the register row walker of ahd_walk.h, one strip per wave, plus the
// bilinear border pixels in extra blocks of the same launch.
#include "ahd_kernel.h"
#include "frh_launch.h"

namespace frh {

// The two-output (RGBA8 + linear RGB16) form is built in develop_ahd_dual.hip.
void launch_ahd_dual(int rp, const DevArgs &a, const WalkPlan &P, uint32_t walk_blocks, uint32_t grid, uint32_t block,
                     hipStream_t stream);

template <int RP>
static void launch_ahd_rp(const DevArgs &a, const WalkPlan &P, uint32_t walk_blocks, uint32_t grid, uint32_t block,
                          hipStream_t stream) {
    const bool fast = fast_rgba8(a), tame = a.ahd_tame != 0;
    if (tame && fast_rgba8_lin16(a))
        launch_ahd_dual(RP, a, P, walk_blocks, grid, block, stream);
    else if (tame && fast_rgba8_rot(a) && a.orientation <= 4)
        hipLaunchKernelGGL((k_develop_ahd_walk<RP, true, true, 2>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
    else if (tame && fast_rgba8_rot(a))
        hipLaunchKernelGGL((k_develop_ahd_walk<RP, true, true, 3>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
    else if (fast && tame)
        hipLaunchKernelGGL((k_develop_ahd_walk<RP, true, true>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
    else if (fast)
        hipLaunchKernelGGL((k_develop_ahd_walk<RP, true, false>), dim3(grid), dim3(block), 0, stream, a,
                           P, walk_blocks);
    else if (tame)
        hipLaunchKernelGGL((k_develop_ahd_walk<RP, false, true>), dim3(grid), dim3(block), 0, stream, a,
                           P, walk_blocks);
    else
        hipLaunchKernelGGL((k_develop_ahd_walk<RP, false, false>), dim3(grid), dim3(block), 0, stream, a,
                           P, walk_blocks);
}

// Every float in [2^-96, 2^128), and +-0, through sqrt_nonneg against the compiler's correctly
// rounded sqrtf; the mismatches are added to *count. 16 values per thread.
constexpr uint32_t SQRT_LO = 0x0F800000u, SQRT_HI = 0x7F800000u;  // 2^-96, +inf
__global__ __launch_bounds__(256) void k_check_sqrt(unsigned long long *count) {
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t bad = 0;
    for (int j = 0; j < 16; j++) {
        const uint64_t b = SQRT_LO + t * 16 + (uint64_t)j;
        if (b < SQRT_HI) {
            const float x = __uint_as_float((uint32_t)b);
            bad += __float_as_uint(sqrt_nonneg(x)) != __float_as_uint(sqrtf(x)) ? 1u : 0u;
        }
    }
    if (t == 0) {
        bad += __float_as_uint(sqrt_nonneg(0.0f)) != __float_as_uint(sqrtf(0.0f)) ? 1u : 0u;
        bad += __float_as_uint(sqrt_nonneg(-0.0f)) != __float_as_uint(sqrtf(-0.0f)) ? 1u : 0u;
    }
    if (bad) atomicAdd(count, (unsigned long long)bad);
}

hipError_t check_sqrt(unsigned long long *count, hipStream_t stream) {
    const uint64_t threads = ((uint64_t)(SQRT_HI - SQRT_LO) + 15) / 16;
    hipLaunchKernelGGL(k_check_sqrt, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, stream, count);
    return hipGetLastError();
}

hipError_t launch_develop_ahd_walk(const DevArgs &a, hipStream_t stream, int rows_per_strip) {
    const int rp = red_position(a);
    if (rp < 0) return hipErrorInvalidValue;
    const WalkPlan P = make_walk_plan(a, rows_per_strip);
    const uint32_t wpb = walk_wpb(a.ahd_wpb, AHD_MAX_WAVES), block = wpb * 64;
    const uint32_t walk_blocks = (uint32_t)((P.strips * P.chunks + wpb - 1) / wpb);
    const uint32_t border_blocks = (P.border.count + block - 1) / block;
    const uint32_t grid = walk_blocks + border_blocks;
    if (grid == 0) return hipSuccess;
    switch (rp) {
    case 0: launch_ahd_rp<0>(a, P, walk_blocks, grid, block, stream); break;
    case 1: launch_ahd_rp<1>(a, P, walk_blocks, grid, block, stream); break;
    case 2: launch_ahd_rp<2>(a, P, walk_blocks, grid, block, stream); break;
    default: launch_ahd_rp<3>(a, P, walk_blocks, grid, block, stream); break;
    }
    return hipGetLastError();
}

}  // namespace frh
