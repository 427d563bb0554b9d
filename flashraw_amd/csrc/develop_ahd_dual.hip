// The AHD walker's two-output form (RGBA8 and linear RGB16 in one pass, tame frames). A
// translation unit of its own: it is built with the machine scheduler's max-ILP strategy, as this
// compiler fails on its instantiations with the iterative ILP strategy of develop_ahd.hip.
#include "ahd_kernel.h"
#include "frh_launch.h"

namespace frh {

void launch_ahd_dual(int rp, const DevArgs &a, const WalkPlan &P, uint32_t walk_blocks, uint32_t grid, uint32_t block,
                     hipStream_t stream) {
    switch (rp) {
    case 0:
        hipLaunchKernelGGL((k_develop_ahd_walk<0, true, true, 1>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
        break;
    case 1:
        hipLaunchKernelGGL((k_develop_ahd_walk<1, true, true, 1>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
        break;
    case 2:
        hipLaunchKernelGGL((k_develop_ahd_walk<2, true, true, 1>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
        break;
    default:
        hipLaunchKernelGGL((k_develop_ahd_walk<3, true, true, 1>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
        break;
    }
}

}  // namespace frh
