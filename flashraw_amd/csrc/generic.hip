// The generic develop (gfx950): every frame that is not a single-component Bayer mosaic developed
// at full size. CPU statement: Kernel::row_cols (crates/flashraw/src/process.rs:967-1129) -
// half size averages the samples of each 2x2 block per colour (975-1036); frames without a CFA
// take RGB from their first three components or grey from the first (1038-1046); other
// single-component CFAs (X-Trans, ...) average the same-colour samples of the 3x3 neighbourhood,
// sum * (1 / n) inside the frame and sum / n over the neighbours inside at the border
// (1098-1126). One thread develops one output pixel and stores it in every encoding; the sums
// run in the reference's order (neighbours row-major, blocks (0,0), (1,0), (0,1), (1,1)).
#include <hip/hip_runtime.h>

#include "develop_common.h"
#include "frh_launch.h"

namespace frh {
namespace {

__device__ __forceinline__ void add_to(uint32_t c, float v, float &a0, float &a1, float &a2, float &n0, float &n1,
                                       float &n2) {
    if (c == 0) {
        a0 += v;
        n0 += 1.0f;
    } else if (c == 1) {
        a1 += v;
        n1 += 1.0f;
    } else if (c == 2) {
        a2 += v;
        n2 += 1.0f;
    }
}

__global__ __launch_bounds__(256) void k_develop_generic(DevArgs a) {
    __shared__ Lds s;
    load_lds(a, s);
    __syncthreads();
    const uint32_t dw = a.dev_w, rows = a.band_y1 - a.band_y0;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= dw * rows) return;
    const uint32_t ox = i % dw, oy = a.band_y0 + i / dw;  // developed coordinates
    float r, g, b;
    if (a.half) {
        const uint32_t x0 = a.crop_x + 2u * ox, y0 = a.crop_y + 2u * oy;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
        for (uint32_t dy = 0; dy < 2; dy++) {
            const uint32_t y = y0 + dy;
            if (y >= a.height) continue;
            for (uint32_t dx = 0; dx < 2; dx++) {
                const uint32_t x = x0 + dx;
                if (x >= a.width) continue;
                if (a.gmode == 1) {
                    add_to(site_at(a, x, y) & 3u, gen_norm(a, x, y, 0), a0, a1, a2, n0, n1, n2);
                } else if (a.cpp == 1) {
                    a1 += gen_norm(a, x, y, 0);
                    n1 += 1.0f;
                } else {
                    a0 += gen_norm(a, x, y, 0);
                    n0 += 1.0f;
                    a1 += gen_norm(a, x, y, 1);
                    n1 += 1.0f;
                    if (a.cpp >= 3) {
                        a2 += gen_norm(a, x, y, 2);
                        n2 += 1.0f;
                    }
                }
            }
        }
        r = n0 > 0.0f ? a0 / n0 : 0.0f;
        g = n1 > 0.0f ? a1 / n1 : 0.0f;
        b = n2 > 0.0f ? a2 / n2 : 0.0f;
        if (a.cpp == 1 && a.gmode != 1) r = g = b = a1 / fmaxf(n1, 1.0f);
    } else {
        const uint32_t x = a.crop_x + ox, y = a.crop_y + oy;
        if (a.gmode == 1) {
            cfa_mean_rgb(a, x, y, r, g, b);
        } else if (a.gmode == 2) {
            r = gen_norm(a, x, y, 0);
            g = gen_norm(a, x, y, 1);
            b = gen_norm(a, x, y, 2);
        } else {
            r = g = b = gen_norm(a, x, y, 0);
        }
    }
    float o0, o1, o2;
    finish(a, r, g, b, false, o0, o1, o2);
    store_both(a, s, ox, oy, o0, o1, o2);
}

}  // namespace

hipError_t launch_develop_generic(const DevArgs &a, hipStream_t stream) {
    const uint64_t n = (uint64_t)a.dev_w * (a.band_y1 - a.band_y0);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_develop_generic, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace frh
