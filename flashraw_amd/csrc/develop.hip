// Fused develop kernels for gfx950: normalise -> demosaic (bilinear or AHD) -> white balance ->
// clip -> camera->sRGB matrix -> levels -> encode -> store, one pass per frame, no full-size
// intermediate. CPU statement of the same arithmetic: crates/flashraw/src/process.rs
// (row_cols 967-1129, finish_rgb 654-707, encoders 1593-1991) and crates/flashraw/src/ahd.rs.
//
// Both kernels are persistent: a block loads the sRGB table and the calibration pattern into LDS
// once, then walks tiles `blockIdx.x, blockIdx.x + gridDim.x, ...` of the developed image.
#include "bil_walk.h"
#include "bil_walk4.h"
#include "develop_common.h"
#include "frh_launch.h"

namespace frh {

// ------------------------------------------------------------------------------------------
// Bilinear develop (the reference's `develop` kernel, develop.wgsl:185-260; CPU
// process.rs:1044-1129). A tile of BW x BH developed pixels; the normalised frame samples it
// reads (one-sample apron) are staged in LDS; each thread develops 2x2 quads.

constexpr int BW = 128, BH = 16;
constexpr int BLW = BW + 2, BLH = BH + 2;

__global__ __launch_bounds__(256) void k_develop_bilinear(DevArgs a, uint32_t tiles_x, uint32_t ntiles) {
    __shared__ Lds s;
    __shared__ float t[BLH][BLW + 2];
    load_lds(a, s);
    const int w = (int)a.width, h = (int)a.height;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
        const int dx0 = (int)(tx * BW), dy0 = (int)(a.band_y0 + ty * BH);
        const int x0 = (int)a.crop_x + dx0, y0 = (int)a.crop_y + dy0;
        __syncthreads();  // LDS tables ready / previous tile consumed
        for (int i = threadIdx.x; i < BLH * BLW; i += blockDim.x) {
            int r = i / BLW, c = i - r * BLW;
            int x = x0 - 1 + c, y = y0 - 1 + r;
            t[r][c] = (x >= 0 && x < w && y >= 0 && y < h) ? norm_at(a, s, x, y) : 0.0f;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < (BW / 2) * (BH / 2); q += blockDim.x) {
            const int qx = q % (BW / 2), qy = q / (BW / 2);
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const int ddx = dx0 + qx * 2 + (p & 1), ddy = dy0 + qy * 2 + (p >> 1);
                if (ddx >= (int)a.dev_w || ddy >= (int)a.band_y1) continue;
                const int x = (int)a.crop_x + ddx, y = (int)a.crop_y + ddy;
                float rgb[3];
                if (x > 0 && y > 0 && x + 1 < w && y + 1 < h) {
                    const int lx = x - x0 + 1, ly = y - y0 + 1;
                    const uint32_t c0 = chan_at(a, x, y);
                    float own = t[ly][lx], l = t[ly][lx - 1], r = t[ly][lx + 1], u = t[ly - 1][lx],
                          d = t[ly + 1][lx];
                    if (c0 == 1) {
                        const uint32_t ch = chan_at(a, x + 1, y);
                        float hv = (l + r) * 0.5f, vv = (u + d) * 0.5f;
                        rgb[1] = own;
                        rgb[0] = ch == 0 ? hv : vv;
                        rgb[2] = ch == 0 ? vv : hv;
                    } else {
                        float orth = (l + r + u + d) * 0.25f;
                        float diag =
                            (t[ly - 1][lx - 1] + t[ly - 1][lx + 1] + t[ly + 1][lx - 1] + t[ly + 1][lx + 1]) * 0.25f;
                        rgb[0] = c0 == 0 ? own : diag;
                        rgb[1] = orth;
                        rgb[2] = c0 == 0 ? diag : own;
                    }
                } else {
                    bilinear_global(a, s, x, y, rgb);
                }
                float o0, o1, o2;
                finish(a, rgb[0], rgb[1], rgb[2], false, o0, o1, o2);
                store_both(a, s, (uint32_t)ddx, (uint32_t)ddy, o0, o1, o2);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// AHD develop (ahd.rs:191-393; GPU statement edge.wgsl:386-548), fused into one pass: a tile of
// AW x AH developed pixels stages the balanced mosaic (apron 5), both green candidates (apron 3),
// both candidates' L, a, b (apron 2) and both homogeneity counts (apron 1) in LDS, then blends,
// finishes and encodes. Pixels nearer the frame border than `ahd_halo` are developed bilinearly
// (strip_to, process.rs:1183-1224).

constexpr int AW = 32, AH = 16;
constexpr int MW = AW + 10, MH = AH + 10;
constexpr int GW = AW + 6, GH = AH + 6;
constexpr int LW = AW + 4, LH = AH + 4;
constexpr int HW = AW + 2, HH = AH + 2;

struct AhdLds {
    float m[MH][MW];
    float g[2][GH][GW];
    float lab[6][LH][LW];
    float hom[2][HH][HW];
};

// Candidate RGB of candidate d at M index (mi, mj) / G index (mi - 2, mj - 2)
// (ahd.rs:249-291; the measured colour kept, the others from colour differences).
__device__ __forceinline__ void cand_rgb(const AhdLds &L, int d, int mi, int mj, uint32_t c, bool hr, float &red,
                                         float &gg, float &blue) {
    const int gi = mi - 2, gj = mj - 2;
#define DF(di, dj) (L.m[mi + (di)][mj + (dj)] - L.g[d][gi + (di)][gj + (dj)])
    const float g = L.g[d][gi][gj];
    const float cm = L.m[mi][mj];
    float diag = g + 0.25f * (DF(-1, -1) + DF(-1, 1) + DF(1, -1) + DF(1, 1));
    float row = g + 0.5f * (DF(0, -1) + DF(0, 1));
    float col = g + 0.5f * (DF(-1, 0) + DF(1, 0));
#undef DF
    float on_r = hr ? row : col, on_b = hr ? col : row;
    red = c == 0 ? cm : (c == 1 ? on_r : diag);
    blue = c == 2 ? cm : (c == 1 ? on_b : diag);
    gg = g;
}

__global__ __launch_bounds__(256) void k_develop_ahd(DevArgs a, uint32_t tiles_x, uint32_t ntiles) {
    __shared__ Lds s;
    __shared__ AhdLds L;
    load_lds(a, s);
    const int w = (int)a.width, h = (int)a.height;
    const int halo = (int)a.ahd_halo;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
        const int dx0 = (int)(tx * AW), dy0 = (int)(a.band_y0 + ty * AH);
        const int x0 = (int)a.crop_x + dx0, y0 = (int)a.crop_y + dy0;
        // Does the tile hold any AHD pixel (frame x in [halo, w - halo), y in [halo, h - halo))?
        const bool any = x0 + AW > halo && x0 < w - halo && y0 + AH > halo && y0 < h - halo;
        __syncthreads();
        if (any) {
            // 1. balanced mosaic, clipped at 1 (ahd.rs:143-151)
            for (int i = threadIdx.x; i < MH * MW; i += blockDim.x) {
                int r = i / MW, c = i - r * MW;
                int x = x0 - 5 + c, y = y0 - 5 + r;
                float v = 0.0f;
                if (x >= 0 && x < w && y >= 0 && y < h) v = fminf(norm_at(a, s, x, y) * wb_of(a, chan_at(a, x, y)), 1.0f);
                L.m[r][c] = v;
            }
            __syncthreads();
            // 2. green along rows and columns (ahd.rs:225-244)
            for (int i = threadIdx.x; i < GH * GW; i += blockDim.x) {
                int gi = i / GW, gj = i - gi * GW;
                int mi = gi + 2, mj = gj + 2;
                int x = x0 - 3 + gj, y = y0 - 3 + gi;
                float c = L.m[mi][mj];
                float gh = c, gv = c;
                if (chan_at(a, x, y) != 1) {
                    float l = L.m[mi][mj - 1], r = L.m[mi][mj + 1], l2 = L.m[mi][mj - 2], r2 = L.m[mi][mj + 2];
                    float u = L.m[mi - 1][mj], d = L.m[mi + 1][mj], u2 = L.m[mi - 2][mj], d2 = L.m[mi + 2][mj];
                    gh = fminf(fmaxf((l + r) * 0.5f + (2.0f * c - l2 - r2) * 0.25f, fminf(l, r)), fmaxf(l, r));
                    gv = fminf(fmaxf((u + d) * 0.5f + (2.0f * c - u2 - d2) * 0.25f, fminf(u, d)), fmaxf(u, d));
                }
                L.g[0][gi][gj] = gh;
                L.g[1][gi][gj] = gv;
            }
            __syncthreads();
            // 3. candidates' perceptual coordinates (ahd.rs:249-298)
            for (int i = threadIdx.x; i < LH * LW; i += blockDim.x) {
                int li = i / LW, lj = i - li * LW;
                int x = x0 - 2 + lj, y = y0 - 2 + li;
                uint32_t c = chan_at(a, x, y);
                bool hr = c == 1 && chan_at(a, x + 1, y) == 0;
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    float red, g, blue;
                    cand_rgb(L, d, li + 3, lj + 3, c, hr, red, g, blue);
                    float sr = sqrtf(fmaxf(red, 0.0f)), sg = sqrtf(fmaxf(g, 0.0f)), sb = sqrtf(fmaxf(blue, 0.0f));
                    L.lab[d * 3 + 0][li][lj] = (sr + 2.0f * sg + sb) * 0.25f;
                    L.lab[d * 3 + 1][li][lj] = sr - sg;
                    L.lab[d * 3 + 2][li][lj] = sb - sg;
                }
            }
            __syncthreads();
            // 4. thresholds and homogeneity counts (ahd.rs:300-351)
            for (int i = threadIdx.x; i < HH * HW; i += blockDim.x) {
                int hi = i / HW, hj = i - hi * HW;
                int li = hi + 1, lj = hj + 1;
#define STEP_L(d, di, dj) fabsf(L.lab[(d) * 3][li + (di)][lj + (dj)] - L.lab[(d) * 3][li][lj])
#define STEP_C(d, di, dj)                                                                           \
    ((L.lab[(d) * 3 + 1][li + (di)][lj + (dj)] - L.lab[(d) * 3 + 1][li][lj]) *                      \
         (L.lab[(d) * 3 + 1][li + (di)][lj + (dj)] - L.lab[(d) * 3 + 1][li][lj]) +                  \
     (L.lab[(d) * 3 + 2][li + (di)][lj + (dj)] - L.lab[(d) * 3 + 2][li][lj]) *                      \
         (L.lab[(d) * 3 + 2][li + (di)][lj + (dj)] - L.lab[(d) * 3 + 2][li][lj]))
                float el = fminf(fmaxf(STEP_L(0, 0, -1), STEP_L(0, 0, 1)), fmaxf(STEP_L(1, -1, 0), STEP_L(1, 1, 0)));
                float ec = fminf(fmaxf(STEP_C(0, 0, -1), STEP_C(0, 0, 1)), fmaxf(STEP_C(1, -1, 0), STEP_C(1, 1, 0)));
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    float cnt = 0.0f;
#pragma unroll
                    for (int di = -1; di <= 1; di++) {
#pragma unroll
                        for (int dj = -1; dj <= 1; dj++) {
                            if (di == 0 && dj == 0) continue;
                            if (STEP_L(d, di, dj) <= el && STEP_C(d, di, dj) <= ec) cnt += 1.0f;
                        }
                    }
                    L.hom[d][hi][hj] = cnt;
                }
#undef STEP_L
#undef STEP_C
            }
            __syncthreads();
        }
        // 5. blend (ahd.rs:353-392), finish and store; border pixels bilinear
        for (int i = threadIdx.x; i < AH * AW; i += blockDim.x) {
            int oi = i / AW, oj = i - oi * AW;
            int ddx = dx0 + oj, ddy = dy0 + oi;
            if (ddx >= (int)a.dev_w || ddy >= (int)a.band_y1) continue;
            int x = x0 + oj, y = y0 + oi;
            float o0, o1, o2;
            if (x >= halo && x < w - halo && y >= halo && y < h - halo) {
                float sh = 0.0f, sv = 0.0f;
#pragma unroll
                for (int di = 0; di < 3; di++)
#pragma unroll
                    for (int dj = 0; dj < 3; dj++) {
                        sh += L.hom[0][oi + di][oj + dj];
                        sv += L.hom[1][oi + di][oj + dj];
                    }
                float wgt = sh + sv > 0.0f ? sh / (sh + sv) : 0.5f;
                uint32_t c = chan_at(a, x, y);
                bool hr = c == 1 && chan_at(a, x + 1, y) == 0;
                float r0, g0, b0, r1, g1, b1;
                cand_rgb(L, 0, oi + 5, oj + 5, c, hr, r0, g0, b0);
                cand_rgb(L, 1, oi + 5, oj + 5, c, hr, r1, g1, b1);
                float iw = 1.0f - wgt;
                float r = wgt * r0 + iw * r1, g = wgt * g0 + iw * g1, b = wgt * b0 + iw * b1;
                finish(a, r, g, b, true, o0, o1, o2);
            } else {
                float rgb[3];
                bilinear_global(a, s, x, y, rgb);
                finish(a, rgb[0], rgb[1], rgb[2], false, o0, o1, o2);
            }
            store_both(a, s, (uint32_t)ddx, (uint32_t)ddy, o0, o1, o2);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Bilinear row walker launcher (the walker is in bil_walk.h).

template <int RP, bool FAST, bool TAME, int ROT = 0>
__global__ __launch_bounds__(BIL_MAX_WAVES * 64) void k_develop_bilinear_walk(DevArgs args, BWalkPlan P,
                                                                            uint32_t walk_blocks) {
    const DevArgs &a = kernel_args(args);
    __shared__ Lds s;
    load_lds(a, s);
    __syncthreads();
    // the walker blocks come first and are renumbered among themselves; border blocks follow
    const uint32_t blk = blockIdx.x < walk_blocks ? xcd_block(blockIdx.x, walk_blocks) : blockIdx.x;
    if (blk >= walk_blocks) {
        const uint32_t i = (blk - walk_blocks) * blockDim.x + threadIdx.x;
        if (i < P.border.count) border_pixel<false>(a, s, P.border, i);
        return;
    }
    // the wave index is the same for every lane; readfirstlane lets the compiler keep all the strip
    // bookkeeping in scalar registers
    const int item = (int)uniform(blk * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (item >= P.strips * P.chunks) return;
    bwalk_item<RP, FAST, TAME, ROT>(a, P, s, item);
}

template <int RP>
static void launch_bil_rp(const DevArgs &a, const BWalkPlan &P, uint32_t walk_blocks, uint32_t grid, uint32_t block,
                          hipStream_t stream) {
    // the tame frame's finish takes the matrix pre-scaled by 4096 (every product is 0 or normal:
    // the bound of ahd_tame covers the bilinear means too)
    const bool fast = fast_rgba8(a), tame = a.ahd_tame != 0;
    if (fast && tame)
        hipLaunchKernelGGL((k_develop_bilinear_walk<RP, true, true>), dim3(grid), dim3(block), 0, stream,
                           a, P, walk_blocks);
    else if (tame && fast_rgba8_rot(a) && a.orientation <= 4)
        hipLaunchKernelGGL((k_develop_bilinear_walk<RP, true, true, 1>), dim3(grid), dim3(block), 0, stream,
                           a, P, walk_blocks);
    else if (tame && fast_rgba8_rot(a))
        hipLaunchKernelGGL((k_develop_bilinear_walk<RP, true, true, 2>), dim3(grid), dim3(block), 0, stream,
                           a, P, walk_blocks);
    else if (fast)
        hipLaunchKernelGGL((k_develop_bilinear_walk<RP, true, false>), dim3(grid), dim3(block), 0,
                           stream, a, P, walk_blocks);
    else
        hipLaunchKernelGGL((k_develop_bilinear_walk<RP, false, false>), dim3(grid), dim3(block), 0,
                           stream, a, P, walk_blocks);
}

// The four-column bilinear walker (RGBA8 fast path of aligned frames), blocks renumbered per XCD
// like the two-column one.
template <int RP, bool TAME>
__global__ __launch_bounds__(BIL_MAX_WAVES * 64) void k_develop_bilinear_walk4(DevArgs args, BWalkPlan P,
                                                                             uint32_t walk_blocks) {
    const DevArgs &a = kernel_args(args);
    __shared__ Lds s;
    load_lds(a, s);
    __syncthreads();
    const uint32_t blk = blockIdx.x < walk_blocks ? xcd_block(blockIdx.x, walk_blocks) : blockIdx.x;
    if (blk >= walk_blocks) {
        const uint32_t i = (blk - walk_blocks) * blockDim.x + threadIdx.x;
        if (i < P.border.count) border_pixel<false>(a, s, P.border, i);
        return;
    }
    const int item = (int)uniform(blk * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (item >= P.strips * P.chunks) return;
    bwalk4_item<RP, TAME>(a, P, s, item);
}

template <int RP>
static void launch_bil4_rp(const DevArgs &a, const BWalkPlan &P, uint32_t walk_blocks, uint32_t grid,
                           uint32_t block, hipStream_t stream) {
    if (a.ahd_tame)
        hipLaunchKernelGGL((k_develop_bilinear_walk4<RP, true>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
    else
        hipLaunchKernelGGL((k_develop_bilinear_walk4<RP, false>), dim3(grid), dim3(block), 0, stream, a, P,
                           walk_blocks);
}

static hipError_t launch_develop_bilinear_walk(const DevArgs &a, hipStream_t stream, int rows_per_strip) {
    const int rp = red_position(a);
    if (rp < 0) return hipErrorInvalidValue;
    if (a.bil_cpl == 4 && walk4_ok(a)) {
        const BWalkPlan P = make_bwalk4_plan(a, rows_per_strip);
        const uint32_t wpb = walk_wpb(a.bil_wpb, BIL_MAX_WAVES), block = wpb * 64;
        const uint32_t walk_blocks = (uint32_t)((P.strips * P.chunks + wpb - 1) / wpb);
        const uint32_t border_blocks = (P.border.count + block - 1) / block;
        const uint32_t grid = walk_blocks + border_blocks;
        if (grid == 0) return hipSuccess;
        switch (rp) {
        case 0: launch_bil4_rp<0>(a, P, walk_blocks, grid, block, stream); break;
        case 1: launch_bil4_rp<1>(a, P, walk_blocks, grid, block, stream); break;
        case 2: launch_bil4_rp<2>(a, P, walk_blocks, grid, block, stream); break;
        default: launch_bil4_rp<3>(a, P, walk_blocks, grid, block, stream); break;
        }
        return hipGetLastError();
    }
    const BWalkPlan P = make_bwalk_plan(a, rows_per_strip);
    const uint32_t wpb = walk_wpb(a.bil_wpb, BIL_MAX_WAVES), block = wpb * 64;
    const uint32_t walk_blocks = (uint32_t)((P.strips * P.chunks + wpb - 1) / wpb);
    const uint32_t border_blocks = (P.border.count + block - 1) / block;
    const uint32_t grid = walk_blocks + border_blocks;
    if (grid == 0) return hipSuccess;
    switch (rp) {
    case 0: launch_bil_rp<0>(a, P, walk_blocks, grid, block, stream); break;
    case 1: launch_bil_rp<1>(a, P, walk_blocks, grid, block, stream); break;
    case 2: launch_bil_rp<2>(a, P, walk_blocks, grid, block, stream); break;
    default: launch_bil_rp<3>(a, P, walk_blocks, grid, block, stream); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Launchers

static uint32_t persistent_grid(uint32_t ntiles, int num_cus, int blocks_per_cu) {
    uint32_t g = (uint32_t)num_cus * (uint32_t)blocks_per_cu;
    return ntiles < g ? ntiles : g;
}

hipError_t launch_develop(const DevArgs &a, bool ahd, hipStream_t stream, int num_cus) {
    if (a.generic) return launch_develop_generic(a, stream);
    if (a.edge_aware) return launch_develop_edge(a, stream);
    const uint32_t rows = a.band_y1 - a.band_y0;
    if (rows == 0 || a.dev_w == 0) return hipSuccess;
    if (ahd && a.ahd_impl == 0) return launch_develop_ahd_walk(a, stream, a.ahd_rows ? (int)a.ahd_rows : 64);
    if (!ahd && a.bil_impl == 0) return launch_develop_bilinear_walk(a, stream, a.bil_rows ? (int)a.bil_rows : 64);
    if (ahd) {
        uint32_t tx = (a.dev_w + AW - 1) / AW, ty = (rows + AH - 1) / AH;
        uint32_t n = tx * ty;
        hipLaunchKernelGGL(k_develop_ahd, dim3(persistent_grid(n, num_cus, 4)), dim3(256), 0, stream, a, tx, n);
    } else {
        uint32_t tx = (a.dev_w + BW - 1) / BW, ty = (rows + BH - 1) / BH;
        uint32_t n = tx * ty;
        hipLaunchKernelGGL(k_develop_bilinear, dim3(persistent_grid(n, num_cus, 8)), dim3(256), 0, stream, a, tx,
                           n);
    }
    return hipGetLastError();
}

}  // namespace frh
