// Fused develop kernels for gfx950: normalise -> demosaic (bilinear or AHD) -> white balance ->
// clip -> camera->sRGB matrix -> levels -> encode -> store, one pass per frame, no full-size
// intermediate. CPU statement of the same arithmetic: crates/flashraw/src/process.rs
// (row_cols 967-1129, finish_rgb 654-707, encoders 1593-1991) and crates/flashraw/src/ahd.rs.
//
// Both kernels are persistent: a block loads the sRGB table and the calibration pattern into LDS
// once, then walks tiles `blockIdx.x, blockIdx.x + gridDim.x, ...` of the developed image.
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
// Bilinear develop, register row walker: lane j owns frame columns xs + 2j and xs + 2j + 1, the
// wave walks down a strip keeping the three normalised rows the 3x3 neighbourhood needs in
// registers, horizontal neighbours come from the adjacent lanes (DPP). With the Bayer phase a
// template parameter and rows unrolled by two, every pixel evaluates only its own formulas
// (process.rs:1076-1096). Lanes 1..62 hold results (124 columns per strip); pixels on the frame
// border are developed by extra blocks with the in-frame neighbour means (process.rs:1109-1126).

constexpr int BWALK_COLS = 124;
constexpr int BWALK_WAVES = 4;
constexpr int BPF = 4;  // rows fetched ahead (each slot is loaded and consumed in place, never moved)

struct BWalk {
    float n[3][2];           // normalised rows R-2, R-1, R; [k] even / odd column
    float nl[3];             // odd column of lane - 1
    float nr[3];             // even column of lane + 1
    uint32_t pf[BPF];        // raw pairs fetched ahead, slot (row - start) % BPF
};

template <int RP, int PY, int SLOT, bool FAST, bool HOT>
__device__ __forceinline__ void bwalk_row(const DevArgs &a, const Lds &s, BWalk &W, int R, int xs, int lane,
                                          int out_row0, int out_row1, int out_x0, int out_x1, bool in0, bool in1,
                                          __amdgpu_buffer_rsrc_t rs, __amdgpu_buffer_rsrc_t rd) {
    const int x0 = xs + 2 * lane;
#pragma unroll
    for (int r = 0; r < 2; r++) {
        W.n[r][0] = W.n[r + 1][0];
        W.n[r][1] = W.n[r + 1][1];
        W.nl[r] = W.nl[r + 1];
        W.nr[r] = W.nr[r + 1];
    }
    const uint32_t raw = W.pf[SLOT];
    W.pf[SLOT] = fetch_pair<HOT>(a, rs, x0, R + BPF);
    norm_pair<PY, HOT>(a, s, raw, x0, R, R >= 0 && R < (int)a.height, in0, in1, W.n[2][0], W.n[2][1]);
    W.nl[2] = from_left(W.n[2][1]);
    W.nr[2] = from_right(W.n[2][0]);
    const int ro = R - 1;
    if (ro < out_row0 || ro >= out_row1 || lane < 1 || lane > 62) return;
    constexpr int PO = PY ^ 1;
    float o[2][3];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float own = W.n[1][k];
        const float l = k == 0 ? W.nl[1] : W.n[1][0], r = k == 0 ? W.n[1][1] : W.nr[1];
        const float u = W.n[0][k], d = W.n[2][k];
        float rgb[3];
        const int col = site<RP>(PO, k);
        if (col == 1) {
            const int ch = site<RP>(PO, k ^ 1);  // the colour of this row's other sites
            const float hv = (l + r) * 0.5f, vv = (u + d) * 0.5f;
            rgb[0] = ch == 0 ? hv : vv;
            rgb[1] = own;
            rgb[2] = ch == 0 ? vv : hv;
        } else {
            const float ul = k == 0 ? W.nl[0] : W.n[0][0], ur = k == 0 ? W.n[0][1] : W.nr[0];
            const float dl = k == 0 ? W.nl[2] : W.n[2][0], dr = k == 0 ? W.n[2][1] : W.nr[2];
            const float orth = (l + r + u + d) * 0.25f;
            const float diag = (ul + ur + dl + dr) * 0.25f;
            rgb[0] = col == 0 ? own : diag;
            rgb[1] = orth;
            rgb[2] = col == 0 ? diag : own;
        }
        if (FAST) {
            o[k][0] = __uint_as_float(finish_rgba8_fast(a, s, rgb[0], rgb[1], rgb[2], false));
        } else {
            finish(a, rgb[0], rgb[1], rgb[2], false, o[k][0], o[k][1], o[k][2]);
        }
    }
    const int dx = x0 - (int)a.crop_x, dy = ro - (int)a.crop_y;
    const bool st0 = x0 >= out_x0 && x0 < out_x1, st1 = x0 + 1 >= out_x0 && x0 + 1 < out_x1;
    if (FAST) {
        const uint32_t w0 = __float_as_uint(o[0][0]), w1 = __float_as_uint(o[1][0]);
        if (st0 && st1) {
            store_rgba8_pair(a, rd, (uint32_t)dx, (uint32_t)dy, w0, w1, 2);
        } else if (st0) {
            store_rgba8_pair(a, rd, (uint32_t)dx, (uint32_t)dy, w0, 0, 1);
        } else if (st1) {
            store_rgba8_pair(a, rd, (uint32_t)dx + 1, (uint32_t)dy, w1, 0, 1);
        }
    } else {
        if (st0) store_both(a, s, (uint32_t)dx, (uint32_t)dy, o[0][0], o[0][1], o[0][2]);
        if (st1) store_both(a, s, (uint32_t)dx + 1, (uint32_t)dy, o[1][0], o[1][1], o[1][2]);
    }
}

// One wave's strip: rows from row0 - 2 (even) until the last output row out_row1 - 1 = R - 1,
// four rows per trip so every prefetch slot is static.
template <int RP, bool FAST, bool HOT>
__device__ __forceinline__ void bwalk_strip(const DevArgs &a, const Lds &s, int xs, int lane, int row0, int out_row0,
                                            int out_row1, int out_x0, int out_x1) {
    BWalk W;
    {
        float *f = reinterpret_cast<float *>(&W);
        for (int i = 0; i < (int)(sizeof(BWalk) / 4) - BPF; i++) f[i] = 0.0f;
    }
    const int x0 = xs + 2 * lane;
    const bool in0 = x0 >= 0 && x0 < (int)a.width, in1 = x0 + 1 >= 0 && x0 + 1 < (int)a.width;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(a.src, src_bytes(a));
    const __amdgpu_buffer_rsrc_t rd = make_rsrc(a.dst, a.dst_w * a.dev_h * 4u);
#pragma unroll
    for (int i = 0; i < BPF; i++) W.pf[i] = fetch_pair<HOT>(a, rs, x0, row0 - 2 + i);
    for (int R = row0 - 2; R < out_row1 + 1; R += BPF) {
        bwalk_row<RP, 0, 0, FAST, HOT>(a, s, W, R, xs, lane, out_row0, out_row1, out_x0, out_x1, in0, in1, rs, rd);
        bwalk_row<RP, 1, 1, FAST, HOT>(a, s, W, R + 1, xs, lane, out_row0, out_row1, out_x0, out_x1, in0, in1, rs,
                                       rd);
        bwalk_row<RP, 0, 2, FAST, HOT>(a, s, W, R + 2, xs, lane, out_row0, out_row1, out_x0, out_x1, in0, in1, rs,
                                       rd);
        bwalk_row<RP, 1, 3, FAST, HOT>(a, s, W, R + 3, xs, lane, out_row0, out_row1, out_x0, out_x1, in0, in1, rs,
                                       rd);
    }
}

struct BWalkPlan {
    int x_first, x_end, y_first, y_end;
    int xs0, ys0;
    int strips, chunks, rows;
    Border border;
};

template <int RP, bool FAST>
__global__ __launch_bounds__(BWALK_WAVES * 64) void k_develop_bilinear_walk(DevArgs a, BWalkPlan P,
                                                                            uint32_t walk_blocks) {
    __shared__ Lds s;
    load_lds(a, s);
    __syncthreads();
    if (blockIdx.x >= walk_blocks) {
        const uint32_t i = (blockIdx.x - walk_blocks) * blockDim.x + threadIdx.x;
        if (i < P.border.count) border_pixel(a, s, P.border, i);
        return;
    }
    const int lane = threadIdx.x & 63;
    // the wave index is the same for every lane; readfirstlane lets the compiler keep all the strip
    // bookkeeping in scalar registers
    const int item = (int)uniform(blockIdx.x * BWALK_WAVES + (threadIdx.x >> 6));
    if (item >= P.strips * P.chunks) return;
    const int sx = item % P.strips, sy = item / P.strips;
    const int xs = P.xs0 + BWALK_COLS * sx - 2;  // lane 1 holds output column xs + 2
    const int row0 = P.ys0 + P.rows * sy;
    const int out_row0 = row0 > P.y_first ? row0 : P.y_first;
    const int out_row1 = row0 + P.rows < P.y_end ? row0 + P.rows : P.y_end;
    const int out_x0 = max(xs + 2, P.x_first), out_x1 = min(xs + 2 + BWALK_COLS, P.x_end);
    const bool hot = a.simple_cal && xs >= 0 && xs + 128 <= (int)a.width && ((uintptr_t)a.src & 3u) == 0 &&
                     (a.stride & 1u) == 0;
    if (hot)
        bwalk_strip<RP, FAST, true>(a, s, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
    else
        bwalk_strip<RP, FAST, false>(a, s, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
}

template <int RP>
static void launch_bil_rp(const DevArgs &a, const BWalkPlan &P, uint32_t walk_blocks, uint32_t grid,
                          hipStream_t stream) {
    if (fast_rgba8(a))
        hipLaunchKernelGGL((k_develop_bilinear_walk<RP, true>), dim3(grid), dim3(BWALK_WAVES * 64), 0, stream, a,
                           P, walk_blocks);
    else
        hipLaunchKernelGGL((k_develop_bilinear_walk<RP, false>), dim3(grid), dim3(BWALK_WAVES * 64), 0, stream, a,
                           P, walk_blocks);
}

static hipError_t launch_develop_bilinear_walk(const DevArgs &a, hipStream_t stream, int rows_per_strip) {
    const int rp = red_position(a);
    if (rp < 0) return hipErrorInvalidValue;
    BWalkPlan P;
    P.border = make_border(a, 1);
    P.rows = rows_per_strip < 2 ? 2 : (rows_per_strip & ~1);
    if (P.border.ix1 > P.border.ix0) {
        P.x_first = P.border.ix0 + (int)a.crop_x;
        P.x_end = P.border.ix1 + (int)a.crop_x;
        P.y_first = P.border.iy0 + (int)a.crop_y;
        P.y_end = P.border.iy1 + (int)a.crop_y;
        P.xs0 = P.x_first & ~1;
        P.ys0 = P.y_first & ~1;
        P.strips = (P.x_end - P.xs0 + BWALK_COLS - 1) / BWALK_COLS;
        P.chunks = (P.y_end - P.ys0 + P.rows - 1) / P.rows;
    } else {
        P.x_first = P.x_end = P.y_first = P.y_end = P.xs0 = P.ys0 = 0;
        P.strips = P.chunks = 0;
    }
    const uint32_t walk_blocks = (uint32_t)((P.strips * P.chunks + BWALK_WAVES - 1) / BWALK_WAVES);
    const uint32_t border_blocks = (P.border.count + BWALK_WAVES * 64 - 1) / (BWALK_WAVES * 64);
    const uint32_t grid = walk_blocks + border_blocks;
    if (grid == 0) return hipSuccess;
    switch (rp) {
    case 0: launch_bil_rp<0>(a, P, walk_blocks, grid, stream); break;
    case 1: launch_bil_rp<1>(a, P, walk_blocks, grid, stream); break;
    case 2: launch_bil_rp<2>(a, P, walk_blocks, grid, stream); break;
    default: launch_bil_rp<3>(a, P, walk_blocks, grid, stream); break;
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
