// Bilinear develop, register row walker (gfx950).
#pragma once

#include "develop_common.h"

namespace frh {

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
    uint32_t hold[2];        // ROT 2: the previous output row's two pixels
    uint32_t pf[BPF];        // raw pairs fetched ahead, slot (row - start) % BPF
};

// Per-lane constants of a strip: columns x0, x0 + 1 inside the frame (in0, in1) and output
// columns of the strip (st0, st1); pair_even[p]: at row parity p the lane's two output pixels start
// at an even destination pixel (one 8-byte store).
struct BStripCols {
    bool in0, in1, st0, st1;
    bool pair_even[2];
    // ROT: destination pixel index of the lane's first pixel on the crop's first row, and its steps
    // per developed column and row (modulo 2^32); ROT 2: rot_lo is 1 when the previous row's pixel
    // of a column has the lower destination index (cy = +1)
    uint32_t rot_lane, rot_cx, rot_cy, rot_lo;
};

// ROT (with FAST and TAME): RGBA8 under EXIF orientation 2..4 (ROT 1: rows stay rows, one 4-byte
// store per pixel) or 5..8 (ROT 2: a column's pixels on consecutive rows are adjacent, and go out
// in pairs).
template <int RP, int PY, int SLOT, bool FAST, bool HOT, bool TAME, bool OUT = true, int ROT = 0>
__device__ __forceinline__ void bwalk_row(const DevArgs &a, const Lds &s, BWalk &W, int R, int xs, int lane,
                                          int out_row0, int out_row1, const BStripCols &C,
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
    // rows past out_row1 feed no output: their prefetch re-reads row out_row1 (a cache hit)
    W.pf[SLOT] = fetch_pair<HOT>(a, rs, x0, min(R + BPF, out_row1));
    norm_pair<PY, HOT>(a, s, raw, x0, R, R >= 0 && R < (int)a.height, C.in0, C.in1, W.n[2][0], W.n[2][1]);
    W.nl[2] = from_left(W.n[2][1]);
    W.nr[2] = from_right(W.n[2][0]);
    if (!OUT) return;
    const int ro = R - 1;
    // FAST: every row of the main loop has its output row inside the strip but the first row of a
    // frame's first strip (the frame border), whose stores are dropped like those of the halo lanes
    if (!FAST && (ro < out_row0 || ro >= out_row1 || lane < 1 || lane > 62)) return;
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
            o[k][0] = __uint_as_float(TAME ? finish_rgba8_scaled(a, s, rgb[0] * a.wb[0], rgb[1] * a.wb[1],
                                                                 rgb[2] * a.wb[2])
                                           : finish_rgba8_fast(a, s, rgb[0], rgb[1], rgb[2], false));
        } else {
            finish(a, rgb[0], rgb[1], rgb[2], false, o[k][0], o[k][1], o[k][2]);
        }
    }
    if (FAST) {
        // every lane finishes its pixels; stores of columns outside the strip's output address past
        // the destination resource, where the hardware drops them (no divergent branch)
        const bool rv = ro >= out_row0 && ro < out_row1;
        const uint32_t off = ((uint32_t)(ro - (int)a.crop_y) * a.dst_w + (uint32_t)(x0 - (int)a.crop_x)) * 4u;
        const bool s0 = rv && C.st0, s1 = rv && C.st1, pe = C.pair_even[PY];
        constexpr uint32_t PAST = 0xfffffff0u;
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        u32x2 v;
        v.x = __float_as_uint(o[0][0]);
        v.y = __float_as_uint(o[1][0]);
        if (ROT == 1) {
            const uint32_t i0 = C.rot_lane + C.rot_cy * (uint32_t)(ro - (int)a.crop_y), i1 = i0 + C.rot_cx;
            __builtin_amdgcn_raw_buffer_store_b32(v.x, rd, s0 ? (int)(i0 * 4u) : (int)PAST, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(v.y, rd, s1 ? (int)(i1 * 4u) : (int)PAST, 0, 0);
        } else if (ROT == 2) {
            // A row whose pixel closes an aligned pair with the previous row's (lower index even)
            // stores both as 8 bytes when the previous row is the strip's; a row that opens one leaves
            // its pixel to the next row unless that row is past the strip's output; any other owned
            // pixel goes out alone. The rows before the strip's first output row store nothing.
            const uint32_t i0 = C.rot_lane + C.rot_cy * (uint32_t)(ro - (int)a.crop_y), i1 = i0 + C.rot_cx;
            const bool rp = ro - 1 >= out_row0, rn = ro + 1 < out_row1, sw = C.rot_lo != 0u;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const uint32_t ic = k ? i1 : i0, lo = ic - C.rot_lo, ok = k ? v.y : v.x;
                const bool own = k ? s1 : s0, closer = !(lo & 1u);
                u32x2 q;
                q.x = sw ? W.hold[k] : ok;
                q.y = sw ? ok : W.hold[k];
                __builtin_amdgcn_raw_buffer_store_b64(q, rd, (own && closer && rp) ? (int)(lo * 4u) : (int)PAST, 0,
                                                      0);
                __builtin_amdgcn_raw_buffer_store_b32(ok, rd, (own && (closer ? !rp : !rn)) ? (int)(ic * 4u) : (int)PAST,
                                                      0, 0);
                W.hold[k] = ok;
            }
        } else {
            __builtin_amdgcn_raw_buffer_store_b64(v, rd, (pe && s0 && s1) ? (int)off : (int)PAST, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(v.x, rd, (s0 && !(pe && s1)) ? (int)off : (int)PAST, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(v.y, rd, (s1 && !(pe && s0)) ? (int)(off + 4u) : (int)PAST, 0,
                                                  0);
        }
    } else {
        const int dx = x0 - (int)a.crop_x, dy = ro - (int)a.crop_y;
        if (C.st0) store_both<false>(a, s, (uint32_t)dx, (uint32_t)dy, o[0][0], o[0][1], o[0][2]);
        if (C.st1) store_both<false>(a, s, (uint32_t)dx + 1, (uint32_t)dy, o[1][0], o[1][1], o[1][2]);
    }
}

// One wave's strip: mosaic rows row0 - 2 (even) .. out_row1; output row R - 1 follows row R.
// FAST (no divergent branch in a row): three prologue rows fill the window (row0 gives output row
// row0 - 1, outside the strip), then whole trips of four rows, each straight-line code the compiler
// schedules as one block (the LDS table reads of one row overlap the next row's arithmetic), then
// at most three rows more. Every row has a static prefetch slot (R - row0 + 2) mod 4 and parity.
template <int RP, bool FAST, bool HOT, bool TAME, int ROT = 0>
__device__ __forceinline__ void bwalk_strip(const DevArgs &a, const Lds &s, int xs, int lane, int row0, int out_row0,
                                            int out_row1, int out_x0, int out_x1) {
    BWalk W;
    {
        float *f = reinterpret_cast<float *>(&W);
        for (int i = 0; i < (int)(sizeof(BWalk) / 4) - BPF; i++) f[i] = 0.0f;
    }
    const int x0 = xs + 2 * lane;
    BStripCols C;
    C.in0 = x0 >= 0 && x0 < (int)a.width;
    C.in1 = x0 + 1 >= 0 && x0 + 1 < (int)a.width;
    C.st0 = x0 >= out_x0 && x0 < out_x1;
    C.st1 = x0 + 1 >= out_x0 && x0 + 1 < out_x1;
    // the output row R - 1 of row parity PY has parity PY ^ 1
    const uint32_t cpar = (uint32_t)(xs - (int)a.crop_x) & 1u;
#pragma unroll
    for (int py = 0; py < 2; py++)
        C.pair_even[py] = ((cpar + (uint32_t)(py ^ 1 ^ (int)(a.crop_y & 1u)) * a.dst_w) & 1u) == 0;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(a.src, src_bytes(a));
    const __amdgpu_buffer_rsrc_t rd = make_rsrc(a.dst, a.dev_w * a.dev_h * 4u);
    C.rot_lane = C.rot_cx = C.rot_cy = C.rot_lo = 0u;
    if (ROT) {
        // destination index c0 + cx dx + cy dy (orient, process.rs:2035-2044)
        uint32_t c0;
        orient_linear(a, c0, C.rot_cx, C.rot_cy);
        C.rot_lane = c0 + C.rot_cx * (uint32_t)(x0 - (int)a.crop_x);
        C.rot_lo = C.rot_cy == 1u ? 1u : 0u;
    }
#pragma unroll
    for (int i = 0; i < BPF; i++) W.pf[i] = fetch_pair<HOT>(a, rs, x0, min(row0 - 2 + i, out_row1));
    if (!FAST) {
        // the general path (lanes and rows leave each row early, a divergent branch): whole trips
        // of four rows from row0 - 2, the rows past out_row1 only shifting the window
        for (int R = row0 - 2; R < out_row1 + 1; R += BPF) {
            bwalk_row<RP, 0, 0, FAST, HOT, TAME, true, ROT>(a, s, W, R, xs, lane, out_row0, out_row1, C, rs, rd);
            bwalk_row<RP, 1, 1, FAST, HOT, TAME, true, ROT>(a, s, W, R + 1, xs, lane, out_row0, out_row1, C, rs, rd);
            bwalk_row<RP, 0, 2, FAST, HOT, TAME, true, ROT>(a, s, W, R + 2, xs, lane, out_row0, out_row1, C, rs, rd);
            bwalk_row<RP, 1, 3, FAST, HOT, TAME, true, ROT>(a, s, W, R + 3, xs, lane, out_row0, out_row1, C, rs, rd);
        }
        return;
    }
    bwalk_row<RP, 0, 0, FAST, HOT, TAME, false, ROT>(a, s, W, row0 - 2, xs, lane, out_row0, out_row1, C, rs, rd);
    bwalk_row<RP, 1, 1, FAST, HOT, TAME, false, ROT>(a, s, W, row0 - 1, xs, lane, out_row0, out_row1, C, rs, rd);
    bwalk_row<RP, 0, 2, FAST, HOT, TAME, false, ROT>(a, s, W, row0, xs, lane, out_row0, out_row1, C, rs, rd);
    int R = row0 + 1;
    for (; R + 3 <= out_row1; R += 4) {
        bwalk_row<RP, 1, 3, FAST, HOT, TAME, true, ROT>(a, s, W, R, xs, lane, out_row0, out_row1, C, rs, rd);
        bwalk_row<RP, 0, 0, FAST, HOT, TAME, true, ROT>(a, s, W, R + 1, xs, lane, out_row0, out_row1, C, rs, rd);
        bwalk_row<RP, 1, 1, FAST, HOT, TAME, true, ROT>(a, s, W, R + 2, xs, lane, out_row0, out_row1, C, rs, rd);
        bwalk_row<RP, 0, 2, FAST, HOT, TAME, true, ROT>(a, s, W, R + 3, xs, lane, out_row0, out_row1, C, rs, rd);
    }
    if (R <= out_row1) bwalk_row<RP, 1, 3, FAST, HOT, TAME, true, ROT>(a, s, W, R, xs, lane, out_row0, out_row1, C, rs, rd);
    if (R + 1 <= out_row1)
        bwalk_row<RP, 0, 0, FAST, HOT, TAME, true, ROT>(a, s, W, R + 1, xs, lane, out_row0, out_row1, C, rs, rd);
    if (R + 2 <= out_row1)
        bwalk_row<RP, 1, 1, FAST, HOT, TAME, true, ROT>(a, s, W, R + 2, xs, lane, out_row0, out_row1, C, rs, rd);
}

struct BWalkPlan {
    int x_first, x_end, y_first, y_end;
    int xs0, ys0;
    int strips, chunks, rows;
    Border border;
};

// The bilinear walker's work split of the developed band of `a`.
__host__ inline BWalkPlan make_bwalk_plan(const DevArgs &a, int rows_per_strip) {
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
    return P;
}

// Strip `item` of plan P, developed by the calling wave.
template <int RP, bool FAST, bool TAME, int ROT = 0>
__device__ __forceinline__ void bwalk_item(const DevArgs &a, const BWalkPlan &P, const Lds &s, int item) {
    const int lane = threadIdx.x & 63;
    const int sx = item % P.strips, sy = item / P.strips;
    const int xs = P.xs0 + BWALK_COLS * sx - 2;  // lane 1 holds output column xs + 2
    const int row0 = P.ys0 + P.rows * sy;
    const int out_row0 = row0 > P.y_first ? row0 : P.y_first;
    const int out_row1 = row0 + P.rows < P.y_end ? row0 + P.rows : P.y_end;
    const int out_x0 = max(xs + 2, P.x_first), out_x1 = min(xs + 2 + BWALK_COLS, P.x_end);
    const bool hot = a.simple_cal && xs >= 0 && xs + 128 <= (int)a.width && ((uintptr_t)a.src & 3u) == 0 &&
                     (a.stride & 1u) == 0;
    if (hot)
        bwalk_strip<RP, FAST, true, TAME, ROT>(a, s, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
    else
        bwalk_strip<RP, FAST, false, TAME, ROT>(a, s, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
}

}  // namespace frh
