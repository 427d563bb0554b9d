// Bilinear develop, register row walker with four columns per lane (gfx950): the RGBA8 fast path
// of frames whose rows and output pixels are aligned to four samples.
#pragma once

#include "bil_walk.h"

namespace frh {

// Lane j owns frame columns x0 = xs + 4j .. x0 + 3 (xs a multiple of 4), loaded as one 8-byte word
// per row and stored as one 16-byte RGBA8 word; the horizontal neighbours x0 - 1 and x0 + 4 come
// from the adjacent lanes (DPP). Output columns of a strip: [xs + 2, xs + 254), 252 per wave, with
// xs + 1 and xs + 254 its halo columns. The method, the phase templates and the exact f32 operation
// order are those of bwalk_row (process.rs:1076-1096).
constexpr int BWALK4_COLS = 252;

// True when a launch can take the four-column walker: the RGBA8 fast path with the scalar
// calibration, 16-byte aligned output pixels (destination width and crop a multiple of 4) and
// 8-byte aligned input rows (stride a multiple of 4, source address a multiple of 8).
__host__ __device__ inline bool walk4_ok(const DevArgs &a) {
    return fast_rgba8(a) && a.simple_cal && (a.dst_w & 3u) == 0 && (a.crop_x & 3u) == 0 && (a.stride & 3u) == 0 &&
           ((uintptr_t)a.src & 7u) == 0 && !a.srcf;
}

struct BWalk4 {
    float n[3][4];  // normalised rows R-2, R-1, R; [k] column x0 + k
    float nl[3];    // column x0 - 1 (lane - 1's k = 3)
    float nr[3];    // column x0 + 4 (lane + 1's k = 0)
    uint32_t pf[BPF][2];  // raw quads fetched ahead, slot (row - start) % BPF
};

// Per-lane constants of a strip: columns inside the frame (in[k]) and output columns (st[k]).
struct B4StripCols {
    bool in[4], st[4];
};

// Raw samples x0 .. x0 + 3 of row y as two words (x0 in the low half of the first). HOT: one
// aligned 8-byte load inside the frame; otherwise per-sample loads at clamped columns.
template <bool HOT>
__device__ __forceinline__ void fetch_quad(const DevArgs &a, __amdgpu_buffer_rsrc_t rs, int x0, int y,
                                           uint32_t q[2]) {
    const int yc = min(max(y, 0), (int)a.height - 1);
    const int soff = (int)uniform((uint32_t)yc * a.stride * 2u);
    if (HOT) {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, x0 * 2, soff, 0);
        q[0] = v.x;
        q[1] = v.y;
        return;
    }
    uint32_t s[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int c = min(max(x0 + k, 0), (int)a.width - 1);
        s[k] = __builtin_amdgcn_raw_buffer_load_b16(rs, c * 2, soff, 0);
    }
    q[0] = s[0] | (s[1] << 16);
    q[1] = s[2] | (s[3] << 16);
}

template <int RP, int PY, int SLOT, bool HOT, bool TAME, bool OUT = true>
__device__ __forceinline__ void bwalk4_row(const DevArgs &a, const Lds &s, BWalk4 &W, int R, int xs, int lane,
                                           int out_row0, int out_row1, const B4StripCols &C,
                                           __amdgpu_buffer_rsrc_t rs, __amdgpu_buffer_rsrc_t rd) {
    const int x0 = xs + 4 * lane;
#pragma unroll
    for (int r = 0; r < 2; r++) {
#pragma unroll
        for (int k = 0; k < 4; k++) W.n[r][k] = W.n[r + 1][k];
        W.nl[r] = W.nl[r + 1];
        W.nr[r] = W.nr[r + 1];
    }
    const uint32_t raw[2] = {W.pf[SLOT][0], W.pf[SLOT][1]};
    // rows past out_row1 feed no output: their prefetch re-reads row out_row1 (a cache hit)
    fetch_quad<HOT>(a, rs, x0, min(R + BPF, out_row1), W.pf[SLOT]);
    {
        // Calib::norm of the four samples (x0 even: columns of parity k & 1); 0 outside the frame
        const bool row_in = R >= 0 && R < (int)a.height;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float v = (float)((raw[k >> 1] >> (16 * (k & 1))) & 0xffffu);
            const float n = (v - a.qb[PY * 2 + (k & 1)]) * a.qs[PY * 2 + (k & 1)];
            W.n[2][k] = row_in && C.in[k] ? n : 0.0f;
        }
    }
    W.nl[2] = from_left(W.n[2][3]);
    W.nr[2] = from_right(W.n[2][0]);
    if (!OUT) return;
    const int ro = R - 1;
    constexpr int PO = PY ^ 1;
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float own = W.n[1][k];
        const float l = k == 0 ? W.nl[1] : W.n[1][k - 1], r = k == 3 ? W.nr[1] : W.n[1][k + 1];
        const float u = W.n[0][k], d = W.n[2][k];
        float rgb[3];
        const int col = site<RP>(PO, k & 1);
        if (col == 1) {
            const int ch = site<RP>(PO, (k & 1) ^ 1);  // the colour of this row's other sites
            const float hv = (l + r) * 0.5f, vv = (u + d) * 0.5f;
            rgb[0] = ch == 0 ? hv : vv;
            rgb[1] = own;
            rgb[2] = ch == 0 ? vv : hv;
        } else {
            const float ul = k == 0 ? W.nl[0] : W.n[0][k - 1], ur = k == 3 ? W.nr[0] : W.n[0][k + 1];
            const float dl = k == 0 ? W.nl[2] : W.n[2][k - 1], dr = k == 3 ? W.nr[2] : W.n[2][k + 1];
            const float orth = (l + r + u + d) * 0.25f;
            const float diag = (ul + ur + dl + dr) * 0.25f;
            rgb[0] = col == 0 ? own : diag;
            rgb[1] = orth;
            rgb[2] = col == 0 ? diag : own;
        }
        o[k] = TAME ? finish_rgba8_scaled(a, s, rgb[0] * a.wb[0], rgb[1] * a.wb[1], rgb[2] * a.wb[2])
                    : finish_rgba8_fast(a, s, rgb[0], rgb[1], rgb[2], false);
    }
    // Every lane finishes its four pixels; a lane whose four are all output columns stores them as
    // one 16-byte word (aligned: xs, crop_x and dst_w are multiples of 4), the others store their
    // runs in pairs and single words. Stores that must not happen address past the destination
    // resource, where the hardware drops them: no branch.
    const bool rv = ro >= out_row0 && ro < out_row1;
    bool v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = rv && C.st[k];
    const bool all = v[0] & v[1] & v[2] & v[3];
    const uint32_t off = ((uint32_t)(ro - (int)a.crop_y) * a.dst_w + (uint32_t)(x0 - (int)a.crop_x)) * 4u;
    constexpr uint32_t PAST = 0xfffffff0u;
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4 w4;
    w4.x = o[0];
    w4.y = o[1];
    w4.z = o[2];
    w4.w = o[3];
    u32x2 lo, hi;
    lo.x = o[0];
    lo.y = o[1];
    hi.x = o[2];
    hi.y = o[3];
    __builtin_amdgcn_raw_buffer_store_b128(w4, rd, all ? (int)off : (int)PAST, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b64(lo, rd, (v[0] & v[1] & !all) ? (int)off : (int)PAST, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b64(hi, rd, (v[2] & v[3] & !all) ? (int)(off + 8u) : (int)PAST, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(o[0], rd, (v[0] & !v[1]) ? (int)off : (int)PAST, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(o[1], rd, (v[1] & !v[0]) ? (int)(off + 4u) : (int)PAST, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(o[2], rd, (v[2] & !v[3]) ? (int)(off + 8u) : (int)PAST, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b32(o[3], rd, (v[3] & !v[2]) ? (int)(off + 12u) : (int)PAST, 0, 0);
}

// One wave's strip, walked as bwalk_strip's fast path: three prologue rows, whole four-row trips of
// straight-line code, at most three rows more.
template <int RP, bool HOT, bool TAME>
__device__ __forceinline__ void bwalk4_strip(const DevArgs &a, const Lds &s, int xs, int lane, int row0,
                                             int out_row0, int out_row1, int out_x0, int out_x1) {
    BWalk4 W;
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (int k = 0; k < 4; k++) W.n[r][k] = 0.0f;
        W.nl[r] = W.nr[r] = 0.0f;
    }
    const int x0 = xs + 4 * lane;
    B4StripCols C;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        C.in[k] = x0 + k >= 0 && x0 + k < (int)a.width;
        C.st[k] = x0 + k >= out_x0 && x0 + k < out_x1;
    }
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(a.src, src_bytes(a));
    const __amdgpu_buffer_rsrc_t rd = make_rsrc(a.dst, a.dst_w * a.dev_h * 4u);
#pragma unroll
    for (int i = 0; i < BPF; i++) fetch_quad<HOT>(a, rs, x0, min(row0 - 2 + i, out_row1), W.pf[i]);
    bwalk4_row<RP, 0, 0, HOT, TAME, false>(a, s, W, row0 - 2, xs, lane, out_row0, out_row1, C, rs, rd);
    bwalk4_row<RP, 1, 1, HOT, TAME, false>(a, s, W, row0 - 1, xs, lane, out_row0, out_row1, C, rs, rd);
    bwalk4_row<RP, 0, 2, HOT, TAME, false>(a, s, W, row0, xs, lane, out_row0, out_row1, C, rs, rd);
    int R = row0 + 1;
    for (; R + 3 <= out_row1; R += 4) {
        bwalk4_row<RP, 1, 3, HOT, TAME>(a, s, W, R, xs, lane, out_row0, out_row1, C, rs, rd);
        bwalk4_row<RP, 0, 0, HOT, TAME>(a, s, W, R + 1, xs, lane, out_row0, out_row1, C, rs, rd);
        bwalk4_row<RP, 1, 1, HOT, TAME>(a, s, W, R + 2, xs, lane, out_row0, out_row1, C, rs, rd);
        bwalk4_row<RP, 0, 2, HOT, TAME>(a, s, W, R + 3, xs, lane, out_row0, out_row1, C, rs, rd);
    }
    if (R <= out_row1) bwalk4_row<RP, 1, 3, HOT, TAME>(a, s, W, R, xs, lane, out_row0, out_row1, C, rs, rd);
    if (R + 1 <= out_row1) bwalk4_row<RP, 0, 0, HOT, TAME>(a, s, W, R + 1, xs, lane, out_row0, out_row1, C, rs, rd);
    if (R + 2 <= out_row1) bwalk4_row<RP, 1, 1, HOT, TAME>(a, s, W, R + 2, xs, lane, out_row0, out_row1, C, rs, rd);
}

// The four-column walker's work split: strips of BWALK4_COLS output columns from xs0 + 2, xs0 a
// multiple of 4 at most x_first - 2, over the interior of bilinear plan P.
__host__ inline BWalkPlan make_bwalk4_plan(const DevArgs &a, int rows_per_strip) {
    BWalkPlan P = make_bwalk_plan(a, rows_per_strip);
    if (P.strips > 0) {
        P.xs0 = (P.x_first - 2) & ~3;
        P.strips = (P.x_end - (P.xs0 + 2) + BWALK4_COLS - 1) / BWALK4_COLS;
    }
    return P;
}

template <int RP, bool TAME>
__device__ __forceinline__ void bwalk4_item(const DevArgs &a, const BWalkPlan &P, const Lds &s, int item) {
    const int lane = threadIdx.x & 63;
    const int sx = item % P.strips, sy = item / P.strips;
    const int xs = P.xs0 + BWALK4_COLS * sx;  // lane 0 holds columns xs .. xs + 3
    const int row0 = P.ys0 + P.rows * sy;
    const int out_row0 = row0 > P.y_first ? row0 : P.y_first;
    const int out_row1 = row0 + P.rows < P.y_end ? row0 + P.rows : P.y_end;
    const int out_x0 = max(xs + 2, P.x_first), out_x1 = min(xs + 2 + BWALK4_COLS, P.x_end);
    const bool hot = xs >= 0 && xs + 256 <= (int)a.width;
    if (hot)
        bwalk4_strip<RP, true, TAME>(a, s, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
    else
        bwalk4_strip<RP, false, TAME>(a, s, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
}

}  // namespace frh
