// AHD develop, register-resident row walker (gfx950).
//
// The method of crates/flashraw/src/ahd.rs:191-393 (two green candidates, colour differences,
// square-root Lab, homogeneity counts, proportional blend) fused with the finish and the encode,
// with the same f32 operations in the same order. Instead of staging planes in LDS, one wave
// walks down a strip of the frame: lane j owns frame columns x_s + 2j and x_s + 2j + 1 (an even
// and an odd column), every stage keeps the rows it still needs in registers, and horizontal
// neighbours come from the adjacent lanes through DPP wave shifts. With the Bayer phase a
// template parameter and the row loop unrolled by two, the colour of every sample a lane touches
// is a compile-time constant, so each pixel evaluates only its own formulas.
//
// Per iteration the wave loads mosaic row R and advances each stage one row:
//   green candidates of row R-2, candidate RGB + Lab of row R-3, pair steps between rows R-4 and
//   R-3 (each step computed once and read from both ends, ahd.rs:300-351), homogeneity counts of
//   row R-4, and the blend, finish and encode of row R-5.
// Lanes 3..60 hold complete results (116 output columns per strip); the halo lanes only feed
// their neighbours. Pixels nearer the frame border than the AHD halo are developed bilinearly by
// extra blocks of the same launch (strip_to, process.rs:1183-1224).
#pragma once

#include "develop_common.h"

namespace frh {

constexpr int WALK_COLS = 116;  // output columns per strip (lanes 3..60)
constexpr int WALK_WAVES = 4;   // waves per block of the batch kernels

// Lightness step and squared chroma distance between two pixels of one candidate
// (ahd.rs:172-180): |L_q - L_p| and (a_q - a_p)^2 + (b_q - b_p)^2 (the same from either end).
struct Step {
    float l, c;
};
__device__ __forceinline__ Step step(float pl, float pa, float pb, float ql, float qa, float qb) {
    Step s;
    s.l = fabsf(ql - pl);
    float da = qa - pa, db = qb - pb;
    s.c = da * da + db * db;
    return s;
}
__device__ __forceinline__ Step step_left(Step s) { return Step{from_left(s.l), from_left(s.c)}; }
__device__ __forceinline__ Step step_right(Step s) { return Step{from_right(s.l), from_right(s.c)}; }

// Everything a lane carries from one row to the next. Index k: 0 the even column, 1 the odd one.
// The walker runs in trips of six rows; inside a trip every buffer is a ring indexed with the trip
// position T (a compile-time constant), so no register is ever copied from one row to the next.
struct Walk {
    float m[6][2];          // balanced mosaic, row R - k in slot (T - k) mod 6
    float g[2][2][2];       // [slot][cand][k] green candidates (0 H, 1 V): R-2 in T mod 2, R-3 in the other
    float df[3][2][2];      // [slot][cand][k] mosaic minus green: R-2 in T mod 3, R-3 in T+2, R-4 in T+1
    float dfl[3][2];        // [slot][cand] df of the odd column of lane - 1
    float dfr[3][2];        // [slot][cand] df of the even column of lane + 1
    float lab[2][2][3][2];  // [slot][cand][L, a, b][k] of a row: R-3 in T mod 2, R-4 in the other
    Step sd[2][2][2], sdl[2][2][2], sdr[2][2][2];  // [slot][cand][k] steps down / down-left / down-right
    Step sr[2][2][2];                              // [slot][cand][k] steps to the right
    uint32_t hom[2][2];     // [slot][k] homogeneity counts H + (V << 16): R-6 in T mod 2, R-5 in the other
    float u[3][2][2][2];    // [slot][cand][k][2] interpolated channels: R-3 in T mod 3, R-4 T+2, R-5 T+1
    uint32_t pf[2];         // raw pairs fetched two rows ahead, slot = row parity (consumed in place)
    uint32_t hold[2];       // OUT 3: the previous output row's two pixels
};

// Per-lane constants of a strip: its columns x0, x0 + 1 are inside the frame (in0, in1) and are
// output columns of the strip (st0, st1); pair_even[p]: at trip position parity p the lane's two
// output pixels start at an even destination pixel (one 8-byte store).
struct StripCols {
    bool in0, in1, st0, st1;
    bool pair_even[2];
    float inf;  // +inf (opaque_inf)
    // rotated RGBA8 (OUT 2, 3): destination pixel index of the lane's first pixel on the crop's
    // first row, and its steps per developed column and row (modulo 2^32); OUT 3: rot_lo is 1 when
    // the previous row's pixel of a column has the lower destination index (cy = +1)
    uint32_t rot_lane, rot_cx, rot_cy, rot_lo;
};

// One walker row at trip position T: load mosaic row R (parity T & 1) and advance every stage.
// OUT (with FAST and TAME): 0 RGBA8 unrotated; 1 RGBA8 to `rd` and linear RGB16 to `rl`; 2 RGBA8
// under EXIF orientation 2..4 (rows stay rows); 3 RGBA8 under orientation 5..8 (rows become
// columns, pixels of consecutive rows stored in pairs).
template <int RP, int T, bool FAST, bool HOT, bool TAME, int OUT = 0>
__device__ __forceinline__ void walk_row(const DevArgs &a, const Lds &s, const float *wt, Walk &W, int R, int xs,
                                         int lane, int out_row0, int out_row1, const StripCols &C,
                                         __amdgpu_buffer_rsrc_t rs, __amdgpu_buffer_rsrc_t rd,
                                         __amdgpu_buffer_rsrc_t rl) {
    constexpr int PY = T & 1;
    // ring slots
    constexpr int M0 = T % 6, M1 = (T + 5) % 6, M2 = (T + 4) % 6, M3 = (T + 3) % 6, M4 = (T + 2) % 6,
                  M5 = (T + 1) % 6;                        // rows R, R-1, ..., R-5
    constexpr int G2 = T % 2, G3 = (T + 1) % 2;            // green rows R-2, R-3
    constexpr int D2 = T % 3, D3 = (T + 2) % 3, D4 = (T + 1) % 3;  // df rows R-2, R-3, R-4
    constexpr int LQ = T % 2, LP = (T + 1) % 2;            // lab / steps: new row (R-3), previous row
    constexpr int H6 = T % 2, H5 = (T + 1) % 2;            // hom rows R-6, R-5 (R-4 goes to H6)
    constexpr int U3 = T % 3, U5 = (T + 1) % 3;            // interpolated rows R-3 (new), R-5
    const int x0 = xs + 2 * lane;

    // ---- balanced mosaic of row R, clipped at 1 (ahd.rs:143-151): min(norm * wb[c], 1) ----
    {
        const uint32_t raw = W.pf[PY];
        W.pf[PY] = fetch_pair<HOT>(a, rs, x0, R + 2);
        float n0, n1;
        norm_pair<PY, HOT>(a, s, raw, x0, R, R >= 0 && R < (int)a.height, C.in0, C.in1, n0, n1);
        W.m[M0][0] = fmin_nn<TAME>(n0 * a.wb[site<RP>(PY, 0)], 1.0f, C.inf);
        W.m[M0][1] = fmin_nn<TAME>(n1 * a.wb[site<RP>(PY, 1)], 1.0f, C.inf);
    }

    // ---- green candidates of row R-2 (ahd.rs:225-244) ----
    constexpr int PG = PY;
    {
        const float me = W.m[M2][0], mo = W.m[M2][1];
        const float pe = from_left(me), po = from_left(mo);   // columns x0 - 2, x0 - 1
        const float ne = from_right(me), no = from_right(mo); // columns x0 + 2, x0 + 3
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const float c = k == 0 ? me : mo;
            float gh = c, gv = c;
            if (site<RP>(PG, k) != 1) {
                const float l = k == 0 ? po : me, r = k == 0 ? mo : ne;
                const float l2 = k == 0 ? pe : po, r2 = k == 0 ? ne : no;
                const float u = W.m[M3][k], dn = W.m[M1][k], u2 = W.m[M4][k], d2 = W.m[M0][k];
                const float eh = (l + r) * 0.5f + (2.0f * c - l2 - r2) * 0.25f;
                const float ev = (u + dn) * 0.5f + (2.0f * c - u2 - d2) * 0.25f;
                if (TAME) {
                    // clamp(e, min(l, r), max(l, r)) is the median of e, l, r (no NaN; no -0 arises
                    // in a tame frame, so the choice between equal zeros cannot differ)
                    gh = __builtin_amdgcn_fmed3f(eh, l, r);
                    gv = __builtin_amdgcn_fmed3f(ev, u, dn);
                } else {
                    gh = fminf(fmaxf(eh, fminf(l, r)), fmaxf(l, r));
                    gv = fminf(fmaxf(ev, fminf(u, dn)), fmaxf(u, dn));
                }
            }
            W.g[G2][0][k] = gh;
            W.g[G2][1][k] = gv;
            W.df[D2][0][k] = c - gh;
            W.df[D2][1][k] = c - gv;
        }
#pragma unroll
        for (int d = 0; d < 2; d++) {
            W.dfl[D2][d] = from_left(W.df[D2][d][1]);
            W.dfr[D2][d] = from_right(W.df[D2][d][0]);
        }
    }

    // ---- candidate RGB and Lab of row R-3 (ahd.rs:249-298) ----
    constexpr int PL = PY ^ 1;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float cm = W.m[M3][k];
        const float sm = sqrt_m0<TAME>(cm, C.inf);  // the measured channel, the same in both candidates
        const int col = site<RP>(PL, k);
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const float g = W.g[G3][d][k];
            // colour differences around (R-3, column k) in rows R-4, R-3, R-2
            auto dleft = [&](int slot) { return k == 0 ? W.dfl[slot][d] : W.df[slot][d][0]; };
            auto dright = [&](int slot) { return k == 1 ? W.dfr[slot][d] : W.df[slot][d][1]; };
            float red, gg, blue, u0, u1;
            if (col == 1) {
                const float row = g + 0.5f * (dleft(D3) + dright(D3));
                const float clm = g + 0.5f * (W.df[D4][d][k] + W.df[D2][d][k]);
                const bool hr = site<RP>(PL, k ^ 1) == 0;  // red neighbours along the row
                red = hr ? row : clm;
                blue = hr ? clm : row;
                gg = cm;
                u0 = red;
                u1 = blue;
            } else {
                const float diag = g + 0.25f * (dleft(D4) + dright(D4) + dleft(D2) + dright(D2));
                red = col == 0 ? cm : diag;
                blue = col == 0 ? diag : cm;
                gg = g;
                u0 = g;
                u1 = diag;
            }
            float sr, sg, sb;
            if (col == 0) {
                sr = sm;
                sg = sqrt_m0<TAME>(gg, C.inf);
                sb = sqrt_m0<TAME>(blue, C.inf);
            } else if (col == 1) {
                sr = sqrt_m0<TAME>(red, C.inf);
                sg = sm;
                sb = sqrt_m0<TAME>(blue, C.inf);
            } else {
                sr = sqrt_m0<TAME>(red, C.inf);
                sg = sqrt_m0<TAME>(gg, C.inf);
                sb = sm;
            }
            // Lightness (sr + 2 sg + sb) / 4. The tame form keeps 4 L: lightness only enters the
            // homogeneity test as differences compared with maxima of differences, and scaling by
            // 4 commutes with their rounding (no value of a tame frame comes near the subnormals).
            const float l4 = sr + 2.0f * sg + sb;
            W.lab[LQ][d][0][k] = TAME ? l4 : l4 * 0.25f;
            W.lab[LQ][d][1][k] = sr - sg;
            W.lab[LQ][d][2][k] = sb - sg;
            W.u[U3][d][k][0] = u0;
            W.u[U3][d][k][1] = u1;
        }
    }

    // ---- pair steps between rows R-4 (p) and R-3 (q); homogeneity of row R-4 (ahd.rs:300-351) ----
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const float(&p)[3][2] = W.lab[LP][d];
        const float(&q)[3][2] = W.lab[LQ][d];
        // q neighbours: even column of lane + 1, odd column of lane - 1
        float qn[3], qp[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            qn[c] = from_right(q[c][0]);
            qp[c] = from_left(q[c][1]);
        }
        W.sd[LQ][d][0] = step(p[0][0], p[1][0], p[2][0], q[0][0], q[1][0], q[2][0]);
        W.sd[LQ][d][1] = step(p[0][1], p[1][1], p[2][1], q[0][1], q[1][1], q[2][1]);
        W.sdr[LQ][d][0] = step(p[0][0], p[1][0], p[2][0], q[0][1], q[1][1], q[2][1]);
        W.sdr[LQ][d][1] = step(p[0][1], p[1][1], p[2][1], qn[0], qn[1], qn[2]);
        W.sdl[LQ][d][0] = step(p[0][0], p[1][0], p[2][0], qp[0], qp[1], qp[2]);
        W.sdl[LQ][d][1] = step(p[0][1], p[1][1], p[2][1], q[0][0], q[1][0], q[2][0]);
        W.sr[LQ][d][0] = step(q[0][0], q[1][0], q[2][0], q[0][1], q[1][1], q[2][1]);
        W.sr[LQ][d][1] = step(q[0][1], q[1][1], q[2][1], qn[0], qn[1], qn[2]);
    }
    // the eight neighbour steps of each pixel of row p, per candidate: up-left, up, up-right, left,
    // right (steps of the previous rows, LP), down-left, down, down-right (this row's, LQ)
    Step nb[2][2][8];
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const Step ul_e = step_left(W.sdr[LP][d][1]);  // down-right from (p-1, x0-1)
        const Step ur_o = step_right(W.sdl[LP][d][0]); // down-left from (p-1, x0+2)
        const Step l_e = step_left(W.sr[LP][d][1]);    // right step of (p, x0-1)
        nb[d][0][0] = ul_e;
        nb[d][0][1] = W.sd[LP][d][0];
        nb[d][0][2] = W.sdl[LP][d][1];
        nb[d][0][3] = l_e;
        nb[d][0][4] = W.sr[LP][d][0];
        nb[d][0][5] = W.sdl[LQ][d][0];
        nb[d][0][6] = W.sd[LQ][d][0];
        nb[d][0][7] = W.sdr[LQ][d][0];
        nb[d][1][0] = W.sdr[LP][d][0];
        nb[d][1][1] = W.sd[LP][d][1];
        nb[d][1][2] = ur_o;
        nb[d][1][3] = W.sr[LP][d][0];
        nb[d][1][4] = W.sr[LP][d][1];
        nb[d][1][5] = W.sdl[LQ][d][1];
        nb[d][1][6] = W.sd[LQ][d][1];
        nb[d][1][7] = W.sdr[LQ][d][1];
    }
    uint32_t hp[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        // thresholds: the smaller of the largest steps along the row in H (left, right) and along
        // the column in V (up, down)
        const Step hl = nb[0][k][3], hrr = nb[0][k][4], vu = nb[1][k][1], vd = nb[1][k][6];
        const float el =
            fmin_nn<TAME>(fmax_nn<TAME>(hl.l, hrr.l, C.inf), fmax_nn<TAME>(vu.l, vd.l, C.inf), C.inf);
        const float ec =
            fmin_nn<TAME>(fmax_nn<TAME>(hl.c, hrr.c, C.inf), fmax_nn<TAME>(vu.c, vd.c, C.inf), C.inf);
        uint32_t ch = 0, cv = 0;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            ch += (nb[0][k][n].l <= el && nb[0][k][n].c <= ec) ? 1u : 0u;
            cv += (nb[1][k][n].l <= el && nb[1][k][n].c <= ec) ? 1u : 0u;
        }
        hp[k] = ch | (cv << 16);
    }

    // ---- blend, finish and store row R-5 (ahd.rs:353-392) ----
    const uint32_t v0 = W.hom[H6][0] + W.hom[H5][0] + hp[0];
    const uint32_t v1 = W.hom[H6][1] + W.hom[H5][1] + hp[1];
    W.hom[H6][0] = hp[0];
    W.hom[H6][1] = hp[1];
    // 3x3 sums (the DPP reads happen with every lane active: a lane switched off by a branch
    // would hand its neighbour a 0)
    const uint32_t sums[2] = {from_left_u(v1) + v0 + v1, v0 + v1 + from_right_u(v0)};
    const int ro = R - 5;
    constexpr int PO = PY ^ 1;
    auto blend = [&](int k, float &r, float &g, float &b) {
        // w = sh / (sh + sv), 0.5 when both are 0 (ahd.rs:388-389), read from the exact table
        const float w = wt[(sums[k] & 0xffffu) * 73u + (sums[k] >> 16)];
        const float cm = W.m[M5][k];
        const int col = site<RP>(PO, k);
        float rgb[2][3];
#pragma unroll
        for (int d = 0; d < 2; d++) {
            const float u0 = W.u[U5][d][k][0], u1 = W.u[U5][d][k][1];
            rgb[d][0] = col == 0 ? cm : (col == 1 ? u0 : u1);
            rgb[d][1] = col == 1 ? cm : u0;
            rgb[d][2] = col == 2 ? cm : u1;
        }
        const float iw = 1.0f - w;
        r = w * rgb[0][0] + iw * rgb[1][0];
        g = w * rgb[0][1] + iw * rgb[1][1];
        b = w * rgb[0][2] + iw * rgb[1][2];
    };
    if (FAST) {
        // Every lane finishes its two pixels (the halo lanes' counts are still 0..8, so their table
        // and code indexes stay in range), and the stores of pixels outside the strip's output, or
        // of rows outside it (pipeline fill, the rows past the strip's end), address past the
        // destination resource, where the hardware drops them: no branch splits the trip.
        uint32_t o[2];
        float c[2][3];  // OUT 1: 4096 clamp(output, 0, 1) per channel
#pragma unroll
        for (int k = 0; k < 2; k++) {
            float r, g, b;
            blend(k, r, g, b);
            if (OUT == 1)
                o[k] = finish_rgba8_scaled_c(a, s, r, g, b, c[k]);
            else
                o[k] = TAME ? finish_rgba8_scaled(a, s, r, g, b) : finish_rgba8_fast(a, s, r, g, b, true);
        }
        const bool rv = ro >= out_row0 && ro < out_row1;
        // destination pixel index of the lane's first pixel, modulo 2^32 (-1 for a lane whose second
        // pixel starts a row: both outputs' offsets are multiples of it, so pixel 1 still lands at 0)
        const uint32_t pix = (uint32_t)(ro - (int)a.crop_y) * a.dst_w + (uint32_t)(x0 - (int)a.crop_x);
        const uint32_t off = pix * 4u;
        const bool s0 = rv && C.st0, s1 = rv && C.st1, pe = C.pair_even[PY];
        constexpr uint32_t PAST = 0xfffffff0u;
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        if (OUT == 2) {
            // rotated: the destination index is linear in the developed position (C0 + Cx dx + Cy dy).
            // One 4-byte store per pixel: pairing pixels into 8-byte stores here (adjacent pixels of
            // a row, or of two rows for the transposed orientations) costs 30-40 spilled registers.
            const uint32_t i0 = C.rot_lane + C.rot_cy * (uint32_t)(ro - (int)a.crop_y), i1 = i0 + C.rot_cx;
            __builtin_amdgcn_raw_buffer_store_b32(o[0], rd, s0 ? (int)(i0 * 4u) : (int)PAST, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(o[1], rd, s1 ? (int)(i1 * 4u) : (int)PAST, 0, 0);
        } else if (OUT == 3) {
            // A row whose pixel closes an aligned pair with the previous row's (lower index even)
            // stores both as 8 bytes when the previous row is the strip's; a row that opens one leaves
            // its pixel to the next row unless that row is past the strip's output; any other owned
            // pixel goes out alone.
            const uint32_t i0 = C.rot_lane + C.rot_cy * (uint32_t)(ro - (int)a.crop_y), i1 = i0 + C.rot_cx;
            const bool rp = ro - 1 >= out_row0, rn = ro + 1 < out_row1, sw = C.rot_lo != 0u;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const uint32_t ic = k ? i1 : i0, lo = ic - C.rot_lo;
                const bool own = k ? s1 : s0, closer = !(lo & 1u);
                u32x2 q;
                q.x = sw ? W.hold[k] : o[k];
                q.y = sw ? o[k] : W.hold[k];
                __builtin_amdgcn_raw_buffer_store_b64(q, rd, (own && closer && rp) ? (int)(lo * 4u) : (int)PAST, 0,
                                                      0);
                __builtin_amdgcn_raw_buffer_store_b32(o[k], rd,
                                                      (own && (closer ? !rp : !rn)) ? (int)(ic * 4u) : (int)PAST, 0, 0);
                W.hold[k] = o[k];
            }
        } else {
            u32x2 v;
            v.x = o[0];
            v.y = o[1];
            __builtin_amdgcn_raw_buffer_store_b64(v, rd, (pe && s0 && s1) ? (int)off : (int)PAST, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(o[0], rd, (s0 && !(pe && s1)) ? (int)off : (int)PAST, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(o[1], rd, (s1 && !(pe && s0)) ? (int)(off + 4u) : (int)PAST, 0,
                                                  0);
        }
        if (OUT == 1) {
            // linear RGB16, 6 bytes per pixel at the same pixel index (the outputs are unrotated and
            // dst_w = dev_w): a pair at an even pixel is one 12-byte store, anything else goes out
            // one 16-bit sample at a time
            uint32_t q[2][3];
#pragma unroll
            for (int k = 0; k < 2; k++)
#pragma unroll
                for (int ch = 0; ch < 3; ch++) q[k][ch] = lin16_of_4k(c[k][ch]);
            const uint32_t off6 = pix * 6u;
            const bool full = pe && s0 && s1;
            typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
            u32x3 t;
            t.x = q[0][0] | (q[0][1] << 16);
            t.y = q[0][2] | (q[1][0] << 16);
            t.z = q[1][1] | (q[1][2] << 16);
            __builtin_amdgcn_raw_buffer_store_b96(t, rl, full ? (int)off6 : (int)PAST, 0, 0);
#pragma unroll
            for (int k = 0; k < 2; k++)
#pragma unroll
                for (int ch = 0; ch < 3; ch++)
                    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)q[k][ch], rl,
                                                          (!full && (k ? s1 : s0)) ? (int)(off6 + 6u * k + 2u * ch)
                                                                                   : (int)PAST,
                                                          0, 0);
        }
    } else if (ro >= out_row0 && ro < out_row1 && lane >= 3 && lane <= 60) {
        float o[2][3];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            float r, g, b;
            blend(k, r, g, b);
            finish(a, r, g, b, true, o[k][0], o[k][1], o[k][2]);
        }
        const int dx = x0 - (int)a.crop_x, dy = ro - (int)a.crop_y;
        if (C.st0) store_both<false>(a, s, (uint32_t)dx, (uint32_t)dy, o[0][0], o[0][1], o[0][2]);
        if (C.st1) store_both<false>(a, s, (uint32_t)dx + 1, (uint32_t)dy, o[1][0], o[1][1], o[1][2]);
    }
}

// One wave's strip: mosaic row R runs from row0 - 6 (even) past the last output row
// out_row1 - 1 = R - 5, in whole trips of six rows (prefetch slot = row parity), so a trip is
// straight-line code the compiler schedules as one block; the rows walked past the strip's end
// read clamped rows and store nothing.
template <int RP, bool FAST, bool HOT, bool TAME, int OUT = 0>
__device__ __forceinline__ void walk_strip(const DevArgs &a, const Lds &s, const float *wt, int xs, int lane,
                                           int row0, int out_row0, int out_row1, int out_x0, int out_x1) {
    Walk W;
    {
        float *f = reinterpret_cast<float *>(&W);
        for (int i = 0; i < (int)(sizeof(Walk) / 4) - 2; i++) f[i] = 0.0f;
        W.hom[0][0] = W.hom[0][1] = W.hom[1][0] = W.hom[1][1] = 0u;
    }
    const int x0 = xs + 2 * lane;
    StripCols C;
    C.in0 = x0 >= 0 && x0 < (int)a.width;
    C.in1 = x0 + 1 >= 0 && x0 + 1 < (int)a.width;
    C.st0 = x0 >= out_x0 && x0 < out_x1;
    C.st1 = x0 + 1 >= out_x0 && x0 + 1 < out_x1;
    C.inf = TAME ? opaque_inf() : __builtin_inff();
    // an 8-byte store needs an even destination pixel: the column parity is the strip's, the row's
    // adds dst_w times the row parity (row R - 5 has parity PY ^ 1 at trip position PY)
    const uint32_t cpar = (uint32_t)(xs - (int)a.crop_x) & 1u;
#pragma unroll
    for (int py = 0; py < 2; py++)
        C.pair_even[py] = ((cpar + (uint32_t)(py ^ 1 ^ (int)(a.crop_y & 1u)) * a.dst_w) & 1u) == 0;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(a.src, src_bytes(a));
    // OUT 1: the RGBA8 and linear RGB16 destinations, in whichever order the outputs name them
    void *const d8 = (OUT == 1 && a.out_mode != OUT_RGBA8) ? a.dst2 : a.dst;
    void *const d16 = (OUT == 1 && a.out_mode == OUT_RGBA8) ? a.dst2 : a.dst;
    const __amdgpu_buffer_rsrc_t rd = make_rsrc(d8, a.dev_w * a.dev_h * 4u);
    const __amdgpu_buffer_rsrc_t rl = make_rsrc(OUT == 1 ? d16 : nullptr, OUT == 1 ? a.dev_w * a.dev_h * 6u : 0u);
    C.rot_lane = C.rot_cx = C.rot_cy = C.rot_lo = 0u;
    W.hold[0] = W.hold[1] = 0u;
    if (OUT == 2 || OUT == 3) {
        // destination index c0 + cx dx + cy dy (orient, process.rs:2035-2044)
        uint32_t c0, cx, cy;
        orient_linear(a, c0, cx, cy);
        C.rot_lane = c0 + cx * (uint32_t)(x0 - (int)a.crop_x);
        C.rot_lo = cy == 1u ? 1u : 0u;
        C.rot_cx = cx;
        C.rot_cy = cy;
    }
#pragma unroll
    for (int i = 0; i < 2; i++) W.pf[i] = fetch_pair<HOT>(a, rs, x0, row0 - 6 + i);
    const int trips = (out_row1 + 5 - (row0 - 6) + 5) / 6;
    int R = row0 - 6;
    for (int t = 0; t < trips; t++, R += 6) {
#define WALK_ROW(t) \
    walk_row<RP, t, FAST, HOT, TAME, OUT>(a, s, wt, W, R + t, xs, lane, out_row0, out_row1, C, rs, rd, rl);
        WALK_ROW(0)
        WALK_ROW(1)
        WALK_ROW(2)
        WALK_ROW(3)
        WALK_ROW(4)
        WALK_ROW(5)
#undef WALK_ROW
    }
}

// Work split of one AHD develop: strips of WALK_COLS columns x `rows` rows over the AHD
// interior of the developed band, plus the bilinear border pixels.
struct WalkPlan {
    int x_first, x_end, y_first, y_end;  // interior, frame coordinates
    int xs0, ys0;                        // even-aligned origins
    int strips, chunks, rows;
    Border border;
};

// The AHD walker's work split of the developed band of `a`, `rows_per_strip` rows per wave.
__host__ inline WalkPlan make_walk_plan(const DevArgs &a, int rows_per_strip) {
    WalkPlan P;
    P.border = make_border(a, (int)a.ahd_halo);
    P.rows = rows_per_strip < 2 ? 2 : (rows_per_strip & ~1);
    if (P.border.ix1 > P.border.ix0) {
        P.x_first = P.border.ix0 + (int)a.crop_x;
        P.x_end = P.border.ix1 + (int)a.crop_x;
        P.y_first = P.border.iy0 + (int)a.crop_y;
        P.y_end = P.border.iy1 + (int)a.crop_y;
        P.xs0 = P.x_first & ~1;
        P.ys0 = P.y_first & ~1;
        P.strips = (P.x_end - P.xs0 + WALK_COLS - 1) / WALK_COLS;
        P.chunks = (P.y_end - P.ys0 + P.rows - 1) / P.rows;
    } else {
        P.x_first = P.x_end = P.y_first = P.y_end = P.xs0 = P.ys0 = 0;
        P.strips = P.chunks = 0;
    }
    return P;
}

// Strip `item` (0 .. strips * chunks - 1) of plan P, developed by the calling wave.
template <int RP, bool FAST, bool TAME, int OUT = 0>
__device__ __forceinline__ void walk_item(const DevArgs &a, const WalkPlan &P, const Lds &s, const float *wt,
                                          int item) {
    const int lane = threadIdx.x & 63;
    const int sx = item % P.strips, sy = item / P.strips;
    const int xs = P.xs0 + WALK_COLS * sx - 6;  // lane 3 holds output column xs + 6
    const int row0 = P.ys0 + P.rows * sy;
    const int out_row0 = row0 > P.y_first ? row0 : P.y_first;
    const int out_row1 = row0 + P.rows < P.y_end ? row0 + P.rows : P.y_end;
    const int out_x0 = max(xs + 6, P.x_first), out_x1 = min(xs + 6 + WALK_COLS, P.x_end);
    const bool hot = a.simple_cal && xs >= 0 && xs + 128 <= (int)a.width && ((uintptr_t)a.src & 3u) == 0 &&
                     (a.stride & 1u) == 0;
    if (hot)
        walk_strip<RP, FAST, true, TAME, OUT>(a, s, wt, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
    else
        walk_strip<RP, FAST, false, TAME, OUT>(a, s, wt, xs, lane, row0, out_row0, out_row1, out_x0, out_x1);
}

}  // namespace frh
