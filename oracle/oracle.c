/*
 * CPU oracle: a C restatement of the reference's CPU develop path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Every function cites the reference lines it follows;
 * floating-point operations are performed in the same order as the reference, in IEEE single
 * precision, with no FMA contraction (build with -ffp-contract=off), so results are bit-identical
 * to the reference's scalar and AVX2 (no-FMA) builds.
 *
 * Rust semantics restated here:
 *   f32::max / f32::min  -> IEEE maxNum / minNum (a NaN operand yields the other operand)
 *   f32::clamp           -> `if x < lo {lo}; if x > hi {hi}; x` (NaN propagates)
 *   `as u32` / `as u16`  -> truncation, saturating, NaN -> 0
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------ */
/* Rust scalar semantics                                                                       */

static inline float rmax(float a, float b) {
    float m = b > a ? b : a;
    return a != a ? b : m;
}
static inline float rmin(float a, float b) {
    float m = b < a ? b : a;
    return a != a ? b : m;
}
static inline float rclamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
static inline uint32_t sat_u32(float x) {
    if (!(x >= 1.0f)) return 0;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}
static inline uint16_t sat_u16(float x) {
    if (!(x >= 1.0f)) return 0;
    if (x >= 65535.0f) return 65535;
    return (uint16_t)x;
}
static inline uint8_t sat_u8(float x) {
    if (!(x >= 1.0f)) return 0;
    if (x >= 255.0f) return 255;
    return (uint8_t)x;
}
static size_t sat_sub(size_t a, size_t b) { return a > b ? a - b : 0; }
static size_t zmin(size_t a, size_t b) { return a < b ? a : b; }
static size_t zmax(size_t a, size_t b) { return a > b ? a : b; }
static size_t zclamp(size_t x, size_t lo, size_t hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* ------------------------------------------------------------------------------------------ */
/* Tables                                                                                      */

/* process.rs:1903-1915: x = i / 4096 (f32), s = 12.92 x or 1.055 x^(1/2.4) - 0.055 (f32 powf),
 * entry = (s * 255 + 0.5).clamp(0, 255) as u8. */
void fro_srgb8_table(uint8_t out[4097]) {
    const float e = 1.0f / 2.4f;
    for (int i = 0; i <= 4096; i++) {
        float x = (float)i / 4096.0f;
        float s = (x <= 0.0031308f) ? 12.92f * x : 1.055f * powf(x, e) - 0.055f;
        out[i] = sat_u8(rclamp(s * 255.0f + 0.5f, 0.0f, 255.0f));
    }
}

/* process.rs:1922-1934: the curve in f64 at 16385 points, times 65535, stored as f32. */
void fro_srgb16_table(float out[16385]) {
    for (int i = 0; i <= 16384; i++) {
        double x = (double)i / 16384.0;
        double s = (x <= 0.0031308) ? 12.92 * x : 1.055 * pow(x, 1.0 / 2.4) - 0.055;
        out[i] = (float)(s * 65535.0);
    }
}

/* process.rs:1978-1991. */
uint16_t fro_f32_to_f16(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    uint32_t sign = (b >> 16) & 0x8000u;
    uint32_t a = b & 0x7fffffffu;
    uint32_t odd = (a >> 13) & 1u;
    uint32_t normal = (a + 0xc8000fffu + odd) >> 13;
    float fa;
    memcpy(&fa, &a, 4);
    float fs = fa + 0.5f;
    uint32_t sb;
    memcpy(&sb, &fs, 4);
    uint32_t sub = sb - 0x3f000000u;
    uint32_t h = (a < 0x38800000u) ? sub : normal;
    uint32_t big = (a > 0x7f800000u) ? 0x7e00u : 0x7bffu;
    h = (a >= 0x477ff000u) ? big : h;
    return (uint16_t)(sign | h);
}

/* ------------------------------------------------------------------------------------------ */
/* Develop kernel state (process.rs:819-849, kernel_with 1244-1343)                            */

#define NO_CHANNEL 3
#define HALO 6        /* ahd.rs:33 */
#define BAND_ROWS 64  /* ahd.rs:35 */
#define CHUNK 192     /* ahd.rs:37 */
#define STRIP 16      /* process.rs:809 */

typedef struct {
    /* Calib of each component (process.rs:246-299): black[c * cal_pw * cal_ph + pos] */
    uint32_t cal_pw, cal_ph;
    float *black, *scale;
    const float *rowd, *cold;
    uint32_t n_rowd, n_cold;
    /* CFA */
    uint32_t pw, ph;
    const uint8_t *cfa;
    uint32_t cfa_w;
    /* neighbour lists per CFA position and channel (process.rs:1256-1286) */
    int8_t ndx[64][4][9], ndy[64][4][9];
    int nn[64][4];
    float inv[64][4];
    int bayer;
    /* frame */
    const uint16_t *raw;
    const float *rawf;
    size_t w, h, ox, oy, cpp;
    int has_cfa, half;
    /* params */
    float wb[4];
    float m[9];
    int has_levels;
    float lscale, loffset;
    size_t crop_x, crop_y;
    /* AHD plan (ahd.rs:42-61) */
    int ahd;
    uint8_t col[4];
    /* edge-aware plan (edgeaware/mod.rs:189-230), Bayer */
    int edge;
    const fro_edge *et;
    /* nonzero taps of g0 per canonical position; taps where some direction is nonzero, per
     * position and colour (mod.rs:130-160) */
    uint8_t g0_idx[36][49];
    int g0_n[36];
    uint8_t dir_idx[36][3][49];
    int dir_n[36][3];
    int period;
} Kern;

static uint8_t chan_of(uint8_t colour) { return colour < 3 ? colour : NO_CHANNEL; }

/* CfaPattern::color_at (crates/flashraw-core/src/types.rs:222-224). */
static uint8_t cfa_colour(const Kern *k, size_t x, size_t y) {
    return k->cfa[(y % k->ph) * k->cfa_w + (x % k->pw)];
}

/* Channel of frame sample (x, y) (process.rs:971-973). */
static int colour_xy(const Kern *k, size_t x, size_t y) {
    return k->has_cfa ? chan_of(cfa_colour(k, x + k->ox, y + k->oy)) : 1;
}

static size_t gcd(size_t a, size_t b) { return b == 0 ? a : gcd(b, a % b); }
static size_t lcm(size_t a, size_t b) { return a / gcd(a, b) * b; }

/* Calib::new for every component (process.rs:256-286). Without a CFA the pattern is 1x1 and
 * the white level of component c is white[min(c, n - 1)]. */
static int calib_new(Kern *k, const fro_frame *f) {
    size_t rows = f->black_rows ? f->black_rows : 1, cols = f->black_cols ? f->black_cols : 1;
    size_t samples = f->black_samples ? f->black_samples : 1;
    size_t cw = f->cfa_w ? f->cfa_w : 1, ch = f->cfa_h ? f->cfa_h : 1;
    size_t pw = zmin(lcm(cw, cols), 64), ph = zmin(lcm(ch, rows), 64);
    k->cal_pw = (uint32_t)pw;
    k->cal_ph = (uint32_t)ph;
    k->black = (float *)malloc(k->cpp * pw * ph * sizeof(float));
    k->scale = (float *)malloc(k->cpp * pw * ph * sizeof(float));
    if (!k->black || !k->scale) return -1;
    for (size_t c = 0; c < k->cpp; c++) {
        for (size_t y = 0; y < ph; y++) {
            for (size_t x = 0; x < pw; x++) {
                size_t idx = ((y % rows) * cols + x % cols) * samples + c % samples;
                float bl = idx < f->n_black_values ? f->black_values[idx] : 0.0f;
                size_t colour = f->cfa_w ? f->cfa[(y % f->cfa_h) * f->cfa_w + x % f->cfa_w] : c;
                size_t wi = f->n_white > 1 ? zmin(colour, f->n_white - 1) : 0;
                uint32_t wv = wi < f->n_white ? f->white_values[wi] : (f->n_white ? f->white_values[0] : 0);
                float white = (float)wv;
                k->black[c * pw * ph + y * pw + x] = bl;
                k->scale[c * pw * ph + y * pw + x] = white > bl ? 1.0f / (white - bl) : 1.0f;
            }
        }
    }
    k->rowd = f->row_deltas;
    k->n_rowd = f->row_deltas ? f->n_row_deltas : 0;
    k->cold = f->col_deltas;
    k->n_cold = f->col_deltas ? f->n_col_deltas : 0;
    return 0;
}

/* Calib::norm (process.rs:288-298) of component c; x, y are sensor coordinates. */
static inline float calib_norm(const Kern *k, size_t c, float v, size_t x, size_t y) {
    size_t i = c * k->cal_pw * k->cal_ph + (y % k->cal_ph) * k->cal_pw + x % k->cal_pw;
    float b = k->black[i];
    if (k->n_rowd) b += (y < k->n_rowd) ? k->rowd[y] : 0.0f;
    if (k->n_cold) b += (x < k->n_cold) ? k->cold[x] : 0.0f;
    return (v - b) * k->scale[i];
}

/* Kernel::normalize_row (process.rs:881-945) into out[w * cpp]: each of its branches computes
 * calib_norm's value for every sample, (f32(v) - black) * scale, u16 or f32 samples. */
static void normalize_row(const Kern *k, size_t y, float *out) {
    const size_t n = k->w * k->cpp;
    if (k->cpp == 1 && !k->n_rowd && !k->n_cold && k->raw) {
        /* one pattern period of black / scale per row, then a modulo-free loop (process.rs:918-944) */
        const uint16_t *restrict row = k->raw + y * k->w;
        size_t period = k->cal_pw, base = ((y + k->oy) % k->cal_ph) * k->cal_pw;
        float blk[64], scl[64];
        for (size_t i = 0; i < period; i++) {
            blk[i] = k->black[base + (i + k->ox) % period];
            scl[i] = k->scale[base + (i + k->ox) % period];
        }
        if (period == 2) {
            for (size_t x = 0; x + 1 < k->w; x += 2) {
                out[x] = ((float)row[x] - blk[0]) * scl[0];
                out[x + 1] = ((float)row[x + 1] - blk[1]) * scl[1];
            }
            if (k->w & 1) out[k->w - 1] = ((float)row[k->w - 1] - blk[0]) * scl[0];
            return;
        }
        for (size_t x = 0, j = 0; x < k->w; x++) {
            out[x] = ((float)row[x] - blk[j]) * scl[j];
            if (++j == period) j = 0;
        }
        return;
    }
    for (size_t i = 0; i < n; i++) {
        float v = k->rawf ? k->rawf[y * n + i] : (float)k->raw[y * n + i];
        out[i] = calib_norm(k, i % k->cpp, v, i / k->cpp + k->ox, y + k->oy);
    }
}

/* finish_rgb, Highlights::Clip (process.rs:654-707). */
static inline void finish_rgb(const float rgb[3], const float wb[4], const float *m, int has_levels, float ls,
                              float lo, float o[3]) {
    float r = rclamp(rgb[0] * wb[0], 0.0f, 1.0f);
    float g = rclamp(rgb[1] * wb[1], 0.0f, 1.0f);
    float b = rclamp(rgb[2] * wb[2], 0.0f, 1.0f);
    o[0] = m[0] * r + m[1] * g + m[2] * b;
    o[1] = m[3] * r + m[4] * g + m[5] * b;
    o[2] = m[6] * r + m[7] * g + m[8] * b;
    if (has_levels) {
        for (int c = 0; c < 3; c++) o[c] = o[c] * ls - lo;
    }
}

/* Kernel::row_cols for a single-component CFA frame, not half size (process.rs:1044-1129):
 * develops output row `oyy` into out[i] for output columns i in [i0, i1). `rows` holds the
 * normalised frame rows y-1, y, y+1 (NULL outside the frame). */
static void row_cols(const Kern *k, const float *up, const float *mid, const float *dn, size_t oyy, float *out,
                     size_t i0, size_t i1) {
    size_t y = k->crop_y + oyy;
    if (!(k->cpp == 1 && k->has_cfa)) {
        /* no demosaic (process.rs:1038-1046): RGB components, or grey from component 0 */
        for (size_t i = i0; i < i1; i++) {
            size_t x = k->crop_x + i;
            float rgb[3];
            for (int c = 0; c < 3; c++) rgb[c] = mid[x * k->cpp + (k->cpp >= 3 ? (size_t)c : 0)];
            finish_rgb(rgb, k->wb, k->m, k->has_levels, k->lscale, k->loffset, out + i * 3);
        }
        return;
    }
    int interior_y = up && dn;
    const float *rows3[3] = {up, mid, dn};
    size_t py = (y + k->oy) % k->ph;
    int row_colors[2] = {colour_xy(k, (2 - k->ox % 2) % 2, y), colour_xy(k, (3 - k->ox % 2) % 2, y)};
    for (size_t i = i0; i < i1; i++) {
        size_t x = k->crop_x + i;
        float rgb[3];
        if (k->bayer && interior_y && x > 0 && x + 1 < k->w) {
            int c0 = row_colors[(x + k->ox) & 1];
            float own = mid[x];
            float l = mid[x - 1], r = mid[x + 1];
            float u = up[x], d = dn[x];
            if (c0 == 1) {
                int ch = row_colors[(x + k->ox + 1) & 1];
                rgb[1] = own;
                rgb[ch] = (l + r) * 0.5f;
                rgb[2 - ch] = (u + d) * 0.5f;
            } else {
                float orth = (l + r + u + d) * 0.25f;
                float diag = (up[x - 1] + up[x + 1] + dn[x - 1] + dn[x + 1]) * 0.25f;
                if (c0 == 0) {
                    rgb[0] = own; rgb[1] = orth; rgb[2] = diag;
                } else {
                    rgb[0] = diag; rgb[1] = orth; rgb[2] = own;
                }
            }
        } else {
            size_t pi = py * k->pw + (x + k->ox) % k->pw;
            if (interior_y && x > 0 && x + 1 < k->w) {
                for (int c = 0; c < 3; c++) {
                    float s = 0.0f;
                    for (int t = 0; t < k->nn[pi][c]; t++)
                        s += rows3[k->ndy[pi][c][t] + 1][(size_t)((ptrdiff_t)x + k->ndx[pi][c][t])];
                    rgb[c] = s * k->inv[pi][c];
                }
            } else {
                for (int c = 0; c < 3; c++) {
                    float s = 0.0f, n = 0.0f;
                    for (int t = 0; t < k->nn[pi][c]; t++) {
                        ptrdiff_t xx = (ptrdiff_t)x + k->ndx[pi][c][t];
                        if (xx < 0 || (size_t)xx >= k->w) continue;
                        const float *r = rows3[k->ndy[pi][c][t] + 1];
                        if (r) {
                            s += r[xx];
                            n += 1.0f;
                        }
                    }
                    rgb[c] = n > 0.0f ? s / n : 0.0f;
                }
            }
        }
        finish_rgb(rgb, k->wb, k->m, k->has_levels, k->lscale, k->loffset, out + i * 3);
    }
}

/* Kernel::row_cols, half size (process.rs:975-1036): output pixel i of row oyy averages, per
 * colour, the samples of the 2x2 block at (crop_x + 2i, crop_y + 2 oyy) inside the frame (rows
 * r0, r1; r1 NULL past the last row), in the order (0,0), (1,0), (0,1), (1,1). Monochrome frames
 * average into green and replicate it; multi-component frames average their first three. */
static void row_half(const Kern *k, const float *r0, const float *r1, size_t oyy, float *out, size_t ow) {
    size_t y0 = k->crop_y + oyy * 2;
    for (size_t i = 0; i < ow; i++) {
        size_t x0 = k->crop_x + i * 2;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, n[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (size_t dy = 0; dy < 2; dy++) {
            const float *row = dy ? r1 : r0;
            if (!row) continue;
            for (size_t dx = 0; dx < 2; dx++) {
                size_t x = x0 + dx;
                if (x >= k->w) continue;
                if (k->cpp == 1 && k->has_cfa) {
                    int c = colour_xy(k, x, y0 + dy);
                    acc[c] += row[x];
                    n[c] += 1.0f;
                } else if (k->cpp == 1) {
                    acc[1] += row[x];
                    n[1] += 1.0f;
                } else {
                    for (size_t c = 0; c < zmin(k->cpp, 3); c++) {
                        acc[c] += row[x * k->cpp + c];
                        n[c] += 1.0f;
                    }
                }
            }
        }
        float rgb[3];
        for (int c = 0; c < 3; c++) rgb[c] = n[c] > 0.0f ? acc[c] / n[c] : 0.0f;
        if (k->cpp == 1 && !k->has_cfa) {
            float g = acc[1] / rmax(n[1], 1.0f);
            rgb[0] = rgb[1] = rgb[2] = g;
        }
        finish_rgb(rgb, k->wb, k->m, k->has_levels, k->lscale, k->loffset, out + i * 3);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* AHD (ahd.rs:127-393)                                                                        */

typedef struct {
    float *l, *c, *dl, *dc, *xl, *xc, *yl, *yc;
} PairSteps;

typedef struct {
    size_t n, ec;
    float *m, *g[2], *rgb[2][3], *lab[2][3], *diff, *vsum[2], *hsum[2], *mask[2][4], *eps[2], *hom[2];
    PairSteps rd[2][2];
    float *pool;
} Work;

static int work_new(Work *wk, size_t er, size_t ec) {
    size_t n = er * ec;
    size_t total = n * (1 + 2 + 6 + 6 + 1 + 2 + 2) + ec * (2 + 8 + 2) + ec * 8 * 4;
    wk->pool = (float *)calloc(total, sizeof(float));
    if (!wk->pool) return -1;
    float *p = wk->pool;
#define TAKE(dst, cnt) do { dst = p; p += (cnt); } while (0)
    TAKE(wk->m, n);
    for (int d = 0; d < 2; d++) TAKE(wk->g[d], n);
    for (int d = 0; d < 2; d++) for (int c = 0; c < 3; c++) TAKE(wk->rgb[d][c], n);
    for (int d = 0; d < 2; d++) for (int c = 0; c < 3; c++) TAKE(wk->lab[d][c], n);
    TAKE(wk->diff, n);
    for (int d = 0; d < 2; d++) TAKE(wk->hsum[d], n);
    for (int d = 0; d < 2; d++) TAKE(wk->hom[d], n);
    for (int d = 0; d < 2; d++) TAKE(wk->vsum[d], ec);
    for (int r = 0; r < 2; r++) for (int q = 0; q < 4; q++) TAKE(wk->mask[r][q], ec);
    for (int d = 0; d < 2; d++) TAKE(wk->eps[d], ec);
    for (int a = 0; a < 2; a++)
        for (int d = 0; d < 2; d++) {
            PairSteps *s = &wk->rd[a][d];
            TAKE(s->l, ec); TAKE(s->c, ec); TAKE(s->dl, ec); TAKE(s->dc, ec);
            TAKE(s->xl, ec); TAKE(s->xc, ec); TAKE(s->yl, ec); TAKE(s->yc, ec);
        }
#undef TAKE
    wk->n = n;
    wk->ec = ec;
    return 0;
}

/* Plan::pair_steps (ahd.rs:164-186). */
static void pair_steps(float *lab[2][3], size_t ec, size_t i, PairSteps out[2]) {
    size_t x0 = 4, n = ec - 8;
    size_t base = i * ec + x0;
    for (int d = 0; d < 2; d++) {
        const float *pl = lab[d][0], *pa = lab[d][1], *pb = lab[d][2];
        const size_t offs[4] = {1, ec, ec + 1, ec - 1};
        float *ol[4] = {out[d].l, out[d].dl, out[d].xl, out[d].yl};
        float *oc[4] = {out[d].c, out[d].dc, out[d].xc, out[d].yc};
        for (int s = 0; s < 4; s++) {
            size_t off = offs[s];
            float *restrict o_l = ol[s] + x0, *restrict o_c = oc[s] + x0;
            const float *restrict L0 = pl + base, *restrict A0 = pa + base, *restrict B0 = pb + base;
            const float *restrict L1 = pl + base + off, *restrict A1 = pa + base + off, *restrict B1 = pb + base + off;
            #pragma GCC ivdep
            for (size_t x = 0; x < n; x++) {
                float l = L0[x], a = A0[x], b = B0[x];
                float ql = L1[x], qa = A1[x], qb = B1[x];
                o_l[x] = fabsf(ql - l);
                o_c[x] = (qa - a) * (qa - a) + (qb - b) * (qb - b);
            }
        }
    }
}

/* Plan::chunk (ahd.rs:191-393): band mosaic `m` (`er` rows of `bw`), chunk columns c0..c1,
 * output `out` of (er - 2 HALO) * (c1 - c0) pixels. */
static void ahd_chunk(const Kern *k, const float *m, size_t bw, size_t y0, size_t xa, size_t c0, size_t c1,
                      size_t er, Work *wk, float *out) {
    size_t ec = c1 - c0 + 2 * HALO;
    for (size_t i = 0; i < er; i++) memcpy(wk->m + i * ec, m + i * bw + c0, ec * sizeof(float));
    size_t py = (y0 + HALO) & 1, px = (xa + HALO + c0) & 1;
    for (int rk = 0; rk < 2; rk++) {
        for (size_t j = 0; j < ec; j++) {
            uint8_t c = k->col[rk * 2 + ((j + px) & 1)];
            uint8_t right = k->col[rk * 2 + ((j + 1 + px) & 1)];
            wk->mask[rk][0][j] = c == 1 ? 1.0f : 0.0f;
            wk->mask[rk][1][j] = c == 0 ? 1.0f : 0.0f;
            wk->mask[rk][2][j] = c == 2 ? 1.0f : 0.0f;
            wk->mask[rk][3][j] = (c == 1 && right == 0) ? 1.0f : 0.0f;
        }
    }
    /* 1. green along rows (H) and columns (V) (ahd.rs:225-244) */
    for (size_t i = 2; i < er - 2; i++) {
        size_t n = ec - 4, base = i * ec + 2;
        const float *restrict gm = wk->mask[(i + py) & 1][0] + 2;
        const float *restrict mm = wk->m;
        float *restrict gh = wk->g[0] + base, *restrict gv = wk->g[1] + base;
        #pragma GCC ivdep
        for (size_t x = 0; x < n; x++) {
            float c = mm[base + x], l = mm[base - 1 + x], r = mm[base + 1 + x];
            float l2 = mm[base - 2 + x], r2 = mm[base + 2 + x];
            float u = mm[base - ec + x], d = mm[base + ec + x];
            float u2 = mm[base - 2 * ec + x], d2 = mm[base + 2 * ec + x];
            float h = rmin(rmax((l + r) * 0.5f + (2.0f * c - l2 - r2) * 0.25f, rmin(l, r)), rmax(l, r));
            float v = rmin(rmax((u + d) * 0.5f + (2.0f * c - u2 - d2) * 0.25f, rmin(u, d)), rmax(u, d));
            gh[x] = gm[x] != 0.0f ? c : h;
            gv[x] = gm[x] != 0.0f ? c : v;
        }
    }
    /* 2. candidate RGB by colour differences and the perceptual coordinates (ahd.rs:249-298) */
    for (int d = 0; d < 2; d++) {
        {
            float *restrict df = wk->diff;
            const float *restrict mm = wk->m, *restrict gg = wk->g[d];
            for (size_t q = 0; q < er * ec; q++) df[q] = mm[q] - gg[q];
        }
        for (size_t i = 3; i < er - 3; i++) {
            size_t n = ec - 6, base = i * ec + 3;
            size_t rk = (i + py) & 1;
            const float *restrict gm = wk->mask[rk][0] + 3, *restrict rm = wk->mask[rk][1] + 3,
                                  *restrict bm = wk->mask[rk][2] + 3, *restrict hr = wk->mask[rk][3] + 3;
            const float *restrict df = wk->diff;
            const float *restrict c = wk->m + base, *restrict g = wk->g[d] + base;
            float *restrict pr = wk->rgb[d][0] + base, *restrict pg = wk->rgb[d][1] + base,
                  *restrict pb = wk->rgb[d][2] + base;
            float *restrict pl = wk->lab[d][0] + base, *restrict pa = wk->lab[d][1] + base,
                  *restrict pbb = wk->lab[d][2] + base;
            #pragma GCC ivdep
            for (size_t x = 0; x < n; x++) {
                size_t o = base + x;
                float ul = df[o - ec - 1], ur = df[o - ec + 1], dl = df[o + ec - 1], dr = df[o + ec + 1];
                float l = df[o - 1], r = df[o + 1], u = df[o - ec], dn = df[o + ec];
                float diag = g[x] + 0.25f * (ul + ur + dl + dr);
                float row = g[x] + 0.5f * (l + r);
                float col = g[x] + 0.5f * (u + dn);
                float on_green_r = hr[x] != 0.0f ? row : col;
                float on_green_b = hr[x] != 0.0f ? col : row;
                float red = gm[x] != 0.0f ? on_green_r : diag;
                red = rm[x] != 0.0f ? c[x] : red;
                float blue = gm[x] != 0.0f ? on_green_b : diag;
                blue = bm[x] != 0.0f ? c[x] : blue;
                pr[x] = red;
                pg[x] = g[x];
                pb[x] = blue;
                float sr = sqrtf(rmax(red, 0.0f)), sg = sqrtf(rmax(g[x], 0.0f)), sb = sqrtf(rmax(blue, 0.0f));
                pl[x] = (sr + 2.0f * sg + sb) * 0.25f;
                pa[x] = sr - sg;
                pbb[x] = sb - sg;
            }
        }
    }
    /* 3. and 4. thresholds and homogeneity (ahd.rs:300-351) */
    PairSteps *prev = wk->rd[0], *cur = wk->rd[1];
    pair_steps(wk->lab, ec, 4, prev);
    for (size_t i = 5; i < er - 5; i++) {
        pair_steps(wk->lab, ec, i, cur);
        size_t x0 = 5, n = ec - 10;
        float *el = wk->eps[0], *ecr = wk->eps[1];
        {
            const PairSteps *h = &cur[0], *v0 = &prev[1], *v1 = &cur[1];
            const float *restrict hl = h->l + x0, *restrict hc = h->c + x0, *restrict vl0 = v0->dl + x0,
                                  *restrict vl1 = v1->dl + x0, *restrict vc0 = v0->dc + x0, *restrict vc1 = v1->dc + x0;
            float *restrict elr = el, *restrict ecrr = ecr;
            #pragma GCC ivdep
            for (size_t x = 0; x < n; x++) {
                elr[x] = rmin(rmax(hl[x - 1], hl[x]), rmax(vl0[x], vl1[x]));
                ecrr[x] = rmin(rmax(hc[x - 1], hc[x]), rmax(vc0[x], vc1[x]));
            }
        }
        for (int d = 0; d < 2; d++) {
            const PairSteps *p = &prev[d], *q = &cur[d];
            float *cnt = wk->hom[d] + i * ec + x0;
            for (size_t x = 0; x < n; x++) cnt[x] = 0.0f;
            const float *sl[8] = {p->xl + x0 - 1, p->dl + x0, p->yl + x0 + 1, q->l + x0 - 1,
                                  q->l + x0,      q->yl + x0, q->dl + x0,     q->xl + x0};
            const float *sc[8] = {p->xc + x0 - 1, p->dc + x0, p->yc + x0 + 1, q->c + x0 - 1,
                                  q->c + x0,      q->yc + x0, q->dc + x0,     q->xc + x0};
            float *restrict cn = cnt;
            const float *restrict e0 = el, *restrict e1 = ecr;
            for (int t = 0; t < 8; t++) {
                const float *restrict a = sl[t], *restrict b = sc[t];
                for (size_t x = 0; x < n; x++) cn[x] += (float)((a[x] <= e0[x]) & (b[x] <= e1[x]));
            }
        }
        PairSteps *tmp = prev;
        prev = cur;
        cur = tmp;
    }
    /* 5. blend by summed homogeneity (ahd.rs:353-392) */
    size_t n = ec - 2 * HALO;
    for (int d = 0; d < 2; d++) {
        for (size_t i = HALO; i < er - HALO; i++) {
            size_t base = i * ec + HALO - 1, len = n + 2;
            const float *restrict u = wk->hom[d] + base - ec, *restrict c = wk->hom[d] + base,
                                  *restrict dn = wk->hom[d] + base + ec;
            float *restrict v = wk->vsum[d];
            for (size_t x = 0; x < len; x++) v[x] = u[x] + c[x] + dn[x];
            float *restrict hs = wk->hsum[d] + (i - HALO) * n;
            for (size_t x = 0; x < n; x++) hs[x] = v[x] + v[x + 1] + v[x + 2];
        }
    }
    for (size_t i = HALO; i < er - HALO; i++) {
        const float *restrict sh = wk->hsum[0] + (i - HALO) * n, *restrict sv = wk->hsum[1] + (i - HALO) * n;
        size_t base = i * ec + HALO;
        const float *restrict r0 = wk->rgb[0][0] + base, *restrict g0 = wk->rgb[0][1] + base,
                              *restrict b0 = wk->rgb[0][2] + base;
        const float *restrict r1 = wk->rgb[1][0] + base, *restrict g1 = wk->rgb[1][1] + base,
                              *restrict b1 = wk->rgb[1][2] + base;
        float *restrict o = out + (i - HALO) * n * 3;
        #pragma GCC ivdep
        for (size_t x = 0; x < n; x++) {
            float a = sh[x], b = sv[x];
            float s = a + b, q = a / s;
            float w = (s > 0.0f) ? q : 0.5f;
            o[x * 3 + 0] = w * r0[x] + (1.0f - w) * r1[x];
            o[x * 3 + 1] = w * g0[x] + (1.0f - w) * g1[x];
            o[x * 3 + 2] = w * b0[x] + (1.0f - w) * b1[x];
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Encoders with orientation (process.rs:1593-1780, 1786-1900, 2018-2051)                     */

typedef struct {
    int enc;
    int orientation;
    size_t w, h;   /* developed size */
    size_t dw;     /* destination width */
    void *out;
    uint8_t lut8[4097];
    float pairs[16384][2];
    int channels;
    /* the 8-bit table in use (the sRGB table or a curve's) and its segments */
    const uint8_t *t8;
    uint32_t n8;
    const fro_tone *tone;
} Encoder;

/* ToneMap's max / min (tone.rs:98-108): plain comparisons. */
static inline float tmax(float a, float b) { return a > b ? a : b; }
static inline float tmin(float a, float b) { return a < b ? a : b; }

/* ToneMap::position + read (tone.rs:148-162): the curve at linear x >= 0. */
static float tone_curve(const fro_tone *t, float x) {
    float tt = sqrtf(tmin(x, t->xmax)) * t->scale;
    float t1 = tmin(tt, (float)(t->n - 1));
    int i = (int)t1;
    float r = tt - (float)i;
    return t->table[2 * i] + t->table[2 * i + 1] * r;
}

/* ToneMap::apply (tone.rs:170-207). */
static void tone_apply(const fro_tone *t, const float p[3], float o[3]) {
    const float *m = t->m;
    float c0 = tmax(m[0] * p[0] + m[1] * p[1] + m[2] * p[2], 0.0f);
    float c1 = tmax(m[3] * p[0] + m[4] * p[1] + m[5] * p[2], 0.0f);
    float c2 = tmax(m[6] * p[0] + m[7] * p[1] + m[8] * p[2], 0.0f);
    float hi = tmax(tmax(c0, c1), c2), lo = tmin(tmin(c0, c1), c2);
    float fh = tone_curve(t, hi), fl = tone_curve(t, lo);
    float d = hi - lo;
    float k = d > 0.0f ? (fh - fl) / d : 0.0f;
    float b = fl - k * lo;
    for (int c = 0; c < 3; c++) o[c] = b * t->white[c] + k * p[c];
}

/* Destination pixel of developed pixel (sx, sy) under EXIF orientation o: the inverse of the
 * source mapping in encode_developed (process.rs:2035-2044). */
static inline void dest_of(const Encoder *e, size_t sx, size_t sy, size_t *ox, size_t *oy) {
    size_t w = e->w, h = e->h;
    switch (e->orientation) {
    case 2: *ox = w - 1 - sx; *oy = sy; break;
    case 3: *ox = w - 1 - sx; *oy = h - 1 - sy; break;
    case 4: *ox = sx; *oy = h - 1 - sy; break;
    case 5: *ox = sy; *oy = sx; break;
    case 6: *ox = h - 1 - sy; *oy = sx; break;
    case 7: *ox = h - 1 - sy; *oy = w - 1 - sx; break;
    case 8: *ox = sy; *oy = w - 1 - sx; break;
    default: *ox = sx; *oy = sy; break;
    }
}

static void encode_px(const Encoder *e, const float p_in[3], size_t sx, size_t sy) {
    size_t ox = sx, oy = sy;
    int oriented = e->enc == FRO_RGBA8 || e->enc == FRO_SRGB16 || e->enc == FRO_RGB8;
    float toned[3];
    const float *p = p_in;
    if (e->tone && oriented) { /* the display outputs (develop.wgsl store_rgb) */
        tone_apply(e->tone, p_in, toned);
        p = toned;
    }
    if (oriented) dest_of(e, sx, sy, &ox, &oy);
    size_t di = (oy * e->dw + ox) * (size_t)e->channels;
    switch (e->enc) {
    case FRO_RGBA8:
    case FRO_RGB8: {
        uint8_t *o = (uint8_t *)e->out + di;
        for (int c = 0; c < 3; c++) {
            uint32_t ix = sat_u32(rclamp(p[c], 0.0f, 1.0f) * (float)e->n8); /* lut8_run, process.rs:1712-1728 */
            o[c] = e->t8[ix < e->n8 ? ix : e->n8];
        }
        if (e->enc == FRO_RGBA8) o[3] = 255;
        break;
    }
    case FRO_SRGB16: {
        uint16_t *o = (uint16_t *)e->out + di;
        for (int c = 0; c < 3; c++) { /* srgb16_run, process.rs:1760-1780 */
            float x = rclamp(p[c], 0.0f, 1.0f) * 16384.0f;
            uint32_t i = sat_u32(x);
            if (i > 16383) i = 16383;
            float t = x - (float)i;
            o[c] = sat_u16(e->pairs[i][0] + e->pairs[i][1] * t + 0.5f);
        }
        break;
    }
    case FRO_LINEAR16: {
        uint16_t *o = (uint16_t *)e->out + di;
        for (int c = 0; c < 3; c++) o[c] = sat_u16(rclamp(p[c], 0.0f, 1.0f) * 65535.0f + 0.5f); /* 1971-1973 */
        break;
    }
    case FRO_LINEAR_F16: {
        uint16_t *o = (uint16_t *)e->out + di;
        for (int c = 0; c < 3; c++) o[c] = fro_f32_to_f16(p[c]);
        break;
    }
    default: {
        float *o = (float *)e->out + di;
        for (int c = 0; c < 3; c++) o[c] = p[c];
        break;
    }
    }
}

/* Encode developed pixels (sx, y) for sx in [x0, x0 + n) (a run of one developed row): the
 * same per-sample operations as encode_px, with the common unrotated cases as straight loops. */
static void encode_run(const Encoder *e, const float *restrict src, size_t n, size_t x0, size_t y) {
    if (e->orientation != 1 || e->tone || e->n8 != 4096) {
        for (size_t i = 0; i < n; i++) encode_px(e, src + i * 3, x0 + i, y);
        return;
    }
    size_t di = (y * e->dw + x0) * (size_t)e->channels;
    if (e->enc == FRO_RGBA8) {
        uint8_t *restrict o = (uint8_t *)e->out + di;
        uint32_t ix[96];
        for (size_t b0 = 0; b0 < n; b0 += 32) {
            size_t m = zmin(32, n - b0);
            for (size_t q = 0; q < m * 3; q++) ix[q] = sat_u32(rclamp(src[b0 * 3 + q], 0.0f, 1.0f) * 4096.0f);
            for (size_t i = 0; i < m; i++) {
                o[(b0 + i) * 4 + 0] = e->lut8[zmin(ix[i * 3 + 0], 4096)];
                o[(b0 + i) * 4 + 1] = e->lut8[zmin(ix[i * 3 + 1], 4096)];
                o[(b0 + i) * 4 + 2] = e->lut8[zmin(ix[i * 3 + 2], 4096)];
                o[(b0 + i) * 4 + 3] = 255;
            }
        }
        return;
    }
    for (size_t i = 0; i < n; i++) encode_px(e, src + i * 3, x0 + i, y);
}

/* ------------------------------------------------------------------------------------------ */
/* Develop driver (process.rs:1160-1225 strip_to, 1475-1591 develop*, 1786-1900 run_oriented)  */

static int build_kernel(Kern *k, const fro_frame *f, const fro_params *p) {
    memset(k, 0, sizeof(*k));
    if ((!f->u16 && !f->f32) || f->width == 0 || f->height == 0 || f->cfa_w > 8 || f->cfa_h > 8 ||
        (f->cfa_w == 0) != (f->cfa_h == 0))
        return -1;
    k->raw = f->u16;
    k->rawf = f->f32;
    k->cpp = f->cpp ? f->cpp : 1;
    k->has_cfa = f->cfa_w != 0;
    k->half = p->half_size != 0;
    k->w = f->width;
    k->h = f->height;
    k->ox = f->origin_x;
    k->oy = f->origin_y;
    k->pw = k->has_cfa ? f->cfa_w : 1;
    k->ph = k->has_cfa ? f->cfa_h : 1;
    k->cfa = f->cfa;
    k->cfa_w = k->pw;
    if (calib_new(k, f)) return -1;
    /* is_bayer (crates/flashraw-core/src/types.rs:243-252) */
    if (f->cfa_w == 2 && f->cfa_h == 2) {
        int cnt[8] = {0};
        for (int i = 0; i < 4; i++) cnt[f->cfa[i] & 7]++;
        k->bayer = cnt[0] == 1 && cnt[1] == 2 && cnt[2] == 1;
    }
    /* neighbour lists (process.rs:1256-1286) */
    for (size_t py = 0; py < k->ph; py++) {
        for (size_t px = 0; px < k->pw; px++) {
            size_t pi = py * k->pw + px;
            for (int dy = -1; dy <= 1; dy++) {
                for (int dx = -1; dx <= 1; dx++) {
                    int c = chan_of(cfa_colour(k, px + dx + k->pw * 4, py + dy + k->ph * 4));
                    int t = k->nn[pi][c]++;
                    k->ndx[pi][c][t] = (int8_t)dx;
                    k->ndy[pi][c][t] = (int8_t)dy;
                }
            }
            int own = chan_of(cfa_colour(k, px, py));
            k->nn[pi][own] = 1;
            k->ndx[pi][own][0] = 0;
            k->ndy[pi][own][0] = 0;
            for (int c = 0; c < 4; c++) k->inv[pi][c] = k->nn[pi][c] ? 1.0f / (float)k->nn[pi][c] : 0.0f;
        }
    }
    memcpy(k->wb, p->wb, sizeof(k->wb));
    memcpy(k->m, p->m, sizeof(k->m));
    k->has_levels = p->has_levels;
    k->lscale = p->levels_scale;
    k->loffset = p->levels_offset;
    if ((p->demosaic == FRO_AHD || p->demosaic == FRO_EDGE) && (k->cpp != 1 || !k->has_cfa || k->half)) return -2;
    if (p->demosaic == FRO_AHD) {
        /* ahd::Plan::new (ahd.rs:49-61) with the frame origin; the pattern must be Bayer. */
        if (f->cfa_w != 2 || f->cfa_h != 2) return -2;
        for (int i = 0; i < 4; i++) k->col[i] = chan_of(cfa_colour(k, k->ox + (i & 1), k->oy + (i >> 1)));
        const uint8_t *col = k->col;
        int greens = (col[0] == 1 && col[3] == 1) || (col[1] == 1 && col[2] == 1);
#define OTHERS(a, b) (((a) == 0 && (b) == 2) || ((a) == 2 && (b) == 0))
        int ok = greens && (col[0] == 1 ? OTHERS(col[1], col[2]) : OTHERS(col[0], col[3]));
#undef OTHERS
        if (!ok) return -2;
        k->ahd = 1;
    } else if (p->demosaic == FRO_EDGE) {
        const int P = p->edge && p->edge->period ? (int)p->edge->period : 2;
        if ((P != 2 && P != 6) || (int)f->cfa_w != P || (int)f->cfa_h != P || !p->edge || !p->edge->g0 ||
            !p->edge->dir)
            return -2;
        k->edge = 1;
        k->et = p->edge;
        k->period = P;
        for (int ph = 0; ph < P * P; ph++) {
            k->g0_n[ph] = 0;
            for (int i = 0; i < 49; i++)
                if (p->edge->g0[ph * 49 + i] != 0.0f) k->g0_idx[ph][k->g0_n[ph]++] = (uint8_t)i;
            for (int c = 0; c < 3; c++) {
                k->dir_n[ph][c] = 0;
                for (int i = 0; i < 49; i++) {
                    int any = 0;
                    for (int d = 0; d < 4; d++) any |= p->edge->dir[((ph * 3 + c) * 4 + d) * 49 + i] != 0.0f;
                    if (any) k->dir_idx[ph][c][k->dir_n[ph][c]++] = (uint8_t)i;
                }
            }
        }
    } else if (p->demosaic != FRO_BILINEAR) {
        return -1;
    }
    return 0;
}

static void free_kernel(Kern *k) {
    free(k->black);
    free(k->scale);
}

/* Normalised frame row y into rows[y - base] (cached by the caller). */
typedef struct {
    float *buf;
    size_t base, count;
} RowBlock;

static const float *rb_row(const RowBlock *rb, const Kern *k, size_t y) {
    if (y >= k->h) return NULL;
    return rb->buf + (y - rb->base) * k->w * k->cpp;
}


/* ------------------------------------------------------------------------------------------ */
/* Edge-aware demosaic, Bayer (crates/flashraw/src/edgeaware/mod.rs): develop_chroma 423-607,   */
/* demosaic 612-796, smooth 803-822. The reference develops column chunks of a band; a pixel    */
/* depends only on its neighbourhood (mod.rs:43-46), so this statement works on the whole band. */

#define EDGE_HALO 8  /* CORE_HALO 5 + 3 (mod.rs:66-71) */
#define EDGE_CORE 5

/* Canonical position index (row * 2 + column) of frame sample (x, y). */
static inline int edge_ph(const Kern *k, size_t x, size_t y) {
    const size_t P = (size_t)k->period;
    return (int)(((y + k->et->cy) % P) * P + (x + k->et->cx) % P);
}

/* The median of nine values (median_row, edgeaware/mod.rs:825-858: a selection, so no rounding). */
static float median9(const float v[9]) {
    float a[9];
    memcpy(a, v, sizeof(a));
    for (int i = 1; i < 9; i++)
        for (int j = i; j > 0 && a[j - 1] > a[j]; j--) {
            float t = a[j];
            a[j] = a[j - 1];
            a[j - 1] = t;
        }
    return a[4];
}

/* Camera RGB (balanced) of frame rows [y0, y1) x columns [xa, xb) into out (3 floats per pixel).
 * row(y) returns the normalised frame row y. Needs EDGE_HALO samples around. */
static void edge_band(const Kern *k, const RowBlock *rb, size_t y0, size_t y1, size_t xa, size_t xb, float *out) {
    const fro_edge *t = k->et;
    const size_t nr = y1 - y0, nc = xb - xa;
    const int h = t->beta > 0.0f ? 1 : 0;
    const size_t m = 2 + (size_t)h;
    /* 1. balanced mosaic over rows/columns +- (m + CORE) */
    const size_t MH = m + EDGE_CORE;
    const size_t mr = nr + 2 * MH, mc = nc + 2 * MH;
    const size_t Y0 = y0 - MH, X0 = xa - MH;
    float *M = (float *)malloc(mr * mc * sizeof(float));
    for (size_t i = 0; i < mr; i++) {
        const float *row = rb_row(rb, k, Y0 + i);
        for (size_t j = 0; j < mc; j++) {
            size_t x = X0 + j;
            float scale = k->wb[t->canon[edge_ph(k, x, Y0 + i)]];
            M[i * mc + j] = rmin(row[x] * scale, 1.0f);
        }
    }
#define MS(y, x) M[((y) - Y0) * mc + ((x) - X0)]
    /* 2. sqrt(G0) over +- (m + 2) */
    const size_t GH = m + 2, gr = nr + 2 * GH, gc = nc + 2 * GH;
    const size_t GY0 = y0 - GH, GX0 = xa - GH;
    float *G = (float *)malloc(gr * gc * sizeof(float));
    for (size_t i = 0; i < gr; i++)
        for (size_t j = 0; j < gc; j++) {
            size_t y = GY0 + i, x = GX0 + j;
            int ph = edge_ph(k, x, y);
            float acc = 0.0f;
            for (int n = 0; n < k->g0_n[ph]; n++) {
                int idx = k->g0_idx[ph][n], ty = idx / 7, tx = idx % 7;
                acc += t->g0[ph * 49 + idx] * MS(y + ty - 3, x + tx - 3);
            }
            G[i * gc + j] = sqrtf(rmax(acc, 0.0f));
        }
#define GS(y, x) G[((y) - GY0) * gc + ((x) - GX0)]
    /* 3. squared gradients over +- (m + 1), 4. their 3x3 sums and the direction weights over +- m */
    const size_t DH = m + 1, dr = nr + 2 * DH, dc = nc + 2 * DH;
    const size_t DY0 = y0 - DH, DX0 = xa - DH;
    float *D = (float *)malloc(4 * dr * dc * sizeof(float));
    const float r2 = 0.70710677f; /* FRAC_1_SQRT_2 */
    for (size_t i = 0; i < dr; i++)
        for (size_t j = 0; j < dc; j++) {
            size_t y = DY0 + i, x = DX0 + j;
            float dh = GS(y, x + 1) - GS(y, x - 1);
            float dv = GS(y + 1, x) - GS(y - 1, x);
            float d1 = (GS(y + 1, x + 1) - GS(y - 1, x - 1)) * r2;
            float d2 = (GS(y + 1, x - 1) - GS(y - 1, x + 1)) * r2;
            D[(0 * dr + i) * dc + j] = dh * dh;
            D[(1 * dr + i) * dc + j] = dv * dv;
            D[(2 * dr + i) * dc + j] = d1 * d1;
            D[(3 * dr + i) * dc + j] = d2 * d2;
        }
    const size_t er = nr + 2 * m, ec = nc + 2 * m, EY0 = y0 - m, EX0 = xa - m;
    float *P[3];
    for (int c = 0; c < 3; c++) P[c] = (float *)malloc(er * ec * sizeof(float));
    for (size_t i = 0; i < er; i++)
        for (size_t j = 0; j < ec; j++) {
            size_t y = EY0 + i, x = EX0 + j;
            float w[4];
            for (int d = 0; d < 4; d++) {
                float hs[3];
                for (int a = 0; a < 3; a++) {
                    const float *s0 = &D[(d * dr + (y - 1 + a - DY0)) * dc + (x - 1 - DX0)];
                    hs[a] = s0[0] + s0[1] + s0[2];
                }
                float u = hs[0] + hs[1] + hs[2] + 1e-3f;
                w[d] = 1.0f / (u * u);
            }
            float inv = 1.0f / (w[0] + w[1] + w[2] + w[3]);
            for (int d = 0; d < 4; d++) w[d] = w[d] * inv;
            /* 5. the directional filters for the two missing colours; the own colour is the sample */
            int ph = edge_ph(k, x, y);
            int own = t->canon[ph];
            for (int c = 0; c < 3; c++) {
                if (c == own) {
                    P[c][i * ec + j] = MS(y, x);
                    continue;
                }
                float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int n = 0; n < k->dir_n[ph][c]; n++) {
                    int idx = k->dir_idx[ph][c][n], ty = idx / 7, tx = idx % 7;
                    float v = MS(y + ty - 3, x + tx - 3);
                    for (int d = 0; d < 4; d++) a[d] += t->dir[((ph * 3 + c) * 4 + d) * 49 + idx] * v;
                }
                P[c][i * ec + j] = w[0] * a[0] + w[1] * a[1] + w[2] * a[2] + w[3] * a[3];
            }
        }
#define PS(c, y, x) P[c][((y) - EY0) * ec + ((x) - EX0)]
    /* 6. pass 1 over +- (1 + h): Bayer, the gated mean of the measured differences of the 3x3
     *    window; X-Trans, the median of the differences of the filtered planes (mod.rs:455-530) */
    const size_t H1 = 1 + (size_t)h, e1r = nr + 2 * H1, e1c = nc + 2 * H1, QY0 = y0 - H1, QX0 = xa - H1;
    float *d1[2];
    for (int cc = 0; cc < 2; cc++) d1[cc] = (float *)malloc(e1r * e1c * sizeof(float));
    for (size_t i = 0; i < e1r; i++)
        for (size_t j = 0; j < e1c; j++) {
            size_t y = QY0 + i, x = QX0 + j;
            float g = PS(1, y, x);
            int own = t->canon[edge_ph(k, x, y)];
            if (k->period != 2) {
                for (int cc = 0; cc < 2; cc++) {
                    const int c = cc ? 2 : 0;
                    float df[9];
                    for (int dy = 0; dy < 3; dy++)
                        for (int dx = 0; dx < 3; dx++)
                            df[dy * 3 + dx] = PS(c, y - 1 + dy, x - 1 + dx) - PS(1, y - 1 + dy, x - 1 + dx);
                    d1[cc][i * e1c + j] = own == c ? df[4] : median9(df);
                }
                continue;
            }
            float inv = 1.0f / (0.2f * rmax(g, 0.0f) + 0.015f);
            float sw[2] = {0, 0}, sd[2] = {0, 0}, su[2] = {0, 0}, n[2] = {0, 0};
            for (int dy = 0; dy < 3; dy++)
                for (int dx = 0; dx < 3; dx++) {
                    size_t yy = y - 1 + dy, xx = x - 1 + dx;
                    float gk = PS(1, yy, xx);
                    float e = (gk - g) * inv;
                    float wt = 1.0f / (1.0f + e * e);
                    int col = t->canon[edge_ph(k, xx, yy)];
                    for (int cc = 0; cc < 2; cc++) {
                        float mk = col == (cc ? 2 : 0) ? 1.0f : 0.0f;
                        float d = PS(cc ? 2 : 0, yy, xx) - gk;
                        sw[cc] += mk * wt;
                        sd[cc] += (mk * wt) * d;
                        su[cc] += mk * d;
                        n[cc] += mk;
                    }
                }
            for (int cc = 0; cc < 2; cc++) {
                float mixed = sw[cc] > 1e-12f ? sd[cc] / sw[cc] : su[cc] / n[cc];
                d1[cc][i * e1c + j] = own == (cc ? 2 : 0) ? PS(cc ? 2 : 0, y, x) - g : mixed;
            }
        }
    /* 7. pass 2 over +- h: the same gate over the differences of every pixel */
    const size_t e2r = nr + 2 * (size_t)h, e2c = nc + 2 * (size_t)h, VY0 = y0 - (size_t)h, VX0 = xa - (size_t)h;
    const float xg = k->period == 6 ? rclamp(1.0f - 2.0f * t->beta, 0.0f, 1.0f) : 0.0f;
    float *V = (float *)malloc(e2r * e2c * 3 * sizeof(float));
    for (size_t i = 0; i < e2r; i++)
        for (size_t j = 0; j < e2c; j++) {
            size_t y = VY0 + i, x = VX0 + j;
            float g = PS(1, y, x);
            float inv = 1.0f / (0.1f * rmax(g, 0.0f) + 0.007f);
            float sw = 0.0f, sdr = 0.0f, sdb = 0.0f;
            for (int dy = 0; dy < 3; dy++)
                for (int dx = 0; dx < 3; dx++) {
                    size_t yy = y - 1 + dy, xx = x - 1 + dx;
                    float gk = PS(1, yy, xx);
                    float e = (gk - g) * inv;
                    float wt = 1.0f / (1.0f + e * e);
                    size_t q = (yy - QY0) * e1c + (xx - QX0);
                    sw += wt;
                    sdr += wt * d1[0][q];
                    sdb += wt * d1[1][q];
                }
            float dr = sdr / sw, db = sdb / sw;
            int own = t->canon[edge_ph(k, x, y)];
            /* X-Trans: green at a red or blue site moves towards the measured sample minus the
             * cleaned difference, by xg = clamp(1 - 2 beta, 0, 1) (mod.rs:547-592) */
            float gv = g;
            if (xg != 0.0f && own == 0) gv = g + xg * ((PS(0, y, x) - dr) - g);
            if (xg != 0.0f && own == 2) gv = g + xg * ((PS(2, y, x) - db) - g);
            float *o = V + (i * e2c + j) * 3;
            o[0] = own == 0 ? PS(0, y, x) : gv + dr;
            o[1] = gv;
            o[2] = own == 2 ? PS(2, y, x) : gv + db;
        }
    /* 8. gated smoothing of the core on noisy frames (beta > 0) */
    for (size_t i = 0; i < nr; i++)
        for (size_t j = 0; j < nc; j++) {
            const float *pc = V + ((i + (size_t)h) * e2c + j + (size_t)h) * 3;
            float *o = out + (i * nc + j) * 3;
            if (!h) {
                o[0] = pc[0];
                o[1] = pc[1];
                o[2] = pc[2];
                continue;
            }
            float inv = 1.0f / (0.2f * rmax(pc[1], 0.0f) + 0.015f);
            float sw = 0.0f, sm[3] = {0.0f, 0.0f, 0.0f};
            for (int dy = 0; dy < 3; dy++)
                for (int dx = 0; dx < 3; dx++) {
                    const float *q = V + ((i + (size_t)dy) * e2c + j + (size_t)dx) * 3;
                    float e = (q[1] - pc[1]) * inv;
                    float wt = 1.0f / (1.0f + e * e);
                    sw += wt;
                    for (int c = 0; c < 3; c++) sm[c] += wt * q[c];
                }
            for (int c = 0; c < 3; c++) o[c] = pc[c] + t->beta * (sm[c] / sw - pc[c]);
        }
#undef MS
#undef GS
#undef PS
    free(M);
    free(G);
    free(D);
    for (int c = 0; c < 3; c++) free(P[c]);
    free(d1[0]);
    free(d1[1]);
    free(V);
}

int fro_develop(const fro_frame *f, const fro_params *p, void *out, uint64_t out_bytes, int threads,
                uint32_t *out_w, uint32_t *out_h) {
    Kern k;
    int rc = build_kernel(&k, f, p);
    if (rc) {
        free_kernel(&k);
        return -1;
    }
    /* output_dims (process.rs:1384-1398); Rect::clamp_to */
    size_t cx = zmin(p->crop_x, k.w), cy = zmin(p->crop_y, k.h);
    size_t ow = zmin((size_t)p->crop_w, k.w - cx), oh = zmin((size_t)p->crop_h, k.h - cy);
    if (k.half && ow && oh) { /* output_dims (process.rs:1384-1398) */
        ow /= 2;
        oh /= 2;
    }
    if (ow == 0 || oh == 0) {
        free_kernel(&k);
        return -1;
    }
    k.crop_x = cx;
    k.crop_y = cy;
    Encoder *e = (Encoder *)malloc(sizeof(Encoder));
    if (!e) {
        free_kernel(&k);
        return -1;
    }
    if (p->encoding < FRO_RGBA8 || p->encoding > FRO_RGB8) {
        free(e);
        free_kernel(&k);
        return -1;
    }
    e->enc = p->encoding;
    e->orientation = (p->orientation >= 1 && p->orientation <= 8) ? p->orientation : 1;
    e->w = ow;
    e->h = oh;
    int oriented = e->enc == FRO_RGBA8 || e->enc == FRO_SRGB16 || e->enc == FRO_RGB8;
    int swap = oriented && e->orientation >= 5;
    e->dw = swap ? oh : ow;
    e->out = out;
    e->channels = e->enc == FRO_RGBA8 ? 4 : 3;
    size_t bps = (e->enc == FRO_RGBA8 || e->enc == FRO_RGB8) ? 1 : (e->enc == FRO_LINEAR_F32 ? 4 : 2);
    if (out_bytes < (uint64_t)ow * oh * e->channels * bps) {
        free(e);
        free_kernel(&k);
        return -1;
    }
    fro_srgb8_table(e->lut8);
    e->t8 = e->lut8;
    e->n8 = 4096;
    if (p->curve8 && p->curve8_len > 1) { /* run.rs:672-675 */
        e->t8 = p->curve8;
        e->n8 = p->curve8_len - 1;
    }
    e->tone = (p->tone && p->tone->n) ? p->tone : NULL;
    {
        float *lut16 = (float *)malloc(16385 * sizeof(float));
        if (!lut16) {
            free(e);
            free_kernel(&k);
            return -1;
        }
        fro_srgb16_table(lut16);
        if (p->curve16 && p->curve16_len == 16385) memcpy(lut16, p->curve16, 16385 * sizeof(float)); /* run.rs:639-643 */
        for (int i = 0; i < 16384; i++) {
            e->pairs[i][0] = lut16[i];
            e->pairs[i][1] = lut16[i + 1] - lut16[i];
        }
        free(lut16);
    }
    if (out_w) *out_w = (uint32_t)(swap ? oh : ow);
    if (out_h) *out_h = (uint32_t)(swap ? ow : oh);

    /* AHD and edge-aware develop their interior in 64-row bands (ahd.rs:35, edgeaware/mod.rs:73) */
    const int banded = k.ahd || k.edge;
    const size_t halo = k.ahd ? HALO : (k.edge ? EDGE_HALO : 0);
    size_t strip = banded ? BAND_ROWS : STRIP;
    size_t nstrips = (oh + strip - 1) / strip;
    /* interior (strip_to, process.rs:1183-1190): output columns [ia, ib), rows [lo, hi). */
    size_t ia = zmin(sat_sub(halo, cx), ow);
    size_t ib = zmax(zmin(sat_sub(sat_sub(k.w, halo), cx), ow), ia);
    int nthreads = threads > 0 ? threads : omp_get_max_threads();
    int failed = 0;
#pragma omp parallel num_threads(nthreads)
    {
        /* per-thread buffers */
        size_t maxrows = 2 * strip + 2 * (halo > 1 ? halo : 1) + 2;
        RowBlock rb = {(float *)malloc(maxrows * k.w * k.cpp * sizeof(float)), 0, 0};
        float *rgbrow = (float *)malloc(ow * 3 * sizeof(float));
        float *band = NULL, *mband = NULL, *chunk_out = NULL;
        Work wk;
        memset(&wk, 0, sizeof(wk));
        int ok = rb.buf && rgbrow;
        if (ok && k.edge) {
            chunk_out = (float *)malloc(strip * (ib > ia ? ib - ia : 1) * 3 * sizeof(float));
            ok = chunk_out != NULL;
        }
        if (ok && k.ahd) {
            size_t er = strip + 2 * HALO;
            size_t bwmax = (ib > ia ? ib - ia : 0) + 2 * HALO;
            mband = (float *)malloc(er * bwmax * sizeof(float));
            chunk_out = (float *)malloc(strip * CHUNK * 3 * sizeof(float));
            band = (float *)malloc(k.w * sizeof(float));
            ok = mband && chunk_out && band && work_new(&wk, er, CHUNK + 2 * HALO) == 0;
        }
        if (!ok) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (size_t t = 0; t < nstrips; t++) {
            if (!ok) continue;
            size_t first = t * strip, count = zmin(strip, oh - first);
            /* normalised frame rows needed by the strip (bilinear: +-1, AHD: +-HALO) */
            size_t fy0 = sat_sub(cy + first, banded ? halo : 1);
            size_t fy1 = zmin(cy + first + count + (banded ? halo : 1), k.h);
            if (k.half) { /* the 2x2 blocks of the strip's output rows */
                fy0 = cy + 2 * first;
                fy1 = zmin(cy + 2 * (first + count), k.h);
            }
            rb.base = fy0;
            rb.count = fy1 - fy0;
            for (size_t y = fy0; y < fy1; y++) normalize_row(&k, y, rb.buf + (y - fy0) * k.w * k.cpp);
            size_t lo = 0, hi = 0;
            if (banded) {
                lo = zclamp(sat_sub(halo, cy), first, first + count);
                hi = zclamp(sat_sub(sat_sub(k.h, halo), cy), lo, first + count);
            }
            for (size_t i = 0; i < count && k.half; i++) {
                size_t oyy = first + i, y0 = cy + 2 * oyy;
                row_half(&k, rb_row(&rb, &k, y0), y0 + 1 < k.h ? rb_row(&rb, &k, y0 + 1) : NULL, oyy, rgbrow, ow);
                encode_run(e, rgbrow, ow, 0, oyy);
            }
            for (size_t i = 0; i < count && !k.half; i++) {
                size_t oyy = first + i, y = cy + oyy;
                const float *up = y > 0 ? rb_row(&rb, &k, y - 1) : NULL;
                const float *mid = rb_row(&rb, &k, y);
                const float *dn = y + 1 < k.h ? rb_row(&rb, &k, y + 1) : NULL;
                if (banded && oyy >= lo && oyy < hi && ib > ia) {
                    row_cols(&k, up, mid, dn, oyy, rgbrow, 0, ia);
                    row_cols(&k, up, mid, dn, oyy, rgbrow, ib, ow);
                    encode_run(e, rgbrow, ia, 0, oyy);
                    encode_run(e, rgbrow + ib * 3, ow - ib, ib, oyy);
                } else {
                    row_cols(&k, up, mid, dn, oyy, rgbrow, 0, ow);
                    encode_run(e, rgbrow, ow, 0, oyy);
                }
            }
            if (k.edge && hi > lo && ib > ia) {
                size_t y0 = cy + lo, y1 = cy + hi, xa = cx + ia, xb = cx + ib, nc = xb - xa;
                edge_band(&k, &rb, y0, y1, xa, xb, chunk_out);
                float kk[3] = {k.wb[0] / k.wb[0], k.wb[1] / k.wb[1], k.wb[2] / k.wb[2]};
                const float one[4] = {1.0f, 1.0f, 1.0f, 1.0f};
                for (size_t r = 0; r < y1 - y0; r++) {
                    float *s = chunk_out + r * nc * 3;
                    for (size_t j = 0; j < nc; j++) {
                        float in[3] = {s[j * 3] * kk[0], s[j * 3 + 1] * kk[1], s[j * 3 + 2] * kk[2]};
                        finish_rgb(in, one, k.m, k.has_levels, k.lscale, k.loffset, s + j * 3);
                    }
                    encode_run(e, s, nc, ia, lo + r);
                }
            }
            if (k.ahd && hi > lo && ib > ia) {
                /* develop_impl (ahd.rs:127-162) over frame rows [y0, y1), columns [xa, xb) */
                size_t y0 = cy + lo, y1 = cy + hi, xa = cx + ia, xb = cx + ib;
                size_t nr = y1 - y0, nc = xb - xa, er = nr + 2 * HALO, bw = nc + 2 * HALO;
                for (size_t i = 0; i < er; i++) {
                    size_t y = y0 + i - HALO;
                    const float *row = rb_row(&rb, &k, y);
                    for (size_t j = 0; j < bw; j++) {
                        size_t x = xa + j - HALO;
                        int c = k.col[(y & 1) * 2 + (x & 1)];
                        mband[i * bw + j] = rmin(row[x] * k.wb[c], 1.0f);
                    }
                }
                /* finish_run's k = wb / dm_wb (process.rs:1209); dm_wb = wb here */
                float kk[3] = {k.wb[0] / k.wb[0], k.wb[1] / k.wb[1], k.wb[2] / k.wb[2]};
                const float one[4] = {1.0f, 1.0f, 1.0f, 1.0f};
                for (size_t c0 = 0; c0 < nc; c0 += CHUNK) {
                    size_t c1 = zmin(c0 + CHUNK, nc), cw = c1 - c0;
                    ahd_chunk(&k, mband, bw, y0, xa, c0, c1, er, &wk, chunk_out);
                    for (size_t r = 0; r < nr; r++) {
                        float *s = chunk_out + r * cw * 3;
                        for (size_t j = 0; j < cw; j++) {
                            float in[3] = {s[j * 3] * kk[0], s[j * 3 + 1] * kk[1], s[j * 3 + 2] * kk[2]};
                            finish_rgb(in, one, k.m, k.has_levels, k.lscale, k.loffset, s + j * 3);
                        }
                        encode_run(e, s, cw, ia + c0, lo + r);
                    }
                }
            }
        }
        free(rb.buf);
        free(rgbrow);
        free(band);
        free(mband);
        free(chunk_out);
        free(wk.pool);
    }
    free(e);
    free_kernel(&k);
    return failed ? -1 : 0;
}

/* ------------------------------------------------------------------------------------------ */
/* CFA-plane DCT block denoiser (build-defined; the DCT-II matrix follows                     */
/* crates/flashraw/src/denoise/mod.rs:626-635, block positions follow positions() 877-884)     */

static size_t dct_positions(size_t n, size_t step, size_t *out) {
    size_t cnt = 0, last = n - 8;
    for (size_t p = 0; p <= last; p += step) out[cnt++] = p;
    if (out[cnt - 1] != last) out[cnt++] = last;
    return cnt;
}

int fro_dct_denoise(const uint16_t *mosaic, uint32_t width, uint32_t height, const float black[4],
                    float sig_a, float sig_b, float thr, uint32_t step, int precision, float *out_f32,
                    uint16_t *out_u16, int threads) {
    if (!mosaic || width % 2 || height % 2 || width < 16 || height < 16 || step == 0 || step > 8) return -1;
    size_t pw = width / 2, ph = height / 2;
    double D[8][8];
    for (int k = 0; k < 8; k++) {
        double a = k == 0 ? sqrt(1.0 / 8.0) : sqrt(2.0 / 8.0);
        for (int i = 0; i < 8; i++) D[k][i] = a * cos(M_PI * (2 * i + 1) * k / 16.0);
    }
    float Df[8][8];
    for (int k = 0; k < 8; k++)
        for (int i = 0; i < 8; i++) Df[k][i] = (float)D[k][i];
    size_t *xs = (size_t *)malloc((pw / step + 2) * sizeof(size_t));
    size_t *ys = (size_t *)malloc((ph / step + 2) * sizeof(size_t));
    double *acc = (double *)calloc(pw * ph, sizeof(double));
    float *cnt = (float *)calloc(pw * ph, sizeof(float));
    if (!xs || !ys || !acc || !cnt) {
        free(xs); free(ys); free(acc); free(cnt);
        return -1;
    }
    size_t nx = dct_positions(pw, step, xs), ny = dct_positions(ph, step, ys);
    int nthreads = threads > 0 ? threads : omp_get_max_threads();
    for (int plane = 0; plane < 4; plane++) {
        size_t dy = plane >> 1, dx = plane & 1;
        float bl = black[plane];
        memset(acc, 0, pw * ph * sizeof(double));
        memset(cnt, 0, pw * ph * sizeof(float));
        /* Block rows on the regular grid (multiples of `step`) are processed in `groups`
         * interleaved sets whose members lie more than 8 rows apart, so blocks running
         * concurrently never overlap and every sum has a fixed order. The last position, which
         * may follow its predecessor by less than `step`, runs in a pass of its own. */
        size_t groups = 8 / step + 1;
        size_t ngrid = (ph - 8) / step + 1;
        for (size_t par = 0; par <= groups; par++) {
            size_t first = par < groups ? par : ngrid, end = par < groups ? ngrid : ny;
            size_t stride = par < groups ? groups : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
            for (size_t by = first; by < end; by += stride) {
                size_t y0 = ys[by];
                for (size_t bx = 0; bx < nx; bx++) {
                    size_t x0 = xs[bx];
                    if (precision == 64) {
                        double X[8][8], T[8][8], C[8][8];
                        for (int i = 0; i < 8; i++)
                            for (int j = 0; j < 8; j++)
                                X[i][j] = (double)mosaic[(2 * (y0 + i) + dy) * width + 2 * (x0 + j) + dx] - (double)bl;
                        for (int k = 0; k < 8; k++)
                            for (int j = 0; j < 8; j++) {
                                double s = 0.0;
                                for (int i = 0; i < 8; i++) s += D[k][i] * X[i][j];
                                T[k][j] = s;
                            }
                        for (int k = 0; k < 8; k++)
                            for (int l = 0; l < 8; l++) {
                                double s = 0.0;
                                for (int j = 0; j < 8; j++) s += T[k][j] * D[l][j];
                                C[k][l] = s;
                            }
                        double mean = C[0][0] / 8.0;
                        double sig = sqrt((double)sig_a * (mean > 0.0 ? mean : 0.0) + (double)sig_b);
                        double t = (double)thr * sig;
                        for (int k = 0; k < 8; k++)
                            for (int l = 0; l < 8; l++)
                                if ((k | l) && !(fabs(C[k][l]) >= t)) C[k][l] = 0.0;
                        for (int i = 0; i < 8; i++)
                            for (int l = 0; l < 8; l++) {
                                double s = 0.0;
                                for (int k = 0; k < 8; k++) s += D[k][i] * C[k][l];
                                T[i][l] = s;
                            }
                        for (int i = 0; i < 8; i++)
                            for (int j = 0; j < 8; j++) {
                                double s = 0.0;
                                for (int l = 0; l < 8; l++) s += T[i][l] * D[l][j];
                                acc[(y0 + i) * pw + x0 + j] += s;
                                cnt[(y0 + i) * pw + x0 + j] += 1.0f;
                            }
                    } else {
                        float X[8][8], T[8][8], C[8][8];
                        for (int i = 0; i < 8; i++)
                            for (int j = 0; j < 8; j++)
                                X[i][j] = (float)mosaic[(2 * (y0 + i) + dy) * width + 2 * (x0 + j) + dx] - bl;
                        for (int k = 0; k < 8; k++)
                            for (int j = 0; j < 8; j++) {
                                float s = 0.0f;
                                for (int i = 0; i < 8; i++) s += Df[k][i] * X[i][j];
                                T[k][j] = s;
                            }
                        for (int k = 0; k < 8; k++)
                            for (int l = 0; l < 8; l++) {
                                float s = 0.0f;
                                for (int j = 0; j < 8; j++) s += T[k][j] * Df[l][j];
                                C[k][l] = s;
                            }
                        float mean = C[0][0] / 8.0f;
                        float sig = sqrtf(sig_a * (mean > 0.0f ? mean : 0.0f) + sig_b);
                        float t = thr * sig;
                        for (int k = 0; k < 8; k++)
                            for (int l = 0; l < 8; l++)
                                if ((k | l) && !(fabsf(C[k][l]) >= t)) C[k][l] = 0.0f;
                        for (int i = 0; i < 8; i++)
                            for (int l = 0; l < 8; l++) {
                                float s = 0.0f;
                                for (int k = 0; k < 8; k++) s += Df[k][i] * C[k][l];
                                T[i][l] = s;
                            }
                        for (int i = 0; i < 8; i++)
                            for (int j = 0; j < 8; j++) {
                                float s = 0.0f;
                                for (int l = 0; l < 8; l++) s += T[i][l] * Df[l][j];
                                acc[(y0 + i) * pw + x0 + j] += (double)s;
                                cnt[(y0 + i) * pw + x0 + j] += 1.0f;
                            }
                    }
                }
            }
        }
#pragma omp parallel for num_threads(nthreads)
        for (size_t y = 0; y < ph; y++) {
            for (size_t x = 0; x < pw; x++) {
                double v = acc[y * pw + x] / (double)cnt[y * pw + x];
                size_t o = (2 * y + dy) * width + 2 * x + dx;
                if (out_f32) out_f32[o] = (float)v;
                if (out_u16) {
                    double u = v + (double)bl + 0.5;
                    out_u16[o] = u <= 0.0 ? 0 : (u >= 65535.0 ? 65535 : (uint16_t)u);
                }
            }
        }
    }
    free(xs);
    free(ys);
    free(acc);
    free(cnt);
    return 0;
}

/* ---- the edge-aware smoothing weight (crates/flashraw/src/edgeaware/mod.rs:369-419) ---- */

int fro_normalize(const fro_frame *f, float *out) {
    fro_params p;
    memset(&p, 0, sizeof(p));
    p.demosaic = FRO_BILINEAR;
    Kern k;
    if (build_kernel(&k, f, &p)) {
        free_kernel(&k);
        return -1;
    }
    for (size_t y = 0; y < k.h; y++) normalize_row(&k, y, out + y * k.w * k.cpp);
    free_kernel(&k);
    return 0;
}

static int cmp_f32(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* Plan::set_noise_from (mod.rs:369-409) over a normalised frame: 48 rows y = 4 + n (h - 9) / 47;
 * per green sample x < w - 2 of a row and each green partner x + d (d = 1, 2) of the same row,
 * with a = row[x] g_gain, b = row[x + d] g_gain and mid = 0.5 (a + b) in (0.03, 0.6), the value
 * |a - b| / sqrt(mid); fewer than 1000 values give 0. Otherwise the value of rank len / 5 (in the
 * total order; every value here is finite and non-negative) is q, sigma = q / (sqrt(2) 0.2533),
 * nu = sigma^2 and beta = 0.5 clamp((nu - 2e-4) / 6e-4, 0, 1). */
float fro_edge_beta(const float *norm, uint32_t w, uint32_t h, const uint8_t *canon, uint32_t period, uint32_t cx,
                    uint32_t cy, float g_gain) {
    if (h < 16 || w < 16) return 0.0f;
    const size_t rows = 48, P = period;
    float *v = (float *)malloc(rows * (size_t)w * 2 * sizeof(float));
    if (!v) return -1.0f;
    size_t n = 0;
    for (size_t r = 0; r < rows; r++) {
        size_t y = 4 + r * (h - 9) / (rows - 1);
        const float *row = norm + y * (size_t)w;
        for (size_t x = 0; x + 2 < w; x++) {
            if (canon[((y + cy) % P) * P + (x + cx) % P] != 1) continue;
            for (size_t d = 1; d <= 2; d++) {
                if (canon[((y + cy) % P) * P + (x + d + cx) % P] != 1) continue;
                float a = row[x] * g_gain, b = row[x + d] * g_gain;
                float mid = 0.5f * (a + b);
                if (mid > 0.03f && mid < 0.6f) v[n++] = fabsf(a - b) / sqrtf(mid);
            }
        }
    }
    if (n < 1000) {
        free(v);
        return 0.0f;
    }
    qsort(v, n, sizeof(float), cmp_f32);
    float q = v[n / 5];
    free(v);
    const float k = 1.41421356f * 0.2533f; /* std::f32::consts::SQRT_2 * 0.2533, in f32 */
    float sigma = q / k;
    float nu = sigma * sigma;
    float t = (nu - 2e-4f) / 6e-4f;
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return 0.5f * t;
}

int fro_normalize_clamped(const fro_frame *f, float *out) {
    if (fro_normalize(f, out)) return -1;
    const size_t n = (size_t)f->width * f->height * (f->cpp ? f->cpp : 1);
    for (size_t i = 0; i < n; i++) out[i] = rmax(out[i], 0.0f);
    return 0;
}

/* ---- packed payloads (crates/flashraw-codecs/src/packed.rs) ---- */

/* stream_bits (packed.rs): the 64-bit little-endian load reordered so that the payload's bit
 * stream reads from the most significant bit (MSB orders) or the least (LSB bytes). */
static uint64_t stream_bits(uint64_t le, int order) {
    switch (order) {
    case FRO_LSB_BYTES: return le;
    case FRO_MSB_BYTES: return __builtin_bswap64(le);
    case FRO_MSB32: return (le << 32) | (le >> 32);
    default: {
        uint64_t halves = (le << 32) | (le >> 32);
        return ((halves & 0x0000FFFF0000FFFFull) << 16) | ((halves >> 16) & 0x0000FFFF0000FFFFull);
    }
    }
}

/* field (packed.rs): `bits` bits at bit offset `off` of the window. */
static uint16_t field(uint64_t window, unsigned off, unsigned bits, int order) {
    if (order == FRO_LSB_BYTES) return (uint16_t)((window >> off) & ((1ull << bits) - 1));
    return (uint16_t)((window << off) >> (64 - bits));
}

int fro_unpack(const uint8_t *src, uint64_t len, uint32_t w, uint32_t h, uint32_t bits, int order, uint64_t stride,
               uint16_t *out) {
    if (bits < 1 || bits > 16 || order < 0 || order > 3) return -1;
    const uint64_t row_bytes = ((uint64_t)w * bits + 7) / 8;
    const size_t word = order == FRO_MSB16 ? 2 : (order == FRO_MSB32 ? 4 : 1);
    for (uint64_t y = 0; y < h; y++) {
        /* the row slice: src[y * stride ..][..min(row_bytes, rest)] */
        uint64_t start = y * stride;
        uint64_t n = start < len ? len - start : 0;
        if (n > row_bytes) n = row_bytes;
        const uint8_t *row = src + (start < len ? start : 0);
        uint64_t pos = 0;
        for (uint32_t x = 0; x < w; x++, pos += bits) {
            /* samples_one_by_one: the window at the storage word holding the first bit,
             * load_le64_padded (zeros past the slice) */
            uint64_t at = (pos / 8) & ~(uint64_t)(word - 1);
            uint64_t le = 0;
            for (unsigned j = 0; j < 8; j++)
                if (at + j < n) le |= (uint64_t)row[at + j] << (8 * j);
            out[y * w + x] = field(stream_bits(le, order), (unsigned)(pos - at * 8), bits, order);
        }
    }
    return 0;
}

void fro_linearize(uint16_t *v, uint64_t n, const uint16_t *lut, uint32_t lut_len) {
    if (!lut || lut_len == 0) return;
    for (uint64_t i = 0; i < n; i++) v[i] = lut[v[i] < lut_len ? v[i] : lut_len - 1];
}
