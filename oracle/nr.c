/*
 * CPU oracle for the luma / chroma noise reduction of developed linear RGB.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). A restatement, operation for operation, of the
 * reference's crates/flashraw/src/denoise/mod.rs: apply 75-130, the strength helpers 132-153,
 * blend 223-229, Variance::at 261-275, box_mean / box_row 280-335, Guide 337-382, level_count /
 * pyramid 384-404, split / merge 433-464, down / tap / up_row / add_up / up_affine 467-555,
 * denoise_plane 567-625, Dct (fwd1 / inv1 / block) 627-703, Model 705-869 (next_level, cap,
 * pixel_variance, sigmas, estimate, estimate_step, from_stats), pad_map / positions 871-898,
 * dct_pass and run_group_impl 900-1131. Row-chunked loops keep the reference's chunk heights
 * (32 rows for box_mean), so the result does not depend on the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NB 8
#define NN 64
#define HARD 2.7f
#define CHI2_MEDIAN 0.4549364f
#define SELECTION_BIAS 1.10f
#define MIN_LEVEL_SIDE 48
#define MAX_BLOCKS 50000
#define LUMA_KEEP_EXP 1.2f
#define LUMA_LIGHT 0.25f
#define CHROMA_RADIUS 6.0f
#define LEVEL_CAP 1.5f
#define PAD 8
#define STEP 2
#define GROUP 16
#define MAXBINS 12

static const float KR = 0.2126f, KG = 0.7152f, KB = 0.0722f;

typedef struct {
    size_t w, h;
    float *v;
} Plane;

static Plane plane_new(size_t w, size_t h) {
    const size_t n = w * h;
    Plane p = {w, h, (float *)calloc(n ? n : 1, sizeof(float))};
    return p;
}
static Plane plane_copy(const Plane *s) {
    Plane p = plane_new(s->w, s->h);
    memcpy(p.v, s->v, s->w * s->h * sizeof(float));
    return p;
}
static void plane_free(Plane *p) {
    free(p->v);
    p->v = NULL;
}

static inline float rmaxf(float a, float b) { return (a != a) ? b : (b != b) ? a : (a > b ? a : b); }
static inline float rminf(float a, float b) { return (a != a) ? b : (b != b) ? a : (a < b ? a : b); }
static inline float rclampf(float v, float lo, float hi) { return v != v ? v : (v < lo ? lo : (v > hi ? hi : v)); }
static inline size_t zmin(size_t a, size_t b) { return a < b ? a : b; }

/* ---- Dct (mod.rs:627-703) ---- */

static float M[NB][NB];

static void dct_init(void) {
    static int done = 0;
    if (done) return;
    for (int k = 0; k < NB; k++) {
        double a = k == 0 ? sqrt(1.0 / NB) : sqrt(2.0 / NB);
        for (int i = 0; i < NB; i++) M[k][i] = (float)(a * cos(M_PI * (double)(2 * i + 1) * (double)k / (double)(2 * NB)));
    }
    done = 1;
}

/* (c0 x0 + c1 x1) + (c2 x2 + c3 x3) */
static inline float dot4(const float c[4], float x0, float x1, float x2, float x3) {
    return (c[0] * x0 + c[1] * x1) + (c[2] * x2 + c[3] * x3);
}

/* out[k] = sum_i m[k][i] x[i] with the even / odd factorisation of fwd1 */
static void fwd1(const float x[NB], float out[NB]) {
    float s[4], d[4];
    for (int i = 0; i < 4; i++) {
        s[i] = x[i] + x[7 - i];
        d[i] = x[i] - x[7 - i];
    }
    float e0 = s[0] + s[3], e1 = s[1] + s[2], e2 = s[0] - s[3], e3 = s[1] - s[2];
    out[0] = M[0][0] * (e0 + e1);
    out[1] = dot4(M[1], d[0], d[1], d[2], d[3]);
    out[2] = M[2][0] * e2 + M[2][1] * e3;
    out[3] = dot4(M[3], d[0], d[1], d[2], d[3]);
    out[4] = M[4][0] * (e0 - e1);
    out[5] = dot4(M[5], d[0], d[1], d[2], d[3]);
    out[6] = M[6][0] * e2 + M[6][1] * e3;
    out[7] = dot4(M[7], d[0], d[1], d[2], d[3]);
}

/* out[i] = sum_k m[k][i] x[k] (inv1) */
static void inv1(const float x[NB], float out[NB]) {
    float p = M[0][0] * (x[0] + x[4]);
    float q = M[0][0] * (x[0] - x[4]);
    float u = M[2][0] * x[2] + M[6][0] * x[6];
    float v = M[2][1] * x[2] + M[6][1] * x[6];
    float e[4] = {p + u, q + v, q - v, p - u};
    float o[4];
    for (int i = 0; i < 4; i++) {
        float c[4] = {M[1][i], M[3][i], M[5][i], M[7][i]};
        o[i] = dot4(c, x[1], x[3], x[5], x[7]);
    }
    for (int i = 0; i < 8; i++) out[i] = i < 4 ? e[i] + o[i] : e[7 - i] - o[7 - i];
}

/* Dct::block: out[l * 8 + k] = horizontal frequency l, vertical frequency k of the block at
 * (x0, y0) of a plane of width w. */
static void dct_block(const float *pv, size_t w, size_t x0, size_t y0, float out[NN]) {
    float v[NB][NB]; /* v[k][j]: vertical frequency k of column j */
    for (int j = 0; j < NB; j++) {
        float col[NB], o[NB];
        for (int i = 0; i < NB; i++) col[i] = pv[(y0 + i) * w + x0 + j];
        fwd1(col, o);
        for (int k = 0; k < NB; k++) v[k][j] = o[k];
    }
    for (int k = 0; k < NB; k++) {
        float o[NB];
        fwd1(v[k], o);
        for (int l = 0; l < NB; l++) out[l * NB + k] = o[l];
    }
}

/* ---- split / merge / blend (mod.rs:223-229, 433-464) ---- */

static void split(const float *rgb, size_t w, size_t h, Plane *y, Plane *u, Plane *v) {
    *y = plane_new(w, h);
    *u = plane_new(w, h);
    *v = plane_new(w, h);
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < w * h; k++) {
        const float *p = rgb + 3 * k;
        float l = KR * p[0] + KG * p[1] + KB * p[2];
        y->v[k] = l;
        u->v[k] = p[2] - l;
        v->v[k] = p[0] - l;
    }
}

static void merge(float *rgb, const Plane *y, const Plane *u, const Plane *v) {
    const size_t n = y->w * y->h;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; k++) {
        float r = v->v[k] + y->v[k], b = u->v[k] + y->v[k];
        rgb[3 * k] = r;
        rgb[3 * k + 1] = (y->v[k] - KR * r - KB * b) / KG;
        rgb[3 * k + 2] = b;
    }
}

static void blend(Plane *d, const Plane *orig, float keep) {
    if (!(keep > 0.0f)) return;
    const size_t n = d->w * d->h;
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; k++) d->v[k] += keep * (orig->v[k] - d->v[k]);
}

/* ---- down / tap / up (mod.rs:467-555) ---- */

static Plane down(const Plane *p) {
    const size_t w = (p->w + 1) / 2, h = (p->h + 1) / 2;
    Plane o = plane_new(w, h);
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++) {
        size_t y0 = 2 * y, y1 = zmin(2 * y + 1, p->h - 1);
        const float *r0 = p->v + y0 * p->w, *r1 = p->v + y1 * p->w;
        size_t pairs = p->w / 2;
        for (size_t x = 0; x < pairs; x++)
            o.v[y * w + x] = 0.25f * (r0[2 * x] + r0[2 * x + 1] + r1[2 * x] + r1[2 * x + 1]);
        if (pairs < w) {
            size_t x = p->w - 1;
            o.v[y * w + w - 1] = 0.25f * (r0[x] + r0[x] + r1[x] + r1[x]);
        }
    }
    return o;
}

typedef struct {
    size_t a, b;
    float f;
} Tap;

static Tap tap(size_t i, size_t n) {
    float s = rmaxf(((float)i + 0.5f) * 0.5f - 0.5f, 0.0f);
    Tap t;
    t.a = zmin((size_t)s, n - 1);
    t.b = zmin(t.a + 1, n - 1);
    t.f = s - (float)t.a;
    return t;
}

static void up_row(const Plane *p, const Tap *xs, size_t y, size_t w, float *out) {
    Tap ty = tap(y, p->h);
    const float *r0 = p->v + ty.a * p->w, *r1 = p->v + ty.b * p->w;
    for (size_t x = 0; x < w; x++) {
        const Tap tx = xs[x];
        float a = r0[tx.a] + (r0[tx.b] - r0[tx.a]) * tx.f;
        float b = r1[tx.a] + (r1[tx.b] - r1[tx.a]) * tx.f;
        out[x] = a + (b - a) * ty.f;
    }
}

static void add_up(Plane *d, const Plane *p) {
    const size_t w = d->w, h = d->h;
    Tap *xs = (Tap *)malloc(w * sizeof(Tap));
    for (size_t x = 0; x < w; x++) xs[x] = tap(x, p->w);
#pragma omp parallel
    {
        float *u = (float *)malloc(w * sizeof(float));
#pragma omp for schedule(static)
        for (size_t y = 0; y < h; y++) {
            up_row(p, xs, y, w, u);
            for (size_t x = 0; x < w; x++) d->v[y * w + x] += u[x];
        }
        free(u);
    }
    free(xs);
}

static Plane up_affine(const Plane *a, const Plane *b, const Plane *g) {
    const size_t w = g->w, h = g->h;
    Tap *xs = (Tap *)malloc(w * sizeof(Tap));
    for (size_t x = 0; x < w; x++) xs[x] = tap(x, a->w);
    Plane o = plane_new(w, h);
#pragma omp parallel
    {
        float *ua = (float *)malloc(w * sizeof(float));
#pragma omp for schedule(static)
        for (size_t y = 0; y < h; y++) {
            up_row(a, xs, y, w, ua);
            up_row(b, xs, y, w, o.v + y * w);
            for (size_t x = 0; x < w; x++) o.v[y * w + x] += ua[x] * g->v[y * w + x];
        }
        free(ua);
    }
    free(xs);
    return o;
}

/* ---- box_mean (mod.rs:280-335) ---- */

/* f(k) as a plane of the same size: box_mean's input values */
static void box_row(const float *a, size_t w, size_t r, float n, float *out) {
#define AT(x) a[(x) < 0 ? 0 : ((x) > (long)w - 1 ? w - 1 : (size_t)(x))]
    const size_t seg = (w + 3) / 4;
    size_t start[4], end[4];
    for (int j = 0; j < 4; j++) start[j] = zmin(j * seg, w);
    for (int j = 0; j < 4; j++) end[j] = j < 3 ? start[j + 1] : w;
    float s[4];
    for (int j = 0; j < 4; j++) {
        float acc = -0.0f;
        for (long d = -(long)r; d <= (long)r; d++) acc += AT((long)start[j] + d);
        s[j] = acc;
    }
    for (size_t i = 0; i < seg; i++)
        for (int j = 0; j < 4; j++) {
            size_t x = start[j] + i;
            if (x < end[j]) {
                out[x] = s[j] / n;
                s[j] += AT((long)x + (long)r + 1) - AT((long)x - (long)r);
            }
        }
#undef AT
}

/* Direct window sums in the order of the reference's GPU kernels (dn_box_v / dn_box_h,
 * crates/flashraw-gpu/shaders/denoise.wgsl:191-218): per sample, the vertical sum over
 * d = -r..r from 0, then the horizontal sum of those over d = -r..r from 0, divided by n. The
 * reference's CPU path (box_mean below) slides running sums instead, whose rounding drifts along
 * its segments; the two orders give different bits. */
static Plane box_mean_direct(const Plane *f, size_t r) {
    const size_t w = f->w, h = f->h;
    const float n = (float)((2 * r + 1) * (2 * r + 1));
    Plane vs = plane_new(w, h), out = plane_new(w, h);
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            float s = 0.0f;
            for (long d = -(long)r; d <= (long)r; d++) {
                long yy = (long)y + d;
                yy = yy < 0 ? 0 : (yy > (long)h - 1 ? (long)h - 1 : yy);
                s += f->v[(size_t)yy * w + x];
            }
            vs.v[y * w + x] = s;
        }
#pragma omp parallel for schedule(static)
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) {
            float s = 0.0f;
            for (long d = -(long)r; d <= (long)r; d++) {
                long xx = (long)x + d;
                xx = xx < 0 ? 0 : (xx > (long)w - 1 ? (long)w - 1 : xx);
                s += vs.v[y * w + (size_t)xx];
            }
            out.v[y * w + x] = s / n;
        }
    plane_free(&vs);
    return out;
}

static Plane box_mean(const Plane *f, size_t r, int direct) {
    if (direct) return box_mean_direct(f, r);
    const size_t w = f->w, h = f->h;
    const float n = (float)((2 * r + 1) * (2 * r + 1));
    Plane out = plane_new(w, h);
    const size_t chunks = (h + 31) / 32;
#pragma omp parallel
    {
        float *acc = (float *)malloc(w * sizeof(float)), *add = (float *)malloc(w * sizeof(float)),
              *sub = (float *)malloc(w * sizeof(float));
#pragma omp for schedule(dynamic, 1)
        for (size_t c = 0; c < chunks; c++) {
            const size_t first = c * 32, rows = zmin(32, h - first);
#define ROW(yy, o)                                                                \
    do {                                                                          \
        long y_ = (yy);                                                           \
        size_t k0 = (size_t)(y_ < 0 ? 0 : (y_ > (long)h - 1 ? (long)h - 1 : y_)) * w; \
        for (size_t x = 0; x < w; x++) (o)[x] = f->v[k0 + x];                     \
    } while (0)
            for (size_t x = 0; x < w; x++) acc[x] = 0.0f;
            for (long dy = -(long)r; dy <= (long)r; dy++) {
                ROW((long)first + dy, add);
                for (size_t x = 0; x < w; x++) acc[x] += add[x];
            }
            for (size_t i = 0; i < rows; i++) {
                long y = (long)(first + i);
                box_row(acc, w, r, n, out.v + (first + i) * w);
                ROW(y + (long)r + 1, add);
                ROW(y - (long)r, sub);
                for (size_t x = 0; x < w; x++) acc[x] += add[x] - sub[x];
            }
#undef ROW
        }
        free(acc);
        free(add);
        free(sub);
    }
    return out;
}

/* ---- the noise model (mod.rs:705-869) ---- */

typedef struct {
    size_t n;
    float m[MAXBINS];
    float s[MAXBINS][NN];
} Model;

typedef struct {
    size_t n;
    float m[MAXBINS];
    float v[MAXBINS];
} Variance;

/* partition_point(|b| b < m): the first index whose mean is not below m */
static size_t partition_point(const float *a, size_t n, float m) {
    size_t i = 0;
    while (i < n && a[i] < m) i++;
    return i;
}

static float variance_at(const Variance *vr, float m) {
    size_t i = partition_point(vr->m, vr->n, m);
    if (i == 0) return vr->v[0];
    if (i >= vr->n) return vr->v[vr->n - 1];
    float a = vr->m[i - 1], b = vr->m[i];
    float t = b > a ? (m - a) / (b - a) : 0.0f;
    return vr->v[i - 1] + (vr->v[i] - vr->v[i - 1]) * t;
}

static Variance next_level(const Model *md) {
    Variance v;
    v.n = md->n;
    for (size_t b = 0; b < md->n; b++) {
        float t = 0.0f;
        for (int l = 0; l < 4; l++)
            for (int k = 0; k < 4; k++) t += md->s[b][l * NB + k] * md->s[b][l * NB + k];
        v.m[b] = md->m[b];
        v.v[b] = 0.25f * t / 15.0f;
    }
    return v;
}

static void model_cap(Model *md, const Variance *pred, float f) {
    for (size_t b = 0; b < md->n; b++) {
        float cap = f * sqrtf(rmaxf(variance_at(pred, md->m[b]), 0.0f));
        for (int k = 0; k < NN; k++) md->s[b][k] = rminf(md->s[b][k], cap);
    }
}

static Variance pixel_variance(const Model *md) {
    Variance v;
    v.n = md->n;
    for (size_t b = 0; b < md->n; b++) {
        float t = -0.0f;
        for (int k = 1; k < NN; k++) t += md->s[b][k] * md->s[b][k];
        v.m[b] = md->m[b];
        v.v[b] = t / (float)(NN - 1);
    }
    return v;
}

/* sigmas(m): rows a, b of the deviations and the weight t of b */
static void model_sigmas(const Model *md, float m, const float **sa, const float **sb, float *t) {
    size_t n = md->n, i = partition_point(md->m, n, m);
    if (i == 0 || n == 1) {
        *sa = *sb = md->s[0];
        *t = 0.0f;
        return;
    }
    if (i >= n) {
        *sa = *sb = md->s[n - 1];
        *t = 0.0f;
        return;
    }
    float a = md->m[i - 1], b = md->m[i];
    *sa = md->s[i - 1];
    *sb = md->s[i];
    *t = b > a ? (m - a) / (b - a) : 0.0f;
}

/* total order of f32 (f32::total_cmp) */
static inline int32_t total_key(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}
static int cmp_total(const void *a, const void *b) {
    int32_t x = total_key(*(const float *)a), y = total_key(*(const float *)b);
    return (x > y) - (x < y);
}

typedef struct {
    float mean, act;
    uint32_t idx;
} Ord;

static int cmp_ord(const void *a, const void *b) {
    const Ord *x = (const Ord *)a, *y = (const Ord *)b;
    int32_t kx = total_key(x->mean), ky = total_key(y->mean);
    if (kx != ky) return (kx > ky) - (kx < ky);
    return (x->idx > y->idx) - (x->idx < y->idx); /* stable */
}

/* the value of rank k of v[0..n) in the total order (select_nth_unstable_by) */
static float nth(float *v, size_t n, size_t k) {
    qsort(v, n, sizeof(float), cmp_total);
    return v[k];
}

/* Model::from_stats: 0 on success, -1 when fewer than 32 finite blocks */
int fro_nr_model_from_stats(const float *stats, uint32_t count, uint32_t *bins_out, float *means, float *sigmas) {
    Ord *order = (Ord *)malloc((count ? count : 1) * sizeof(Ord));
    size_t n = 0;
    for (uint32_t i = 0; i < count; i++) {
        const float *s = stats + (size_t)i * NN;
        float a = 0.0f;
        for (int k = 1; k < NN; k++) a = a + s[k];
        if (isfinite(s[0]) && isfinite(a)) {
            order[n].mean = s[0];
            order[n].act = a;
            order[n].idx = i;
            n++;
        }
    }
    if (n < 32) {
        free(order);
        return -1;
    }
    qsort(order, n, sizeof(Ord), cmp_ord);
    size_t bins = n / 96;
    bins = bins < 1 ? 1 : (bins > MAXBINS ? MAXBINS : bins);
    size_t per = (n + bins - 1) / bins;
    float *act = (float *)malloc(per * sizeof(float)), *vals = (float *)malloc(per * sizeof(float));
    const float **flat = (const float **)malloc(per * sizeof(float *));
    Model md;
    md.n = 0;
    for (size_t b0 = 0; b0 < n; b0 += per) {
        size_t len = zmin(per, n - b0);
        for (size_t i = 0; i < len; i++) act[i] = order[b0 + i].act;
        float cut = nth(act, len, len / 2);
        size_t nf = 0;
        for (size_t i = 0; i < len; i++)
            if (order[b0 + i].act <= cut) flat[nf++] = stats + (size_t)order[b0 + i].idx * NN;
        float *sig = md.s[md.n];
        sig[0] = 0.0f;
        for (int k = 1; k < NN; k++) {
            for (size_t i = 0; i < nf; i++) vals[i] = flat[i][k];
            float med = nth(vals, nf, nf / 2);
            sig[k] = sqrtf(med / CHI2_MEDIAN * SELECTION_BIAS);
        }
        md.m[md.n] = order[b0 + len / 2].mean;
        md.n++;
    }
    if (md.n > 2) {
        Model orig = md;
        for (size_t i = 0; i < md.n; i++) {
            size_t a = i ? i - 1 : 0, b = zmin(i + 1, md.n - 1);
            for (int k = 0; k < NN; k++) md.s[i][k] = 0.25f * orig.s[a][k] + 0.5f * orig.s[i][k] + 0.25f * orig.s[b][k];
        }
    }
    *bins_out = (uint32_t)md.n;
    memcpy(means, md.m, md.n * sizeof(float));
    memcpy(sigmas, md.s, md.n * NN * sizeof(float));
    free(order);
    free(act);
    free(vals);
    free(flat);
    return 0;
}

/* estimate_step (mod.rs:790-798): 0 when too small to measure */
static size_t estimate_step(size_t w, size_t h) {
    size_t bx = w / NB, by = h / NB;
    if (bx * by < 32) return 0;
    size_t s = (size_t)ceil(sqrt((double)(bx * by) / (double)MAX_BLOCKS));
    return s < 1 ? 1 : s;
}

uint32_t fro_nr_estimate_step(uint32_t w, uint32_t h) { return (uint32_t)estimate_step(w, h); }

/* The measured blocks of p binned by guide (estimate's statistics, mod.rs:761-788): per block
 * the guide mean then the squared coefficients; returns the count (0: too small). */
uint32_t fro_nr_stats(const float *p, const float *guide, uint32_t w, uint32_t h, float *out) {
    dct_init();
    size_t step = estimate_step(w, h);
    if (!step) return 0;
    size_t bx = w / NB, by = h / NB, cols = (bx + step - 1) / step, rows = (by + step - 1) / step;
    if (!out) return (uint32_t)(cols * rows);
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < rows; r++)
        for (size_t c = 0; c < cols; c++) {
            size_t x0 = c * step * NB, y0 = r * step * NB;
            float mean = 0.0f;
            for (int i = 0; i < NB; i++)
                for (int j = 0; j < NB; j++) mean += guide[(y0 + i) * w + x0 + j];
            float coef[NN];
            dct_block(p, w, x0, y0, coef);
            float *st = out + (r * cols + c) * NN;
            for (int k = 0; k < NN; k++) st[k] = coef[k] * coef[k];
            st[0] = mean / (float)NN;
        }
    return (uint32_t)(cols * rows);
}

static int estimate(const Plane *p, const Plane *guide, Model *out) {
    uint32_t n = fro_nr_stats(p->v, guide->v, (uint32_t)p->w, (uint32_t)p->h, NULL);
    if (!n) return 0;
    float *st = (float *)malloc((size_t)n * NN * sizeof(float));
    fro_nr_stats(p->v, guide->v, (uint32_t)p->w, (uint32_t)p->h, st);
    uint32_t bins = 0;
    int rc = fro_nr_model_from_stats(st, n, &bins, out->m, &out->s[0][0]);
    free(st);
    if (rc) return 0;
    out->n = bins;
    return 1;
}

/* ---- the DCT passes (mod.rs:871-1131) ---- */

static size_t *pad_map(size_t n, size_t pad, size_t extra, size_t *len) {
    *len = n + 2 * pad + extra;
    size_t *m = (size_t *)malloc(*len * sizeof(size_t));
    for (size_t i0 = 0; i0 < *len; i0++) {
        long i = (long)i0 - (long)pad, nn = (long)n;
        long r = i < 0 ? -i - 1 : (i >= nn ? 2 * nn - i - 1 : i);
        r = r < 0 ? 0 : (r > nn - 1 ? nn - 1 : r);
        m[i0] = (size_t)r;
    }
    return m;
}

static size_t *positions(size_t n, size_t step, size_t *len) {
    size_t last = n - NB;
    size_t cnt = last / step + 1;
    size_t *v = (size_t *)malloc((cnt + 1) * sizeof(size_t));
    for (size_t i = 0; i < cnt; i++) v[i] = i * step;
    if (v[cnt - 1] != last) v[cnt++] = last;
    *len = cnt;
    return v;
}

typedef struct {
    size_t first, rows;
    float *acc, *wsum;
} Part;

/* One DCT pass over src: hard thresholding with threshold `hard` deviations (pilot NULL), the
 * empirical Wiener filter of pilot otherwise (dct_pass / run_group_impl). */
static Plane dct_pass(const Plane *src, const Plane *guide, const Plane *pilot, const Model *md, float gain,
                      float hard, size_t step) {
    const size_t w = src->w, h = src->h;
    const size_t pw = w + 2 * PAD, extra = (pw + NB - 1) / NB * NB - pw;
    size_t nys, nxs, fw, nymap;
    size_t *ys = positions(h + 2 * PAD, step, &nys);
    size_t *xs = positions(pw, step, &nxs);
    size_t *xmap = pad_map(w, PAD, extra, &fw);
    size_t *ymap = pad_map(h, PAD, 0, &nymap);
    const size_t ngroups = (nys + GROUP - 1) / GROUP;
    Part *parts = (Part *)calloc(ngroups, sizeof(Part));
    const float thr = hard * gain, n2scale = gain * gain, tiny = 1e-30f;
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t gi = 0; gi < ngroups; gi++) {
        const size_t *g = ys + gi * GROUP;
        const size_t gl = zmin(GROUP, nys - gi * GROUP);
        const size_t first = g[0], rows = g[gl - 1] + NB - first;
        float *sp = (float *)malloc(rows * fw * sizeof(float)), *gp = (float *)malloc(rows * fw * sizeof(float));
        float *pp = pilot ? (float *)malloc(rows * fw * sizeof(float)) : NULL;
        for (size_t i = 0; i < rows; i++) {
            size_t sy = ymap[first + i];
            for (size_t x = 0; x < fw; x++) {
                sp[i * fw + x] = src->v[sy * w + xmap[x]];
                gp[i * fw + x] = guide->v[sy * w + xmap[x]];
                if (pp) pp[i * fw + x] = pilot->v[sy * w + xmap[x]];
            }
        }
        float *acc = (float *)calloc(rows * fw, sizeof(float)), *wsum = (float *)calloc(rows * fw, sizeof(float));
        float(*vert)[NB] = (float(*)[NB])malloc(fw * sizeof(*vert));
        float(*vpil)[NB] = pp ? (float(*)[NB])malloc(fw * sizeof(*vpil)) : NULL;
        float(*racc)[NB] = (float(*)[NB])malloc(fw * sizeof(*racc));
        float *rw = (float *)malloc(fw * sizeof(float)), *gcol = (float *)malloc(fw * sizeof(float));
        for (size_t gy = 0; gy < gl; gy++) {
            const size_t r0 = g[gy] - first;
            for (size_t x = 0; x < fw; x++) {
                float col[NB];
                for (int i = 0; i < NB; i++) col[i] = sp[(r0 + i) * fw + x];
                fwd1(col, vert[x]);
                if (pp) {
                    for (int i = 0; i < NB; i++) col[i] = pp[(r0 + i) * fw + x];
                    fwd1(col, vpil[x]);
                }
            }
            for (size_t x = 0; x < fw; x++) gcol[x] = 0.0f;
            for (int i = 0; i < NB; i++)
                for (size_t x = 0; x < fw; x++) gcol[x] += gp[(r0 + i) * fw + x];
            for (size_t x = 0; x < fw; x++) {
                rw[x] = 0.0f;
                for (int k = 0; k < NB; k++) racc[x][k] = 0.0f;
            }
            for (size_t xi = 0; xi < nxs; xi++) {
                const size_t x0 = xs[xi];
                float coef[NB][NB]; /* coef[l][k] */
                for (int k = 0; k < NB; k++) {
                    float row[NB], o[NB];
                    for (int j = 0; j < NB; j++) row[j] = vert[x0 + j][k];
                    fwd1(row, o);
                    for (int l = 0; l < NB; l++) coef[l][k] = o[l];
                }
                float gs = -0.0f;
                for (int j = 0; j < NB; j++) gs += gcol[x0 + j];
                const float mean = gs / (float)NN;
                const float *sa, *sb;
                float t;
                model_sigmas(md, mean, &sa, &sb, &t);
                float weight;
                if (pp) {
                    float pc[NB][NB];
                    for (int k = 0; k < NB; k++) {
                        float row[NB], o[NB];
                        for (int j = 0; j < NB; j++) row[j] = vpil[x0 + j][k];
                        fwd1(row, o);
                        for (int l = 0; l < NB; l++) pc[l][k] = o[l];
                    }
                    const float dc = coef[0][0];
                    float e[NB] = {0};
                    for (int l = 0; l < NB; l++)
                        for (int k = 0; k < NB; k++) {
                            float a = sa[l * NB + k], b = sb[l * NB + k];
                            float sg = a + (b - a) * t;
                            float n2 = sg * sg * n2scale;
                            float p2 = pc[l][k] * pc[l][k];
                            float s = p2 / (p2 + n2 + tiny);
                            coef[l][k] = coef[l][k] * s;
                            e[k] = e[k] + s * s * n2;
                        }
                    coef[0][0] = dc;
                    float es = ((e[0] + e[4]) + (e[1] + e[5])) + ((e[2] + e[6]) + (e[3] + e[7]));
                    weight = 1.0f / (es + 1e-20f);
                } else {
                    uint32_t kept = 0;
                    for (int l = 0; l < NB; l++)
                        for (int k = 0; k < NB; k++) {
                            float a = sa[l * NB + k], b = sb[l * NB + k];
                            float sg = a + (b - a) * t;
                            int keep = fabsf(coef[l][k]) >= sg * thr;
                            kept += (uint32_t)keep;
                            if (!keep) coef[l][k] = 0.0f;
                        }
                    weight = 1.0f / (float)kept;
                }
                for (int l = 0; l < NB; l++)
                    for (int k = 0; k < NB; k++) coef[l][k] = coef[l][k] * weight;
                /* horizontal inverse: column x0 + j, vertical frequency k */
                for (int k = 0; k < NB; k++) {
                    float in[NB], o[NB];
                    for (int l = 0; l < NB; l++) in[l] = coef[l][k];
                    inv1(in, o);
                    for (int j = 0; j < NB; j++) racc[x0 + j][k] = racc[x0 + j][k] + o[j];
                }
                for (int j = 0; j < NB; j++) rw[x0 + j] = rw[x0 + j] + weight;
            }
            /* vertical inverse of the summed blocks */
            for (size_t x = 0; x < fw; x++) {
                float px[NB];
                inv1(racc[x], px);
                for (int i = 0; i < NB; i++) {
                    acc[(r0 + i) * fw + x] = acc[(r0 + i) * fw + x] + px[i];
                    wsum[(r0 + i) * fw + x] = wsum[(r0 + i) * fw + x] + rw[x];
                }
            }
        }
        parts[gi].first = first;
        parts[gi].rows = rows;
        parts[gi].acc = acc;
        parts[gi].wsum = wsum;
        free(sp);
        free(gp);
        free(pp);
        free(vert);
        free(vpil);
        free(racc);
        free(rw);
        free(gcol);
    }
    Plane out = plane_new(w, h);
#pragma omp parallel
    {
        float *a = (float *)malloc(w * sizeof(float)), *ws = (float *)malloc(w * sizeof(float));
#pragma omp for schedule(static)
        for (size_t yy = 0; yy < h; yy++) {
            const size_t y = yy + PAD;
            for (size_t x = 0; x < w; x++) a[x] = ws[x] = 0.0f;
            size_t k = 0;
            while (k < ngroups && parts[k].first + parts[k].rows <= y) k++;
            for (; k < ngroups && parts[k].first <= y; k++) {
                const size_t o = (y - parts[k].first) * fw + PAD;
                for (size_t x = 0; x < w; x++) {
                    a[x] += parts[k].acc[o + x];
                    ws[x] += parts[k].wsum[o + x];
                }
            }
            for (size_t x = 0; x < w; x++) out.v[yy * w + x] = ws[x] > 0.0f ? a[x] / ws[x] : src->v[yy * w + x];
        }
        free(a);
        free(ws);
    }
    for (size_t gi = 0; gi < ngroups; gi++) {
        free(parts[gi].acc);
        free(parts[gi].wsum);
    }
    free(parts);
    free(ys);
    free(xs);
    free(xmap);
    free(ymap);
    return out;
}

/* ---- levels and the plane pipeline (mod.rs:384-404, 567-625) ---- */

static size_t level_count(size_t w, size_t h, size_t max) {
    size_t n = 1;
    while (n < max && (w + 1) / 2 >= MIN_LEVEL_SIDE && (h + 1) / 2 >= MIN_LEVEL_SIDE) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
        n++;
    }
    return n;
}

uint32_t fro_nr_level_count(uint32_t w, uint32_t h, uint32_t max) { return (uint32_t)level_count(w, h, max); }

typedef struct {
    float gain;
    const size_t *steps;
    size_t nsteps;
    int hard_only0;
} NrPlan;

/* levels[0..nl) the plane's pyramid, guides the luma pyramid (same sizes); first: the level-0
 * model when already measured (NULL: measure) */
static Plane denoise_plane(Plane *const *levels, Plane *const *guides, size_t nl, const Model *first,
                           const NrPlan *plan) {
    Model *models = (Model *)calloc(nl, sizeof(Model));
    int *have = (int *)calloc(nl, sizeof(int));
    for (size_t l = 0; l < nl; l++) {
        if (l == 0 && first) {
            models[0] = *first;
            have[0] = 1;
        } else {
            have[l] = estimate(levels[l], guides[l], &models[l]);
        }
        if (l > 0) {
            if (have[l] && have[l - 1]) {
                Variance pred = next_level(&models[l - 1]);
                model_cap(&models[l], &pred, LEVEL_CAP);
            } else {
                have[l] = 0;
            }
        }
    }
    Plane result = {0, 0, NULL};
    for (size_t li = nl; li-- > 0;) {
        const Plane *src = levels[li], *guide = guides[li];
        Plane d;
        if (have[li]) {
            size_t step = li < plan->nsteps ? plan->steps[li] : STEP;
            Plane pilot = dct_pass(src, guide, NULL, &models[li], plan->gain, HARD, step);
            if (li == 0 && plan->hard_only0) {
                d = pilot;
            } else {
                d = dct_pass(src, guide, &pilot, &models[li], plan->gain, HARD, step);
                plane_free(&pilot);
            }
        } else {
            d = plane_copy(src);
        }
        if (!result.v) {
            result = d;
        } else {
            Plane diff = down(&d);
            for (size_t k = 0; k < diff.w * diff.h; k++) diff.v[k] = result.v[k] - diff.v[k];
            add_up(&d, &diff);
            plane_free(&diff);
            plane_free(&result);
            result = d;
        }
    }
    free(models);
    free(have);
    return result;
}

static float luma_keep(float s) { return powf(1.0f - s, LUMA_KEEP_EXP); }
static float chroma_keep(float s) { return rmaxf(1.0f - 2.0f * s, 0.0f); }
static float chroma_gain(float s) { return 1.0f + 2.0f * rmaxf(s - 0.5f, 0.0f); }
static size_t chroma_radius(float s) {
    float r = rmaxf(roundf(CHROMA_RADIUS * (0.5f + s)), 1.0f);
    return (size_t)r;
}

/* denoise::apply (mod.rs:75-130) on interleaved RGB of w x h, in place. */
int fro_nr_apply(float *rgb, uint32_t wi, uint32_t hi, float luma, float chroma) {
    return fro_nr_apply_order(rgb, wi, hi, luma, chroma, 0);
}

int fro_nr_apply_order(float *rgb, uint32_t wi, uint32_t hi, float luma, float chroma, int box_direct) {
    dct_init();
    luma = rclampf(luma, 0.0f, 1.0f);
    chroma = rclampf(chroma, 0.0f, 1.0f);
    if (luma <= 0.0f && chroma <= 0.0f) return 0;
    const size_t w = wi, h = hi;
    if (w < NB || h < NB) return 0;
    Plane y, u, v;
    split(rgb, w, h, &y, &u, &v);
    const size_t levels = level_count(w, h, chroma > 0.0f ? 6 : 4);
    Plane *g = (Plane *)calloc(levels, sizeof(Plane));
    Plane **guides = (Plane **)calloc(levels, sizeof(Plane *));
    guides[0] = &y;
    for (size_t l = 1; l < levels; l++) {
        g[l] = down(guides[l - 1]);
        guides[l] = &g[l];
    }
    Model ym;
    const int have_ym = estimate(&y, &y, &ym);
    Plane y2;
    if (luma > 0.0f) {
        const size_t n = level_count(w, h, 4);
        NrPlan plan = {1.0f, NULL, 0, luma <= LUMA_LIGHT};
        y2 = denoise_plane(guides, guides, n, have_ym ? &ym : NULL, &plan);
        blend(&y2, &y, luma_keep(luma));
    } else {
        y2 = plane_copy(&y);
    }
    Plane u2, v2;
    if (chroma > 0.0f) {
        static const size_t csteps[2] = {4, 4};
        NrPlan plan = {chroma_gain(chroma), csteps, 2, 1};
        Plane dn[2];
        Plane *src[2] = {&u, &v};
        for (int c = 0; c < 2; c++) {
            Plane *pc = (Plane *)calloc(levels, sizeof(Plane));
            Plane **pyr = (Plane **)calloc(levels, sizeof(Plane *));
            pyr[0] = src[c];
            for (size_t l = 1; l < levels; l++) {
                pc[l] = down(pyr[l - 1]);
                pyr[l] = &pc[l];
            }
            dn[c] = denoise_plane(pyr, guides, levels, NULL, &plan);
            for (size_t l = 1; l < levels; l++) plane_free(&pc[l]);
            free(pc);
            free(pyr);
        }
        /* the guided filter (Guide::new / filter) */
        const size_t r = (chroma_radius(chroma) + 1) / 2 < 1 ? 1 : (chroma_radius(chroma) + 1) / 2;
        Variance noise;
        if (have_ym) noise = pixel_variance(&ym);
        Plane gi = down(&y2);
        Plane mean = box_mean(&gi, r, box_direct);
        Plane sq = plane_new(gi.w, gi.h);
        for (size_t k = 0; k < gi.w * gi.h; k++) sq.v[k] = gi.v[k] * gi.v[k];
        Plane var = box_mean(&sq, r, box_direct);
        for (size_t k = 0; k < var.w * var.h; k++) var.v[k] = rmaxf(var.v[k] - mean.v[k] * mean.v[k], 0.0f);
        Plane den = plane_new(gi.w, gi.h);
        for (size_t k = 0; k < den.w * den.h; k++) {
            float e = have_ym ? variance_at(&noise, mean.v[k]) : 1e-6f;
            den.v[k] = var.v[k] + rmaxf(e, 1e-9f);
        }
        Plane outp[2];
        for (int c = 0; c < 2; c++) {
            Plane ps = down(&dn[c]);
            Plane mp = box_mean(&ps, r, box_direct);
            Plane prod = plane_new(gi.w, gi.h);
            for (size_t k = 0; k < gi.w * gi.h; k++) prod.v[k] = gi.v[k] * ps.v[k];
            Plane a = box_mean(&prod, r, box_direct);
            for (size_t k = 0; k < a.w * a.h; k++) a.v[k] = (a.v[k] - mean.v[k] * mp.v[k]) / den.v[k];
            Plane bb = plane_new(gi.w, gi.h);
            for (size_t k = 0; k < gi.w * gi.h; k++) bb.v[k] = mp.v[k] - a.v[k] * mean.v[k];
            Plane mb = box_mean(&bb, r, box_direct);
            Plane ma = box_mean(&a, r, box_direct);
            outp[c] = up_affine(&ma, &mb, &y2);
            plane_free(&ps);
            plane_free(&mp);
            plane_free(&prod);
            plane_free(&a);
            plane_free(&bb);
            plane_free(&mb);
            plane_free(&ma);
            plane_free(&dn[c]);
        }
        plane_free(&gi);
        plane_free(&mean);
        plane_free(&sq);
        plane_free(&var);
        plane_free(&den);
        const float keep = chroma_keep(chroma);
        blend(&outp[0], &u, keep);
        blend(&outp[1], &v, keep);
        u2 = outp[0];
        v2 = outp[1];
    } else {
        u2 = plane_copy(&u);
        v2 = plane_copy(&v);
    }
    merge(rgb, &y2, &u2, &v2);
    for (size_t l = 1; l < levels; l++) plane_free(&g[l]);
    free(g);
    free(guides);
    plane_free(&y);
    plane_free(&u);
    plane_free(&v);
    plane_free(&y2);
    plane_free(&u2);
    plane_free(&v2);
    return 0;
}

/* ---- pieces, for pinning ---- */

void fro_nr_dct_block(const float *p, uint32_t w, uint32_t x0, uint32_t y0, float *out) {
    dct_init();
    dct_block(p, w, x0, y0, out);
}

void fro_nr_inverse_block(const float *coef, float *out) {
    dct_init();
    float t[NB][NB]; /* t[j][k]: column j, vertical frequency k */
    for (int k = 0; k < NB; k++) {
        float in[NB], o[NB];
        for (int l = 0; l < NB; l++) in[l] = coef[l * NB + k];
        inv1(in, o);
        for (int j = 0; j < NB; j++) t[j][k] = o[j];
    }
    for (int j = 0; j < NB; j++) {
        float o[NB];
        inv1(t[j], o);
        for (int i = 0; i < NB; i++) out[i * NB + j] = o[i];
    }
}

int fro_nr_box_mean(const float *in, uint32_t w, uint32_t h, uint32_t r, float *out) {
    return fro_nr_box_mean_order(in, w, h, r, out, 0);
}

int fro_nr_box_mean_order(const float *in, uint32_t w, uint32_t h, uint32_t r, float *out, int direct) {
    Plane p = {w, h, (float *)in};
    Plane o = box_mean(&p, r, direct);
    memcpy(out, o.v, (size_t)w * h * sizeof(float));
    plane_free(&o);
    return 0;
}

int fro_nr_down(const float *in, uint32_t w, uint32_t h, float *out) {
    Plane p = {w, h, (float *)in};
    Plane o = down(&p);
    memcpy(out, o.v, o.w * o.h * sizeof(float));
    plane_free(&o);
    return 0;
}

/* d + up(coarse) for a plane d of w x h and coarse its down level */
int fro_nr_add_up(float *d, uint32_t w, uint32_t h, const float *coarse) {
    Plane dp = {w, h, d};
    Plane c = {(w + 1) / 2, (h + 1) / 2, (float *)coarse};
    add_up(&dp, &c);
    return 0;
}

/* One DCT pass with an explicit model (bins means / sigmas), for pinning dct_pass itself. */
int fro_nr_dct_pass(const float *src, const float *guide, const float *pilot, uint32_t w, uint32_t h,
                    uint32_t bins, const float *means, const float *sigmas, float gain, float hard, uint32_t step,
                    float *out) {
    dct_init();
    if (w < NB || h < NB || bins < 1 || bins > MAXBINS || step < 1) return -1;
    Model md;
    md.n = bins;
    memcpy(md.m, means, bins * sizeof(float));
    memcpy(md.s, sigmas, bins * NN * sizeof(float));
    Plane s = {w, h, (float *)src}, g = {w, h, (float *)guide}, pp = {w, h, (float *)pilot};
    Plane o = dct_pass(&s, &g, pilot ? &pp : NULL, &md, gain, hard, step);
    memcpy(out, o.v, (size_t)w * h * sizeof(float));
    plane_free(&o);
    return 0;
}
