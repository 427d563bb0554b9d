/*
 * Seeded synthetic Bayer mosaics for the develop workloads (host code, OpenMP).
 *
 * Scene: per channel, a sum of smooth 2-D gradients (a linear ramp plus three low-frequency
 * cosines), `n_edges` step edges (each a half-plane, axis- or diagonal-aligned, limited to a
 * random box), `n_discs` discs and `n_stripes` patches of square-wave stripes with periods of
 * 2-6 pixels. The scene is scaled so its range maps to 5-95 % of [black, white], mosaicked with
 * the CFA, then given Poisson-Gaussian noise: electrons = signal_DN * gain (Poisson; exact
 * inversion below a mean of 30 electrons, a normal approximation above), back to DN, plus a
 * Gaussian read noise of `read_sigma` DN, plus black, rounded and clipped to [0, white].
 * The parameters of every object come from one SplitMix64 stream of `seed`; the noise of pixel i
 * from a counter-based stream of (seed, i), so the mosaic does not depend on the thread count.
 * frs_scene_clean writes the same mosaic before noise and rounding (black + signal, f32 DN).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
static inline double urange(uint64_t *s, double a, double b) { return a + (b - a) * u01(s); }

typedef struct {
    int kind;       /* 0 horizontal step, 1 vertical, 2 diagonal, 3 anti-diagonal */
    double cx, cy;  /* line position */
    int x0, y0, x1, y1;
    float amp[3];
} Edge;

typedef struct {
    double cx, cy, r;
    float amp[3];
} Disc;

typedef struct {
    int x0, y0, x1, y1;
    int period, dir;
    float amp[3];
} Stripe;

static void fill_box(uint64_t *rng, int w, int h, double fmin, double fmax, int *x0, int *y0, int *x1, int *y1) {
    int bw = (int)(w * urange(rng, fmin, fmax)), bh = (int)(h * urange(rng, fmin, fmax));
    if (bw < 2) bw = 2;
    if (bh < 2) bh = 2;
    int cx = (int)(u01(rng) * w), cy = (int)(u01(rng) * h);
    *x0 = cx - bw / 2 < 0 ? 0 : cx - bw / 2;
    *y0 = cy - bh / 2 < 0 ? 0 : cy - bh / 2;
    *x1 = cx + bw / 2 > w ? w : cx + bw / 2;
    *y1 = cy + bh / 2 > h ? h : cy + bh / 2;
}

/* Poisson variate of mean lam from the stream s. */
static double poisson(uint64_t *s, double lam) {
    if (lam <= 0.0) return 0.0;
    if (lam < 30.0) {
        double u = u01(s), p = exp(-lam), c = p;
        int k = 0;
        while (u > c && k < 1000) {
            k++;
            p *= lam / k;
            c += p;
        }
        return (double)k;
    }
    double u1 = u01(s), u2 = u01(s);
    double n = sqrt(-2.0 * log(u1 > 1e-300 ? u1 : 1e-300)) * cos(2.0 * M_PI * u2);
    double v = floor(lam + sqrt(lam) * n + 0.5);
    return v < 0.0 ? 0.0 : v;
}

static double gauss(uint64_t *s) {
    double u1 = u01(s), u2 = u01(s);
    return sqrt(-2.0 * log(u1 > 1e-300 ? u1 : 1e-300)) * cos(2.0 * M_PI * u2);
}

/* cfa: colour (0 R, 1 G, 2 B) of sensor parity (y & 1) * 2 + (x & 1). One of out (noisy,
 * rounded) and clean (noise-free, f32) is written. Returns 0 or -1. */
static int make_scene(uint16_t *out, float *clean, uint32_t w, uint32_t h, const uint8_t cfa[4], float black, float white,
                      float gain, float read_sigma, uint32_t n_edges, uint32_t n_discs, uint32_t n_stripes, uint64_t seed,
                      int threads) {
    if ((!out && !clean) || w == 0 || h == 0 || !(white > black) || gain <= 0.0f) return -1;
    int nt = threads > 0 ? threads : omp_get_max_threads();
    uint64_t rng = seed * 0x2545f4914f6cdd1dull + 1;
    /* smooth gradients */
    double ramp[3][3], cosp[3][3][4];
    for (int c = 0; c < 3; c++) {
        ramp[c][0] = urange(&rng, 0.2, 0.5);
        ramp[c][1] = urange(&rng, -0.2, 0.2);
        ramp[c][2] = urange(&rng, -0.2, 0.2);
        for (int k = 0; k < 3; k++) {
            cosp[c][k][0] = urange(&rng, 0.05, 0.15);   /* amplitude */
            cosp[c][k][1] = urange(&rng, 0.0, 1.5);     /* cycles across the width */
            cosp[c][k][2] = urange(&rng, 0.0, 1.5);     /* cycles across the height */
            cosp[c][k][3] = urange(&rng, 0.0, 2 * M_PI); /* phase */
        }
    }
    Edge *edges = (Edge *)calloc(n_edges ? n_edges : 1, sizeof(Edge));
    Disc *discs = (Disc *)calloc(n_discs ? n_discs : 1, sizeof(Disc));
    Stripe *stripes = (Stripe *)calloc(n_stripes ? n_stripes : 1, sizeof(Stripe));
    float *scene = (float *)malloc((size_t)w * h * 3 * sizeof(float));
    if (!edges || !discs || !stripes || !scene) {
        free(edges); free(discs); free(stripes); free(scene);
        return -1;
    }
    for (uint32_t i = 0; i < n_edges; i++) {
        Edge *e = &edges[i];
        e->kind = (int)(splitmix64(&rng) % 4);
        fill_box(&rng, (int)w, (int)h, 0.1, 0.35, &e->x0, &e->y0, &e->x1, &e->y1);
        e->cx = urange(&rng, e->x0, e->x1);
        e->cy = urange(&rng, e->y0, e->y1);
        for (int c = 0; c < 3; c++) e->amp[c] = (float)urange(&rng, -0.3, 0.3);
    }
    double side = w < h ? w : h;
    for (uint32_t i = 0; i < n_discs; i++) {
        Disc *d = &discs[i];
        d->cx = u01(&rng) * w;
        d->cy = u01(&rng) * h;
        d->r = side * urange(&rng, 0.01, 0.08);
        for (int c = 0; c < 3; c++) d->amp[c] = (float)urange(&rng, -0.3, 0.3);
    }
    for (uint32_t i = 0; i < n_stripes; i++) {
        Stripe *s = &stripes[i];
        int bw = (int)urange(&rng, 32, 128), bh = (int)urange(&rng, 32, 128);
        s->x0 = (int)(u01(&rng) * (w > (uint32_t)bw ? w - bw : 1));
        s->y0 = (int)(u01(&rng) * (h > (uint32_t)bh ? h - bh : 1));
        s->x1 = s->x0 + bw > (int)w ? (int)w : s->x0 + bw;
        s->y1 = s->y0 + bh > (int)h ? (int)h : s->y0 + bh;
        s->period = 2 + (int)(splitmix64(&rng) % 5);
        s->dir = (int)(splitmix64(&rng) % 4);
        for (int c = 0; c < 3; c++) s->amp[c] = (float)urange(&rng, 0.1, 0.3);
    }
    /* base gradients */
#pragma omp parallel for num_threads(nt) schedule(static)
    for (uint32_t y = 0; y < h; y++) {
        double fy = (double)y / h;
        for (uint32_t x = 0; x < w; x++) {
            double fx = (double)x / w;
            for (int c = 0; c < 3; c++) {
                double v = ramp[c][0] + ramp[c][1] * fx + ramp[c][2] * fy;
                for (int k = 0; k < 3; k++)
                    v += cosp[c][k][0] * cos(2 * M_PI * (cosp[c][k][1] * fx + cosp[c][k][2] * fy) + cosp[c][k][3]);
                scene[((size_t)y * w + x) * 3 + c] = (float)v;
            }
        }
    }
    /* objects, in a fixed order, row-parallel (each object touches each pixel at most once) */
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
    for (uint32_t y = 0; y < h; y++) {
        float *row = scene + (size_t)y * w * 3;
        for (uint32_t i = 0; i < n_edges; i++) {
            const Edge *e = &edges[i];
            if ((int)y < e->y0 || (int)y >= e->y1) continue;
            for (int x = e->x0; x < e->x1; x++) {
                double dx = x - e->cx, dy = (double)y - e->cy;
                double t = e->kind == 0 ? dy : e->kind == 1 ? dx : e->kind == 2 ? dx + dy : dx - dy;
                if (t > 0)
                    for (int c = 0; c < 3; c++) row[(size_t)x * 3 + c] += e->amp[c];
            }
        }
        for (uint32_t i = 0; i < n_discs; i++) {
            const Disc *d = &discs[i];
            double dy = (double)y - d->cy;
            if (fabs(dy) >= d->r) continue;
            double half = sqrt(d->r * d->r - dy * dy);
            int x0 = (int)ceil(d->cx - half), x1 = (int)floor(d->cx + half);
            if (x0 < 0) x0 = 0;
            if (x1 > (int)w - 1) x1 = (int)w - 1;
            for (int x = x0; x <= x1; x++)
                for (int c = 0; c < 3; c++) row[(size_t)x * 3 + c] += d->amp[c];
        }
        for (uint32_t i = 0; i < n_stripes; i++) {
            const Stripe *s = &stripes[i];
            if ((int)y < s->y0 || (int)y >= s->y1) continue;
            for (int x = s->x0; x < s->x1; x++) {
                int t = s->dir == 0 ? (int)y : s->dir == 1 ? x : s->dir == 2 ? x + (int)y : x - (int)y + (int)h;
                float sgn = ((t % s->period) * 2 < s->period) ? 1.0f : -1.0f;
                for (int c = 0; c < 3; c++) row[(size_t)x * 3 + c] += sgn * s->amp[c];
            }
        }
    }
    /* range -> 5..95 % */
    float lo = 1e30f, hi = -1e30f;
#pragma omp parallel for num_threads(nt) reduction(min : lo) reduction(max : hi)
    for (size_t i = 0; i < (size_t)w * h * 3; i++) {
        float v = scene[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    double span = hi > lo ? hi - lo : 1.0;
    double range = white - black;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            int c = cfa[(y & 1) * 2 + (x & 1)];
            double s01 = 0.05 + 0.9 * (scene[i * 3 + c] - lo) / span;
            double sig = s01 * range;
            if (clean) {
                clean[i] = (float)(black + sig);
                continue;
            }
            uint64_t st = seed ^ (i * 0xd1b54a32d192ed03ull) ^ 0x8bb84b93962eacc9ull;
            double e = poisson(&st, sig * gain) / gain;
            double v = black + e + read_sigma * gauss(&st);
            v = floor(v + 0.5);
            v = v < 0 ? 0 : (v > white ? white : v);
            out[i] = (uint16_t)v;
        }
    }
    free(edges);
    free(discs);
    free(stripes);
    free(scene);
    return 0;
}

int frs_scene(uint16_t *out, uint32_t w, uint32_t h, const uint8_t cfa[4], float black, float white, float gain,
              float read_sigma, uint32_t n_edges, uint32_t n_discs, uint32_t n_stripes, uint64_t seed, int threads) {
    if (!out) return -1;
    return make_scene(out, NULL, w, h, cfa, black, white, gain, read_sigma, n_edges, n_discs, n_stripes, seed, threads);
}

int frs_scene_clean(float *clean, uint32_t w, uint32_t h, const uint8_t cfa[4], float black, float white, float gain,
                    uint32_t n_edges, uint32_t n_discs, uint32_t n_stripes, uint64_t seed, int threads) {
    if (!clean) return -1;
    return make_scene(NULL, clean, w, h, cfa, black, white, gain, 0.0f, n_edges, n_discs, n_stripes, seed, threads);
}
