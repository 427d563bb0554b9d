// Luma / chroma noise reduction of developed linear RGB (gfx950): the kernels of the method of
// crates/flashraw/src/denoise/mod.rs on f32 planes in HBM.
//
// Pointwise and separable steps are one thread per sample: split / merge into luminance and the
// two colour differences (433-464), 2x2 means (down 467-486), bilinear upsampling (tap / up_row
// 488-515, add_up, up_affine), strength blends, box means (280-335, direct sums instead of the
// CPU's running sums) and the guided filter's coefficients (337-382). The noise estimate measures
// one block per thread (estimate 761-788). The DCT passes (dct_pass / run_group_impl 900-1131)
// run in tiles of 64 x 64 block positions, one wave per tile: the wave's 64 lanes take the 64
// blocks of one position class (positions congruent modulo 8 do not overlap), transform them in
// registers (the reference's even / odd 8-point factorisation, vertical then horizontal),
// threshold or Wiener-filter them against the model's deviations at the block's luma mean, and
// add the weighted inverse blocks to the tile's 72 x 72 sums in LDS, class after class, so no
// two lanes add to one sample at once and the sums run in a fixed order. A gather kernel adds the
// (at most four) tile partials of every sample in a fixed order and divides by the weights.
#include <hip/hip_runtime.h>

#include "develop_common.h"
#include "nr.h"

namespace frh {
namespace {

constexpr float KR = 0.2126f, KG = 0.7152f, KB = 0.0722f;

__device__ __forceinline__ float dot4(float c0, float c1, float c2, float c3, float x0, float x1, float x2,
                                      float x3) {
    return (c0 * x0 + c1 * x1) + (c2 * x2 + c3 * x3);
}

// Global-memory views of the planes a pass descriptor (itself read from memory) points at: named
// as such, their loads and stores are global ones instead of flat ones, which would also count
// against the LDS wait counter.
typedef __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ const gfloat *gl(const float *p) { return (const gfloat *)p; }
__device__ __forceinline__ gfloat *gl(float *p) { return (gfloat *)p; }

// The DCT matrix entries the factorisation reads (Dct::fwd1 / inv1, denoise/mod.rs:643-686).
struct D8 {
    float m00, m20, m21, m40, m60, m61;
    float o[4][4];  // odd rows 1, 3, 5, 7, columns 0..3
};

template <class T>
__device__ __forceinline__ D8 load_d8(const T *__restrict__ d) {
    D8 r;
    r.m00 = d[0];
    r.m20 = d[16];
    r.m21 = d[17];
    r.m40 = d[32];
    r.m60 = d[48];
    r.m61 = d[49];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) r.o[q][i] = d[(2 * q + 1) * 8 + i];
    return r;
}

// out[k] = sum_i m[k][i] x[i]
__device__ __forceinline__ void fwd1(const D8 &D, const float x[8], float out[8]) {
    float s[4], d[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s[i] = x[i] + x[7 - i];
        d[i] = x[i] - x[7 - i];
    }
    const float e0 = s[0] + s[3], e1 = s[1] + s[2], e2 = s[0] - s[3], e3 = s[1] - s[2];
    out[0] = D.m00 * (e0 + e1);
    out[1] = dot4(D.o[0][0], D.o[0][1], D.o[0][2], D.o[0][3], d[0], d[1], d[2], d[3]);
    out[2] = D.m20 * e2 + D.m21 * e3;
    out[3] = dot4(D.o[1][0], D.o[1][1], D.o[1][2], D.o[1][3], d[0], d[1], d[2], d[3]);
    out[4] = D.m40 * (e0 - e1);
    out[5] = dot4(D.o[2][0], D.o[2][1], D.o[2][2], D.o[2][3], d[0], d[1], d[2], d[3]);
    out[6] = D.m60 * e2 + D.m61 * e3;
    out[7] = dot4(D.o[3][0], D.o[3][1], D.o[3][2], D.o[3][3], d[0], d[1], d[2], d[3]);
}

// out[i] = sum_k m[k][i] x[k]
__device__ __forceinline__ void inv1(const D8 &D, const float x[8], float out[8]) {
    const float p = D.m00 * (x[0] + x[4]);
    const float q = D.m00 * (x[0] - x[4]);
    const float u = D.m20 * x[2] + D.m60 * x[6];
    const float v = D.m21 * x[2] + D.m61 * x[6];
    const float e[4] = {p + u, q + v, q - v, p - u};
    float o[4];
#pragma unroll
    for (int i = 0; i < 4; i++) o[i] = dot4(D.o[0][i], D.o[1][i], D.o[2][i], D.o[3][i], x[1], x[3], x[5], x[7]);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        out[i] = e[i] + o[i];
        out[7 - i] = e[i] - o[i];
    }
}

// Four consecutive floats from a 4-byte-aligned address (one dwordx4 load).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef __attribute__((address_space(1))) f4u gf4u;
__device__ __forceinline__ f4u ld4(const gfloat *q) { return *(const gf4u *)q; }
__device__ __forceinline__ f4u ld4(const float *q) { return *(const f4u *)q; }

// Columns 4h .. 4h + 3 of a block, b[i][j] at p[row[i] + col[4h + j]]: one 16-byte load per row
// when the block's columns are consecutive (every block off the reflected border), else one load
// per sample.
template <class T>
__device__ __forceinline__ void load_half(const T *__restrict__ p, const uint32_t row[8], const uint32_t col[8], int h,
                                          float b[8][4]) {
    if (col[7] == col[0] + 7) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const f4u q = ld4(p + row[i] + col[0] + 4 * h);
            b[i][0] = q.x;
            b[i][1] = q.y;
            b[i][2] = q.z;
            b[i][3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) b[i][j] = p[row[i] + col[4 * h + j]];
    }
}

// The 2-D transform of a block given its rows as addresses: c[l][k] is horizontal frequency l,
// vertical frequency k (Dct::block: the columns' vertical transforms, then the horizontal one).
// The vertical stage of a block's transform: v[k][j], vertical frequency k of column j.
template <class T>
__device__ __forceinline__ void block_vert(const D8 &D, const T *__restrict__ p, const uint32_t row[8],
                                           const uint32_t col[8], float v[8][8]) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float b[8][4];
        load_half(p, row, col, h, b);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float x[8], o[8];
#pragma unroll
            for (int i = 0; i < 8; i++) x[i] = b[i][j];
            fwd1(D, x, o);
#pragma unroll
            for (int k = 0; k < 8; k++) v[k][4 * h + j] = o[k];
        }
    }
}

template <class T>
__device__ __forceinline__ void block_fwd(const D8 &D, const T *__restrict__ p, const uint32_t row[8],
                                          const uint32_t col[8], float c[8][8]) {
    float v[8][8];  // v[k][j]
    block_vert(D, p, row, col, v);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        float o[8];
        fwd1(D, v[k], o);
#pragma unroll
        for (int l = 0; l < 8; l++) c[l][k] = o[l];
    }
}

// pad_map (denoise/mod.rs:871-886): plane index of padded position q of a side of n samples.
__device__ __forceinline__ uint32_t reflect8(int q, int n) {
    const int i = q - 8;
    const int r = i < 0 ? -i - 1 : (i >= n ? 2 * n - i - 1 : i);
    return (uint32_t)min(max(r, 0), n - 1);
}

// tap (denoise/mod.rs:488-494): the two taps and the second one's weight of bilinear upsampling
// to position i from a level of side n.
__device__ __forceinline__ void tap(uint32_t i, uint32_t n, uint32_t &a, uint32_t &b, float &f) {
    const float s = fmaxf(((float)i + 0.5f) * 0.5f - 0.5f, 0.0f);
    a = min((uint32_t)s, n - 1);
    b = min(a + 1, n - 1);
    f = s - (float)a;
}

__device__ __forceinline__ float up_at(const float *__restrict__ p, uint32_t pw, uint32_t ph, uint32_t x, uint32_t y) {
    uint32_t x0, x1, y0, y1;
    float fx, fy;
    tap(x, pw, x0, x1, fx);
    tap(y, ph, y0, y1, fy);
    const float *r0 = p + (size_t)y0 * pw, *r1 = p + (size_t)y1 * pw;
    const float a = r0[x0] + (r0[x1] - r0[x0]) * fx;
    const float b = r1[x0] + (r1[x1] - r1[x0]) * fx;
    return a + (b - a) * fy;
}

__global__ __launch_bounds__(256) void k_split(const float *__restrict__ rgb, uint64_t n, float *__restrict__ y,
                                               float *__restrict__ u, float *__restrict__ v) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const float r = rgb[3 * k], g = rgb[3 * k + 1], b = rgb[3 * k + 2];
    const float l = KR * r + KG * g + KB * b;
    y[k] = l;
    u[k] = b - l;
    v[k] = r - l;
}

// 2x2 mean (the last row or column repeated on odd sides); with `minus`, minus - mean.
__global__ __launch_bounds__(256) void k_down(const float *__restrict__ p, uint32_t w, uint32_t h,
                                              float *__restrict__ o, const float *__restrict__ minus) {
    const uint32_t dw = (w + 1) / 2, dh = (h + 1) / 2;
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= dw || y >= dh) return;
    const uint32_t y0 = 2 * y, y1 = min(2 * y + 1, h - 1);
    const float *r0 = p + (size_t)y0 * w, *r1 = p + (size_t)y1 * w;
    float m;
    if (2 * x + 1 < w) {
        m = 0.25f * (r0[2 * x] + r0[2 * x + 1] + r1[2 * x] + r1[2 * x + 1]);
    } else {
        const uint32_t c = w - 1;
        m = 0.25f * (r0[c] + r0[c] + r1[c] + r1[c]);
    }
    const size_t k = (size_t)y * dw + x;
    o[k] = minus ? minus[k] - m : m;
}

__global__ __launch_bounds__(256) void k_add_up(float *__restrict__ d, uint32_t w, uint32_t h,
                                                const float *__restrict__ c) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    d[(size_t)y * w + x] += up_at(c, (w + 1) / 2, (h + 1) / 2, x, y);
}

__global__ __launch_bounds__(256) void k_blend(float *__restrict__ d, const float *__restrict__ o, float keep,
                                               uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n) d[k] += keep * (o[k] - d[k]);
}

__global__ __launch_bounds__(256) void k_mul(const float *__restrict__ a, const float *__restrict__ b,
                                             float *__restrict__ o, uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < n) o[k] = a[k] * b[k];
}

// Box sums of radius r with repeated edges: vertical, then horizontal divided by (2r + 1)^2.
__global__ __launch_bounds__(256) void k_box_v(const float *__restrict__ a, uint32_t w, uint32_t h, uint32_t r,
                                               float *__restrict__ o) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    float s = 0.0f;
    for (int d = -(int)r; d <= (int)r; d++) {
        const int yy = min(max((int)y + d, 0), (int)h - 1);
        s += a[(size_t)yy * w + x];
    }
    o[(size_t)y * w + x] = s;
}

__global__ __launch_bounds__(256) void k_box_h(const float *__restrict__ a, uint32_t w, uint32_t h, uint32_t r,
                                               float n, float *__restrict__ o) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const float *row = a + (size_t)y * w;
    float s = 0.0f;
    for (int d = -(int)r; d <= (int)r; d++) s += row[min(max((int)x + d, 0), (int)w - 1)];
    o[(size_t)y * w + x] = s / n;
}

// Variance::at (denoise/mod.rs:261-275)
__device__ __forceinline__ float var_at(const NrVar *__restrict__ nv, float m) {
    const uint32_t n = nv->n;
    uint32_t i = 0;
    while (i < n && nv->m[i] < m) i++;
    if (i == 0) return nv->v[0];
    if (i >= n) return nv->v[n - 1];
    const float a = nv->m[i - 1], b = nv->m[i];
    const float t = b > a ? (m - a) / (b - a) : 0.0f;
    return nv->v[i - 1] + (nv->v[i] - nv->v[i - 1]) * t;
}

// The guided filter's denominator: max(E[i^2] - E[i]^2, 0) + max(eps(E[i]), 1e-9), eps the luma
// noise variance per pixel at the window's mean (1e-6 without a model).
__global__ __launch_bounds__(256) void k_den(float *__restrict__ var, const float *__restrict__ mean,
                                             const NrVar *__restrict__ nv, uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const float m = mean[k];
    const float v = fmaxf(var[k] - m * m, 0.0f);
    const float e = nv->n ? var_at(nv, m) : 1e-6f;
    var[k] = v + fmaxf(e, 1e-9f);
}

__global__ __launch_bounds__(256) void k_gf_coef(const float *__restrict__ cov, const float *__restrict__ mp,
                                                 const float *__restrict__ mean, const float *__restrict__ den,
                                                 float *__restrict__ a, float *__restrict__ b, uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const float av = (cov[k] - mean[k] * mp[k]) / den[k];
    a[k] = av;
    b[k] = mp[k] - av * mean[k];
}

// up(mb) + up(ma) * g (up_affine, denoise/mod.rs:537-555)
__global__ __launch_bounds__(256) void k_gf_apply(const float *__restrict__ ma, const float *__restrict__ mb,
                                                  const float *__restrict__ g, uint32_t w, uint32_t h,
                                                  float *__restrict__ o) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint32_t cw = (w + 1) / 2, ch = (h + 1) / 2;
    const size_t k = (size_t)y * w + x;
    o[k] = up_at(mb, cw, ch, x, y) + up_at(ma, cw, ch, x, y) * g[k];
}

// The measured blocks (estimate, denoise/mod.rs:761-788): block (c, r) of the grid of `step`
// blocks; the guide's mean over the block (row-major sum / 64), then the squared coefficients.
__global__ __launch_bounds__(64) void k_stats(const float *__restrict__ p, const float *__restrict__ g, uint32_t w,
                                              uint32_t step, uint32_t cols, uint32_t rows,
                                              const float *__restrict__ dct, float *__restrict__ out) {
    // one block per lane; the 64 blocks' rows go out through LDS, so each store instruction writes
    // consecutive words (the row pitch 65 keeps the lanes' LDS writes on distinct banks)
    __shared__ float o[64][65];
    const uint32_t b0 = blockIdx.x * 64, idx = b0 + threadIdx.x, n = cols * rows;
    if (idx < n) {
        const D8 D = load_d8(dct);
        const uint32_t x0 = (idx % cols) * step * 8, y0 = (idx / cols) * step * 8;
        uint32_t row[8], col[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            row[i] = (y0 + i) * w;
            col[i] = x0 + i;
        }
        float mean = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 8; j++) mean += g[row[i] + col[j]];
        float c[8][8];
        block_fwd(D, p, row, col, c);
#pragma unroll
        for (int l = 0; l < 8; l++)
#pragma unroll
            for (int k = 0; k < 8; k++) o[threadIdx.x][l * 8 + k] = (l | k) ? c[l][k] * c[l][k] : mean / 64.0f;
    }
    __syncthreads();
    const uint32_t nb = min(64u, n - b0);
    float *dst = out + (size_t)b0 * 64;
    for (uint32_t t = threadIdx.x; t < nb * 64; t += 64) dst[t] = o[t / 64][t % 64];
}

// The pass of a launch's list that item `i` (a tile or a gather chunk) belongs to: the last one
// whose first item is at most i (`first` the member holding it). Wave-uniform.
template <uint32_t NrPass::*first>
__device__ __forceinline__ const NrPass &pass_of(const NrPass *__restrict__ list, uint32_t n, uint32_t i) {
    uint32_t p = 0;
    while (p + 1 < n && i >= list[p + 1].*first) p++;
    return list[__builtin_amdgcn_readfirstlane(p)];
}

// One DCT pass, one wave per tile of 64 x 64 block positions (see the file comment).
template <bool WIENER>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_pass(const NrPass *__restrict__ list,
                                                                                    uint32_t npass) {
    __shared__ float acc[NR_EXT * NR_EXT];
    __shared__ float sm[NR_BINS], ss[NR_BINS * 64];
    const NrPass &P = pass_of<&NrPass::tile0>(list, npass, blockIdx.x);
    const uint32_t tile = blockIdx.x - P.tile0;
    const int lane = threadIdx.x;
    const int tx = (int)(tile % P.tiles_x), ty = (int)(tile / P.tiles_x);
    const int TX = tx * NR_TILE, TY = ty * NR_TILE;
    for (int i = lane; i < NR_EXT * NR_EXT; i += 64) acc[i] = 0.0f;
    const __attribute__((address_space(1))) NrModel *model = (const __attribute__((address_space(1))) NrModel *)P.model;
    const uint32_t nb = model->n;
    for (int i = lane; i < (int)nb * 64; i += 64) ss[i] = model->s[i / 64][i % 64];
    if (lane < (int)nb) sm[lane] = model->m[lane];
    __syncthreads();
    const D8 D = load_d8(gl(P.dct));
    const int step = (int)P.step, ncls = 8 / step;
    const int xlast = (int)P.xlast, ylast = (int)P.ylast;
    const bool xirr = xlast % step != 0, yirr = ylast % step != 0;
    const int w = (int)P.w, h = (int)P.h;
    const int bi = lane & 7, bj = lane >> 3;
    for (int cy = 0; cy <= ncls; cy++) {
        if (cy == ncls && !yirr) break;
        for (int cx = 0; cx <= ncls; cx++) {
            if (cx == ncls && !xirr) break;
            int x0, y0;
            bool valid;
            if (cx < ncls) {
                x0 = TX + cx * step + 8 * bi;
                valid = x0 <= xlast;
            } else {
                x0 = xlast;
                valid = bi == 0 && xlast >= TX && xlast < TX + NR_TILE;
            }
            if (cy < ncls) {
                y0 = TY + cy * step + 8 * bj;
                valid = valid && y0 <= ylast;
            } else {
                y0 = ylast;
                valid = valid && bj == 0 && ylast >= TY && ylast < TY + NR_TILE;
            }
            if (valid) {
                uint32_t row[8], col[8];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    row[i] = reflect8(y0 + i, h) * (uint32_t)w;
                    col[i] = reflect8(x0 + i, w);
                }
                // the guide's mean over the block: column sums, then their sum / 64
                float gs = 0.0f;
                {
                    float gcol[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) gcol[j] = 0.0f;
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        float gb[8][4];
                        load_half(gl(P.guide), row, col, h, gb);
#pragma unroll
                        for (int i = 0; i < 8; i++)
#pragma unroll
                            for (int j = 0; j < 4; j++) gcol[4 * h + j] += gb[i][j];
                    }
#pragma unroll
                    for (int j = 0; j < 8; j++) gs += gcol[j];
                }
                const float mean = gs / 64.0f;
                // Model::sigmas: the bins around the mean and the weight of the upper one
                uint32_t bin = 0;
                while (bin < nb && sm[bin] < mean) bin++;
                uint32_t ia, ib;
                float t;
                if (bin == 0 || nb == 1) {
                    ia = ib = 0;
                    t = 0.0f;
                } else if (bin >= nb) {
                    ia = ib = nb - 1;
                    t = 0.0f;
                } else {
                    ia = bin - 1;
                    ib = bin;
                    const float a = sm[ia], b = sm[ib];
                    t = b > a ? (mean - a) / (b - a) : 0.0f;
                }
                const float *sa = ss + ia * 64, *sb = ss + ib * 64;
                float c[8][8];
                float weight;
                if (WIENER) {
                    // the source block's coefficients first; then the pilot's, one vertical
                    // frequency k at a time (its horizontal transform), each shrinking the source
                    // coefficients of that k as it appears, so at most 64 + 64 values are live
                    block_fwd(D, gl(P.src), row, col, c);
                    const float dc = c[0][0];
                    float vp[8][8];
                    block_vert(D, gl(P.pilot), row, col, vp);
                    float e[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        float pc[8];
                        fwd1(D, vp[k], pc);  // pc[l]: the pilot's coefficient (l, k)
                        e[k] = 0.0f;
#pragma unroll
                        for (int l = 0; l < 8; l++) {
                            const float a = sa[l * 8 + k], b = sb[l * 8 + k];
                            const float sg = a + (b - a) * t;
                            const float n2 = sg * sg * P.n2scale;
                            const float p2 = pc[l] * pc[l];
                            const float f = p2 / (p2 + n2 + 1e-30f);
                            c[l][k] = c[l][k] * f;
                            e[k] = e[k] + f * f * n2;
                        }
                    }
                    c[0][0] = dc;
                    const float es = ((e[0] + e[4]) + (e[1] + e[5])) + ((e[2] + e[6]) + (e[3] + e[7]));
                    weight = 1.0f / (es + 1e-20f);
                } else {
                    block_fwd(D, gl(P.src), row, col, c);
                    uint32_t kept = 0;
#pragma unroll
                    for (int l = 0; l < 8; l++)
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            const float a = sa[l * 8 + k], b = sb[l * 8 + k];
                            const float sg = a + (b - a) * t;
                            const bool keep = fabsf(c[l][k]) >= sg * P.thr;
                            kept += keep ? 1u : 0u;
                            c[l][k] = keep ? c[l][k] : 0.0f;
                        }
                    weight = 1.0f / (float)kept;
                }
                // the block's weight in the grid of block positions
                {
                    const uint32_t ix = cx < ncls ? (uint32_t)x0 / (uint32_t)step : P.gx - 1;
                    const uint32_t iy = cy < ncls ? (uint32_t)y0 / (uint32_t)step : P.gy - 1;
                    gl(P.wgrid)[(size_t)iy * P.gx + ix] = weight;
                }
                // weighted inverse: horizontal (per vertical frequency), then vertical (per column)
                float tcol[8][8];  // tcol[j][k]
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    float x[8], o[8];
#pragma unroll
                    for (int l = 0; l < 8; l++) x[l] = c[l][k] * weight;
                    inv1(D, x, o);
#pragma unroll
                    for (int j = 0; j < 8; j++) tcol[j][k] = o[j];
                }
                const int ox = x0 - TX, oy = y0 - TY;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    float o[8];
                    inv1(D, tcol[j], o);
#pragma unroll
                    for (int i = 0; i < 8; i++) acc[(oy + i) * NR_EXT + ox + j] += o[i];
                }
            }
            __syncthreads();
        }
    }
    gfloat *pa = gl(P.pacc) + (size_t)tile * (NR_EXT * NR_EXT);
    for (int i = lane; i < NR_EXT * NR_EXT; i += 64) pa[i] = acc[i];
}

// The block positions along one padded side that cover padded coordinate c: the regular ones
// lo..hi (hi < lo: none) and, when the last position is off the step and covers c, index `last_i`.
struct Cover {
    int lo, hi;
    bool last;
};
__device__ __forceinline__ Cover cover(int c, int last, int step) {
    Cover r;
    r.lo = c >= 7 ? (c - 7 + step - 1) / step : 0;
    r.hi = min(c / step, last / step);
    r.last = last % step != 0 && c >= last && c <= last + 7;
    return r;
}

// Every sample: the partials of the tiles whose blocks cover it, in tile order, over its cell's
// sum of the covering blocks' weights (the sample itself where no block has weight). Four
// consecutive samples of a row per thread (a chunk is 1024 samples of one row).
__global__ __launch_bounds__(256) void k_gather(const NrPass *__restrict__ list, uint32_t npass) {
    const NrPass &P = pass_of<&NrPass::chunk0>(list, npass, blockIdx.x);
    const uint32_t cpr = (P.w + 1023) / 1024, local = blockIdx.x - P.chunk0;
    const uint32_t y = local / cpr, xb = (local % cpr) * 1024 + threadIdx.x * 4;
    if (y >= P.h) return;
    const int py = (int)y + 8, ty1 = py / NR_TILE;
    const uint32_t step = P.step, ncx = P.xlast / step + 1;
    const gfloat *wrow = gl(P.wsum) + (size_t)((uint32_t)py / step) * ncx;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t x = xb + (uint32_t)q;
        if (x >= P.w) break;
        const int px = (int)x + 8, tx1 = px / NR_TILE;
        float a = 0.0f;
        for (int ty = ty1 - 1; ty <= ty1; ty++) {
            if (ty < 0 || ty >= (int)P.tiles_y || py - ty * NR_TILE >= NR_EXT) continue;
            for (int tx = tx1 - 1; tx <= tx1; tx++) {
                if (tx < 0 || tx >= (int)P.tiles_x || px - tx * NR_TILE >= NR_EXT) continue;
                a += gl(P.pacc)[(size_t)(ty * (int)P.tiles_x + tx) * (NR_EXT * NR_EXT) +
                                (size_t)((py - ty * NR_TILE) * NR_EXT + (px - tx * NR_TILE))];
            }
        }
        const float ws = wrow[(uint32_t)px / step];
        const size_t k = (size_t)y * P.w + x;
        gl(P.out)[k] = ws > 0.0f ? a / ws : gl(P.src)[k];
    }
}

// The weights of the blocks covering each cell of samples (padded coordinates with the same
// px / step, py / step), summed in position order: rows of block positions from the top, each row
// from the left. The last position, off the step, covers only padding beyond the plane (it
// starts at w + 8, past the last sample's w + 7), so it adds to no cell a sample reads.
__global__ __launch_bounds__(256) void k_wsum(const NrPass *__restrict__ list, uint32_t npass) {
    const NrPass &P = pass_of<&NrPass::cell0>(list, npass, blockIdx.x);
    const uint32_t step = P.step, ncx = P.xlast / step + 1, ncy = P.ylast / step + 1;
    const uint32_t c = (blockIdx.x - P.cell0) * 256 + threadIdx.x;
    if (c >= ncx * ncy) return;
    const Cover cxr = cover((int)((c % ncx) * step), (int)P.xlast, (int)step);
    const Cover cyr = cover((int)((c / ncx) * step), (int)P.ylast, (int)step);
    float ws = 0.0f;
    auto row = [&](uint32_t iy) {
        const gfloat *g = gl(P.wgrid) + (size_t)iy * P.gx;
        for (int ix = cxr.lo; ix <= cxr.hi; ix++) ws += g[ix];
        if (cxr.last) ws += g[P.gx - 1];
    };
    for (int iy = cyr.lo; iy <= cyr.hi; iy++) row((uint32_t)iy);
    if (cyr.last) row(P.gy - 1);
    gl(P.wsum)[c] = ws;
}

// merge (denoise/mod.rs:450-464) and the develop's encoders (tone, curve, orientation)
__global__ __launch_bounds__(256) void k_encode(DevArgs a, const float *__restrict__ py,
                                                const float *__restrict__ pu, const float *__restrict__ pv) {
    __shared__ Lds s;
    load_lds(a, s);
    __syncthreads();
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= a.dev_w || y >= a.dev_h) return;
    const size_t k = (size_t)y * a.dev_w + x;
    const float yv = py[k];
    const float r = pv[k] + yv, b = pu[k] + yv;
    const float g = (yv - KR * r - KB * b) / KG;
    store_both(a, s, x, y, r, g, b);
}

inline dim3 grid2(uint32_t w, uint32_t h) { return dim3((w + 255) / 256, h); }
inline uint32_t grid1(uint64_t n) { return (uint32_t)((n + 255) / 256); }


// ---- the noise model fit (Model::from_stats, denoise/mod.rs:803-848, on the device) ----

// f32::total_cmp order as unsigned keys (the float's total-order key with its sign bit flipped),
// and back.
__device__ __forceinline__ uint32_t order_key(float f) {
    const int32_t i = __float_as_int(f);
    return (uint32_t)(i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1)) ^ 0x80000000u;
}
__device__ __forceinline__ float order_float(uint32_t k) {
    const int32_t t = (int32_t)(k ^ 0x80000000u);
    return __int_as_float(t ^ (int32_t)(((uint32_t)(t >> 31)) >> 1));
}

constexpr int FT = 256;  // threads per block of the fit kernels

// Bins of n finite blocks: at most 12, n / 96 of them, `per` blocks each (the last may hold
// fewer, and there may be fewer bins than asked).
__device__ __forceinline__ void fit_bins(uint32_t n, uint32_t &nb, uint32_t &per) {
    const uint32_t bins = min(max(n / 96u, 1u), (uint32_t)NR_BINS);
    per = (n + bins - 1) / bins;
    nb = (n + per - 1) / per;
}

// The key of rank q (0-based; q below the number of candidates) among get(i), i in [0, n), for
// which keep(i) holds: four passes over 8-bit digits, most significant first. *rest receives the
// rank among the keys equal to the result. Every thread of the block calls it; `hist` is 258
// words of LDS.
template <class Get, class Keep>
__device__ uint32_t select_rank(Get get, Keep keep, uint32_t n, uint32_t q, uint32_t *hist, uint32_t *rest) {
    uint32_t prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < 256; i += FT) hist[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += FT) {
            if (!keep(i)) continue;
            const uint32_t k = get(i);
            if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t cum = 0, d = 0;
            while (d < 255 && cum + hist[d] <= q) cum += hist[d++];
            hist[256] = d;
            hist[257] = q - cum;
        }
        __syncthreads();
        prefix |= hist[256] << shift;
        mask |= 255u << shift;
        q = hist[257];
        __syncthreads();
    }
    *rest = q;
    return prefix;
}

// Per block: the activity (the squared AC coefficients summed in order), the mean's key (all ones
// for a block whose mean or activity is not finite) and the region's count of finite blocks. 64
// blocks per workgroup, read in consecutive words into LDS (row pitch 65), summed one per lane.
__global__ __launch_bounds__(256) void k_fit_keys(const float *__restrict__ stats, uint32_t total,
                                                  const NrRegion *__restrict__ reg, uint32_t nreg,
                                                  uint32_t *__restrict__ key, float *__restrict__ act, NrFit *fit) {
    __shared__ float rowsl[64][65];
    const uint32_t b0 = blockIdx.x * 64, nb = min(64u, total - b0);
    const float *src = stats + (size_t)b0 * 64;
    for (uint32_t t = threadIdx.x; t < nb * 64; t += 256) rowsl[t / 64][t % 64] = src[t];
    __syncthreads();
    if (threadIdx.x >= nb) return;
    const uint32_t i = b0 + threadIdx.x;
    const float *s = rowsl[threadIdx.x];
    float a = 0.0f;
    for (int k = 1; k < 64; k++) a = a + s[k];
    const bool fin = __builtin_isfinite(s[0]) && __builtin_isfinite(a);
    key[i] = fin ? order_key(s[0]) : 0xffffffffu;
    act[i] = a;
    uint32_t r = 0;
    while (r + 1 < nreg && i >= reg[r + 1].first) r++;
    // one atomic per wave when its blocks lie in one region (the count's address is shared by
    // the whole region, so per-block atomics would serialise)
    const uint32_t r0 = __builtin_amdgcn_readfirstlane(r);
    const uint64_t finite = __ballot(fin);
    if (__ballot(r != r0) == 0) {
        if (threadIdx.x == (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x) && finite)
            atomicAdd(&fit[r0].n, (uint32_t)__popcll(finite));
    } else if (fin) {
        atomicAdd(&fit[r].n, 1u);
    }
}

// Block (r, q): the block of rank b * per (q = b) or the bin's median rank b * per + len / 2
// (q = NR_BINS + b) in the order by mean, ties in block order.
__global__ __launch_bounds__(FT) void k_fit_bounds(const float *__restrict__ stats, const NrRegion *__restrict__ reg,
                                                   const uint32_t *__restrict__ key, NrFit *fit) {
    __shared__ uint32_t hist[258], part[FT + 1];
    const NrRegion R = reg[blockIdx.x];
    NrFit &F = fit[blockIdx.x];
    const uint32_t n = F.n;
    if (n < 32) return;
    uint32_t nb, per;
    fit_bins(n, nb, per);
    const uint32_t b = blockIdx.y % NR_BINS;
    const bool med = blockIdx.y >= NR_BINS;
    if (b >= nb) return;
    if (blockIdx.y == 0 && threadIdx.x == 0) {
        F.nbins = nb;
        F.per = per;
    }
    const uint32_t len = min(per, n - b * per);
    const uint32_t *kr = key + R.first;
    uint32_t rest;
    const uint32_t K = select_rank([&](uint32_t i) { return kr[i]; }, [](uint32_t) { return true; }, R.count,
                                   b * per + (med ? len / 2 : 0), hist, &rest);
    // the rest-th block, in block order, whose key is K
    const uint32_t chunk = (R.count + FT - 1) / FT, lo = min(threadIdx.x * chunk, R.count);
    const uint32_t hi = min(lo + chunk, R.count);
    uint32_t c = 0;
    for (uint32_t i = lo; i < hi; i++) c += kr[i] == K ? 1u : 0u;
    part[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int t = 0; t < FT; t++) {
            const uint32_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    const uint32_t before = part[threadIdx.x];
    if (rest >= before && rest < before + c) {
        uint32_t m = rest - before;
        for (uint32_t i = lo; i < hi; i++) {
            if (kr[i] != K) continue;
            if (m == 0) {
                if (med)
                    F.mean[b] = stats[(size_t)(R.first + i) * 64];
                else
                    F.start[b] = i;
                break;
            }
            m--;
        }
    }
}

// Bin (r, b): its blocks, their median activity, and the blocks at most that (the quieter half)
// into flat[first + b * per ...].
__global__ __launch_bounds__(FT) void k_fit_flat(const NrRegion *__restrict__ reg, const uint32_t *__restrict__ key,
                                                 const float *__restrict__ act, uint32_t *__restrict__ flat,
                                                 NrFit *fit) {
    __shared__ uint32_t hist[258], lidx[NR_FIT_LIST], lkey[NR_FIT_LIST], cnt[2];
    const NrRegion R = reg[blockIdx.x];
    NrFit &F = fit[blockIdx.x];
    const uint32_t n = F.n;
    if (n < 32) return;
    uint32_t nb, per;
    fit_bins(n, nb, per);
    const uint32_t b = blockIdx.y;
    if (b >= nb) return;
    const uint32_t *kr = key + R.first;
    const float *ar = act + R.first;
    const uint32_t i0 = b > 0 ? F.start[b] : 0u, k0 = b > 0 ? kr[i0] : 0u;
    const uint32_t i1 = b + 1 < nb ? F.start[b + 1] : 0u, k1 = b + 1 < nb ? kr[i1] : 0u;
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < R.count; j += FT) {
        const uint32_t k = kr[j];
        const bool in = k != 0xffffffffu && (b == 0 || k > k0 || (k == k0 && j >= i0)) &&
                        (b + 1 == nb || k < k1 || (k == k1 && j < i1));
        if (in) {
            const uint32_t m = atomicAdd(&cnt[0], 1u);
            if (m < NR_FIT_LIST) {
                lidx[m] = j;
                lkey[m] = order_key(ar[j]);
            }
        }
    }
    __syncthreads();
    const uint32_t len = min(cnt[0], NR_FIT_LIST);
    uint32_t rest;
    const float cut = order_float(
        select_rank([&](uint32_t i) { return lkey[i]; }, [](uint32_t) { return true; }, len, len / 2, hist, &rest));
    uint32_t *out = flat + R.first + b * per;
    for (uint32_t m = threadIdx.x; m < len; m += FT) {
        const uint32_t j = lidx[m];
        if (ar[j] <= cut) out[atomicAdd(&cnt[1], 1u)] = j;
    }
    __syncthreads();
    if (threadIdx.x == 0) F.nflat[b] = cnt[1];
}

// (r, b, k): the median of squared coefficient k over bin b's quieter blocks.
__global__ __launch_bounds__(FT) void k_fit_sigma(const float *__restrict__ stats, const NrRegion *__restrict__ reg,
                                                  const uint32_t *__restrict__ flat, NrFit *fit) {
    __shared__ uint32_t hist[258], v[NR_FIT_LIST];
    const NrRegion R = reg[blockIdx.x];
    NrFit &F = fit[blockIdx.x];
    const uint32_t n = F.n;
    if (n < 32) return;
    uint32_t nb, per;
    fit_bins(n, nb, per);
    const uint32_t b = blockIdx.y, k = blockIdx.z + 1;
    if (b >= nb) return;
    const uint32_t nf = min(F.nflat[b], NR_FIT_LIST);
    const uint32_t *fl = flat + R.first + b * per;
    for (uint32_t m = threadIdx.x; m < nf; m += FT) v[m] = order_key(stats[(size_t)(R.first + fl[m]) * 64 + k]);
    __syncthreads();
    uint32_t rest;
    const uint32_t sel =
        select_rank([&](uint32_t i) { return v[i]; }, [](uint32_t) { return true; }, nf, nf / 2, hist, &rest);
    if (threadIdx.x == 0) F.sig[b][k] = (int32_t)(sel ^ 0x80000000u);
}

}  // namespace

hipError_t nr_split(const float *rgb, uint32_t w, uint32_t h, float *y, float *u, float *v, hipStream_t st) {
    const uint64_t n = (uint64_t)w * h;
    hipLaunchKernelGGL(k_split, dim3(grid1(n)), dim3(256), 0, st, rgb, n, y, u, v);
    return hipGetLastError();
}

hipError_t nr_down(const float *src, uint32_t w, uint32_t h, float *dst, const float *minus, hipStream_t st) {
    hipLaunchKernelGGL(k_down, grid2((w + 1) / 2, (h + 1) / 2), dim3(256), 0, st, src, w, h, dst, minus);
    return hipGetLastError();
}

hipError_t nr_add_up(float *d, uint32_t w, uint32_t h, const float *coarse, hipStream_t st) {
    hipLaunchKernelGGL(k_add_up, grid2(w, h), dim3(256), 0, st, d, w, h, coarse);
    return hipGetLastError();
}

hipError_t nr_blend(float *d, const float *orig, float keep, uint64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_blend, dim3(grid1(n)), dim3(256), 0, st, d, orig, keep, n);
    return hipGetLastError();
}

hipError_t nr_mul(const float *a, const float *b, float *out, uint64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_mul, dim3(grid1(n)), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

hipError_t nr_box(const float *a, uint32_t w, uint32_t h, uint32_t r, float *tmp, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_box_v, grid2(w, h), dim3(256), 0, st, a, w, h, r, tmp);
    const float n = (float)((2 * r + 1) * (2 * r + 1));
    hipLaunchKernelGGL(k_box_h, grid2(w, h), dim3(256), 0, st, tmp, w, h, r, n, out);
    return hipGetLastError();
}

hipError_t nr_den(float *var, const float *mean, const NrVar *nv, uint64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_den, dim3(grid1(n)), dim3(256), 0, st, var, mean, nv, n);
    return hipGetLastError();
}

hipError_t nr_gf_coef(const float *cov, const float *mp, const float *mean, const float *den, float *a, float *b,
                      uint64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_gf_coef, dim3(grid1(n)), dim3(256), 0, st, cov, mp, mean, den, a, b, n);
    return hipGetLastError();
}

hipError_t nr_gf_apply(const float *ma, const float *mb, const float *g, uint32_t w, uint32_t h, float *out,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_gf_apply, grid2(w, h), dim3(256), 0, st, ma, mb, g, w, h, out);
    return hipGetLastError();
}

hipError_t nr_stats(const float *p, const float *g, uint32_t w, uint32_t step, uint32_t cols, uint32_t rows,
                    const float *dct, float *out, hipStream_t st) {
    const uint64_t n = (uint64_t)cols * rows;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_stats, dim3((uint32_t)((n + 63) / 64)), dim3(64), 0, st, p, g, w, step, cols, rows, dct, out);
    return hipGetLastError();
}

hipError_t nr_fit(const float *stats, uint32_t total, const NrRegion *reg, uint32_t nreg, uint32_t *key, float *act,
                  uint32_t *flat, NrFit *fit, hipStream_t st) {
    if (nreg == 0 || total == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(fit, 0, nreg * sizeof(NrFit), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fit_keys, dim3((total + 63) / 64), dim3(256), 0, st, stats, total, reg, nreg, key, act, fit);
    hipLaunchKernelGGL(k_fit_bounds, dim3(nreg, 2 * NR_BINS), dim3(FT), 0, st, stats, reg, key, fit);
    hipLaunchKernelGGL(k_fit_flat, dim3(nreg, NR_BINS), dim3(FT), 0, st, reg, key, act, flat, fit);
    hipLaunchKernelGGL(k_fit_sigma, dim3(nreg, NR_BINS, 63), dim3(FT), 0, st, stats, reg, flat, fit);
    return hipGetLastError();
}

hipError_t nr_passes(const NrPass *d_list, uint32_t n, uint32_t tiles, uint32_t cell_blocks, uint32_t chunks,
                     bool wiener, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (wiener)
        hipLaunchKernelGGL(k_pass<true>, dim3(tiles), dim3(64), 0, st, d_list, n);
    else
        hipLaunchKernelGGL(k_pass<false>, dim3(tiles), dim3(64), 0, st, d_list, n);
    hipLaunchKernelGGL(k_wsum, dim3(cell_blocks), dim3(256), 0, st, d_list, n);
    hipLaunchKernelGGL(k_gather, dim3(chunks), dim3(256), 0, st, d_list, n);
    return hipGetLastError();
}

hipError_t nr_encode(const DevArgs &a, const float *y, const float *u, const float *v, hipStream_t st) {
    hipLaunchKernelGGL(k_encode, grid2(a.dev_w, a.dev_h), dim3(256), 0, st, a, y, u, v);
    return hipGetLastError();
}

}  // namespace frh
