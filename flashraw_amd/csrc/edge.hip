// Edge-aware demosaic (Bayer) fused with the finish and the encode (gfx950).
//
// The method of crates/flashraw/src/edgeaware/mod.rs:423-822 with the reference's f32 operations
// in the reference's order: a 7x7 non-directional green estimate G0 per canonical position, the
// square root of it, four directional gradient energies summed over 3x3 giving the weights
// w_d = 1 / (E_d + 1e-3)^2 (normalised), one 7x7 filter per position, colour and direction for
// the missing colours, two gated passes over the colour differences and, when beta > 0, a gated
// smoothing. Taps whose coefficient is zero (in every direction, for the directional filters)
// are skipped, as the reference's tap lists skip them (mod.rs:130-160).
//
// One block develops a 32 x 16 output tile. Each stage writes its plane (with the halo the next
// stage needs) to LDS: balanced mosaic (+-8), sqrt(G0) (+-5), squared gradients (+-4), weights
// and the three colour planes (+-3), pass-1 differences (+-2), pass-2 colours (+-1). The filter
// stages run phase by phase so that every wave walks one tap list; the tap lists sit in the
// constant address space and are read with scalar loads. Pixels nearer the frame border than the
// halo are developed bilinearly by extra blocks of the same launch (strip_to,
// process.rs:1160-1225).
#include <hip/hip_runtime.h>

#include "develop_common.h"
#include "frh_internal.h"
#include "frh_launch.h"

namespace frh {
namespace {

constexpr int ETW = 32, ETH = 16;      // output tile
constexpr int EH = 8;                  // halo of the mosaic plane
constexpr int MW = ETW + 2 * EH, MHt = ETH + 2 * EH;   // mosaic plane 48 x 32
constexpr int SW = ETW + 10, SH = ETH + 10;            // sqrt(G0) +-5
constexpr int DW = ETW + 8, DH = ETH + 8;              // gradients +-4
constexpr int PW = ETW + 6, PH = ETH + 6;              // weights, colour planes +-3
constexpr int QW = ETW + 4, QH = ETH + 4;              // pass-1 differences +-2
constexpr int VW = ETW + 2, VH = ETH + 2;              // pass-2 colours +-1

typedef __attribute__((address_space(4))) const EdgeTab ctab_t;

struct EdgeLds {
    float m[MHt * MW];
    union {
        float sg[SH * SW];
        float d1[2][QH * QW];
    } r2;
    union {
        float dsq[4][DH * DW];
        float v[3][VH * VW];
    } r3;
    float w[4][PH * PW];
    float p[3][PH * PW];
};

__device__ __forceinline__ int edge_phase(const DevArgs &a, int x, int y) {
    return (int)((((uint32_t)y + a.ecy) & 1u) * 2u + (((uint32_t)x + a.ecx) & 1u));
}
// Colour (0 R, 1 G, 2 B) of canonical position ph.
__device__ __forceinline__ uint32_t canon_of(const DevArgs &a, int ph) { return (a.ecanon_bits >> (2 * ph)) & 3u; }

// Pixels of canonical phase k in a plane of rw x rh starting at frame (x0, y0): the first row
// and column of that phase and their counts.
__device__ __forceinline__ void phase_grid(const DevArgs &a, int k, int x0, int y0, int rw, int rh, int &fr, int &fc,
                                           int &nr, int &nc) {
    fr = (int)((((uint32_t)(k >> 1)) - ((uint32_t)y0 + a.ecy)) & 1u);
    fc = (int)((((uint32_t)(k & 1)) - ((uint32_t)x0 + a.ecx)) & 1u);
    nr = rh > fr ? (rh - fr + 1) / 2 : 0;
    nc = rw > fc ? (rw - fc + 1) / 2 : 0;
}

// 6x6 mosaics (X-Trans): canonical position of frame sample (x, y), x, y >= 0.
__device__ __forceinline__ int edge_phase6(const DevArgs &a, int x, int y) {
    return (int)((((uint32_t)y + a.ecy) % 6u) * 6u + ((uint32_t)x + a.ecx) % 6u);
}

// Colour of frame sample (x, y): Bayer from the packed canonical colours, 6x6 from the table.
template <int P>
__device__ __forceinline__ uint32_t colour_at(const DevArgs &a, const ctab_t *T, int x, int y) {
    if constexpr (P == 2)
        return canon_of(a, edge_phase(a, x, y));
    else
        return T->canon[edge_phase6(a, x, y)];
}

// Phase-major enumeration of a plane of RW x RH starting at frame (x0, y0) for period 6: item i is
// pixel j of canonical position k = i / (NX * NY), NX x NY being the most columns and rows one
// position has in the plane; returns false for the items beyond a position's own count. A wave
// then walks at most two positions' tap lists.
template <int RW, int RH>
struct Phase6 {
    static constexpr int NX = (RW + 5) / 6, NY = (RH + 5) / 6, ITEMS = 36 * NX * NY;
    __device__ static __forceinline__ bool at(const DevArgs &a, int i, int x0, int y0, int &k, int &ry, int &rx) {
        k = i / (NX * NY);
        const int j = i - k * (NX * NY);
        const int fr = (int)((uint32_t)(k / 6) + 6u - ((uint32_t)y0 + a.ecy) % 6u) % 6;
        const int fc = (int)((uint32_t)(k % 6) + 6u - ((uint32_t)x0 + a.ecx) % 6u) % 6;
        ry = fr + 6 * (j / NX);
        rx = fc + 6 * (j % NX);
        return ry < RH && rx < RW;
    }
};

// Median of three (med3, edgeaware/mod.rs:857-860).
__device__ __forceinline__ float med3(float a, float b, float c) { return fmaxf(fminf(a, b), fminf(fmaxf(a, b), c)); }

// Median of the 3x3 window of differences centred on (y, x) of the colour planes, by the column
// scheme of median_row (edgeaware/mod.rs:825-855): each column of three sorted, then the median of
// (largest column minimum, median of the column medians, smallest column maximum). A median is
// one of the inputs, so no rounding is involved.
__device__ __forceinline__ float median_diff(const float *pc, const float *pg, int o, int pw) {
    float lo[3], mi[3], hi[3];
#pragma unroll
    for (int dx = 0; dx < 3; dx++) {
        const int q = o - pw - 1 + dx;
        const float a0 = pc[q] - pg[q], b0 = pc[q + pw] - pg[q + pw], c0 = pc[q + 2 * pw] - pg[q + 2 * pw];
        const float x1 = fminf(a0, b0), y1 = fmaxf(a0, b0);
        const float y2 = fminf(y1, c0), z2 = fmaxf(y1, c0);
        lo[dx] = fminf(x1, y2);
        mi[dx] = fmaxf(x1, y2);
        hi[dx] = z2;
    }
    const float l = fmaxf(fmaxf(lo[0], lo[1]), lo[2]);
    const float h = fminf(fminf(hi[0], hi[1]), hi[2]);
    return med3(l, med3(mi[0], mi[1], mi[2]), h);
}

// Border pixel i of a 6x6 mosaic: the generic CFA mean (process.rs:1098-1126).
__device__ __forceinline__ void border_pixel_cfa(const DevArgs &a, const Lds &s, const Border &B, uint32_t i) {
    int dx, dy;
    border_xy(a, B, i, dx, dy);
    float r, g, b;
    cfa_mean_rgb(a, a.crop_x + (uint32_t)dx, a.crop_y + (uint32_t)dy, r, g, b);
    float o0, o1, o2;
    finish(a, r, g, b, false, o0, o1, o2);
    store_both(a, s, (uint32_t)dx, (uint32_t)dy, o0, o1, o2);
}

template <int P>
__global__ __launch_bounds__(256) void k_develop_edge(DevArgs args, Border B, int tiles_x, uint32_t tiles) {
    const DevArgs &a = kernel_args(args);
    __shared__ Lds s;
    __shared__ EdgeLds E;
    load_lds(a, s);
    if (blockIdx.x >= tiles) {
        __syncthreads();
        const uint32_t i = (blockIdx.x - tiles) * blockDim.x + threadIdx.x;
        if (i < B.count) {
            if constexpr (P == 2)
                border_pixel(a, s, B, i);
            else
                border_pixel_cfa(a, s, B, i);
        }
        return;
    }
    const ctab_t *T = (const ctab_t *)(uintptr_t)a.etab;
    const int tid = threadIdx.x;
    // interior of this band in frame coordinates
    const int ix0 = B.ix0 + (int)a.crop_x, ix1 = B.ix1 + (int)a.crop_x;
    const int iy0 = B.iy0 + (int)a.crop_y, iy1 = B.iy1 + (int)a.crop_y;
    const int X0 = ix0 + (int)(blockIdx.x % (uint32_t)tiles_x) * ETW;
    const int Y0 = iy0 + (int)(blockIdx.x / (uint32_t)tiles_x) * ETH;
    const int W = (int)a.width, H = (int)a.height;
    __syncthreads();
    // 1. balanced mosaic min(norm * wb[colour], 1) (mod.rs:627-640), clamped reads beyond the frame
    for (int i = tid; i < MHt * MW; i += 256) {
        const int y = min(max(Y0 - EH + i / MW, 0), H - 1), x = min(max(X0 - EH + i % MW, 0), W - 1);
        const uint32_t c = colour_at<P>(a, T, x, y);
        const float wbv = c == 0 ? a.wb[0] : (c == 1 ? a.wb[1] : a.wb[2]);
        E.m[i] = fminf(norm_at(a, s, (uint32_t)x, (uint32_t)y) * wbv, 1.0f);
    }
    __syncthreads();
    // 2. sqrt(G0) over +-5, phase by phase (mod.rs:667-694)
    auto g0_px = [&](int k, int ry, int rx) {  // (ry, rx) in the sqrt(G0) plane
        const float *mp = &E.m[ry * MW + rx];  // tap (0, 0) of the 7x7 window: frame (x - 3, y - 3)
        const int n = T->g0_n[k];
        float acc = 0.0f;
        for (int t = 0; t < n; t++) {
            const uint32_t idx = T->g0_idx[k][t];
            acc += T->g0_c[k][t] * mp[(idx / 7u) * MW + idx % 7u];
        }
        E.r2.sg[ry * SW + rx] = sqrt_max0(acc);
    };
    if constexpr (P == 2) {
        for (int k = 0; k < 4; k++) {
            int fr, fc, nr, nc;
            phase_grid(a, k, X0 - 5, Y0 - 5, SW, SH, fr, fc, nr, nc);
            for (int i = tid; i < nr * nc; i += 256) g0_px(k, fr + 2 * (i / nc), fc + 2 * (i % nc));
        }
    } else {
        using G6 = Phase6<SW, SH>;
        for (int i = tid; i < G6::ITEMS; i += 256) {
            int k, ry, rx;
            if (G6::at(a, i, X0 - 5, Y0 - 5, k, ry, rx)) g0_px(k, ry, rx);
        }
    }
    __syncthreads();
    // 3. squared gradients of sqrt(G0) over +-4 (mod.rs:698-723)
    const float r2 = 0.70710677f;
    for (int i = tid; i < DH * DW; i += 256) {
        const int y = i / DW + 1, x = i % DW + 1;  // in the sqrt(G0) plane
        const float *g = E.r2.sg;
        const float dh = g[y * SW + x + 1] - g[y * SW + x - 1];
        const float dv = g[(y + 1) * SW + x] - g[(y - 1) * SW + x];
        const float e1 = (g[(y + 1) * SW + x + 1] - g[(y - 1) * SW + x - 1]) * r2;
        const float e2 = (g[(y + 1) * SW + x - 1] - g[(y - 1) * SW + x + 1]) * r2;
        E.r3.dsq[0][i] = dh * dh;
        E.r3.dsq[1][i] = dv * dv;
        E.r3.dsq[2][i] = e1 * e1;
        E.r3.dsq[3][i] = e2 * e2;
    }
    __syncthreads();
    // 4. direction weights over +-3: 3x3 sums, 1 / (E + eps)^2, normalised (mod.rs:724-759)
    for (int i = tid; i < PH * PW; i += 256) {
        const int y = i / PW + 1, x = i % PW + 1;  // in the gradient plane
        float w[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const float *q = E.r3.dsq[d];
            const float h0 = q[(y - 1) * DW + x - 1] + q[(y - 1) * DW + x] + q[(y - 1) * DW + x + 1];
            const float h1 = q[y * DW + x - 1] + q[y * DW + x] + q[y * DW + x + 1];
            const float h2 = q[(y + 1) * DW + x - 1] + q[(y + 1) * DW + x] + q[(y + 1) * DW + x + 1];
            const float u = h0 + h1 + h2 + 1e-3f;
            w[d] = 1.0f / (u * u);
        }
        const float inv = 1.0f / (w[0] + w[1] + w[2] + w[3]);
#pragma unroll
        for (int d = 0; d < 4; d++) E.w[d][i] = w[d] * inv;
    }
    __syncthreads();
    // 5. the colour planes over +-3: the own colour is the sample, the missing ones the weighted
    //    directional filters (mod.rs:762-795), phase by phase
    auto colour_px = [&](int k, int ry, int rx) {  // (ry, rx) in the colour planes
        const uint32_t own = P == 2 ? canon_of(a, k) : T->canon[k];
        const float *mp = &E.m[(ry + 5 - 3) * MW + (rx + 5 - 3)];
        const int o = ry * PW + rx;
        const float w0 = E.w[0][o], w1 = E.w[1][o], w2 = E.w[2][o], w3 = E.w[3][o];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if ((uint32_t)c == own) {
                E.p[c][o] = mp[3 * MW + 3];
                continue;
            }
            float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
            const int n = T->dir_n[k][c];
            for (int t = 0; t < n; t++) {
                const uint32_t idx = T->dir_idx[k][c][t];
                const float v = mp[(idx / 7u) * MW + idx % 7u];
                acc0 += T->dir_c[k][c][t][0] * v;
                acc1 += T->dir_c[k][c][t][1] * v;
                acc2 += T->dir_c[k][c][t][2] * v;
                acc3 += T->dir_c[k][c][t][3] * v;
            }
            E.p[c][o] = w0 * acc0 + w1 * acc1 + w2 * acc2 + w3 * acc3;
        }
    };
    if constexpr (P == 2) {
        for (int k = 0; k < 4; k++) {
            int fr, fc, nr, nc;
            phase_grid(a, k, X0 - 3, Y0 - 3, PW, PH, fr, fc, nr, nc);
            for (int i = tid; i < nr * nc; i += 256) colour_px(k, fr + 2 * (i / nc), fc + 2 * (i % nc));
        }
    } else {
        using G6 = Phase6<PW, PH>;
        for (int i = tid; i < G6::ITEMS; i += 256) {
            int k, ry, rx;
            if (G6::at(a, i, X0 - 3, Y0 - 3, k, ry, rx)) colour_px(k, ry, rx);
        }
    }
    __syncthreads();
    // 6. pass 1 over +-2: Bayer, the gated mean of the measured differences of the 3x3 window
    //    (mod.rs:476-530); 6x6, the median of the differences of the filtered planes over the 3x3
    //    window (mod.rs:455-474), the measured difference at the colour's own sites
    for (int i = tid; i < QH * QW; i += 256) {
        const int y = i / QW + 1, x = i % QW + 1;  // in the colour planes
        const int fy = Y0 - 3 + y, fx = X0 - 3 + x;
        if constexpr (P == 6) {
            const uint32_t own = colour_at<P>(a, T, fx, fy);
            const int o = y * PW + x;
            E.r2.d1[0][i] = own == 0 ? E.p[0][o] - E.p[1][o] : median_diff(E.p[0], E.p[1], o, PW);
            E.r2.d1[1][i] = own == 2 ? E.p[2][o] - E.p[1][o] : median_diff(E.p[2], E.p[1], o, PW);
            continue;
        }
        const float g = E.p[1][y * PW + x];
        const float inv = 1.0f / (0.2f * fmaxf(g, 0.0f) + 0.015f);
        float sw0 = 0.0f, sd0 = 0.0f, su0 = 0.0f, n0 = 0.0f, sw1 = 0.0f, sd1 = 0.0f, su1 = 0.0f, n1 = 0.0f;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int q = (y + dy) * PW + x + dx;
                const float gk = E.p[1][q];
                const float e = (gk - g) * inv;
                const float wt = 1.0f / (1.0f + e * e);
                const uint32_t col = colour_at<P>(a, T, fx + dx, fy + dy);
                const float m0 = col == 0 ? 1.0f : 0.0f, m1 = col == 2 ? 1.0f : 0.0f;
                const float d0 = E.p[0][q] - gk, d2 = E.p[2][q] - gk;
                sw0 += m0 * wt;
                sd0 += (m0 * wt) * d0;
                su0 += m0 * d0;
                n0 += m0;
                sw1 += m1 * wt;
                sd1 += (m1 * wt) * d2;
                su1 += m1 * d2;
                n1 += m1;
            }
        const uint32_t own = colour_at<P>(a, T, fx, fy);
        const float mixed0 = sw0 > 1e-12f ? sd0 / sw0 : su0 / n0;
        const float mixed1 = sw1 > 1e-12f ? sd1 / sw1 : su1 / n1;
        E.r2.d1[0][i] = own == 0 ? E.p[0][y * PW + x] - g : mixed0;
        E.r2.d1[1][i] = own == 2 ? E.p[2][y * PW + x] - g : mixed1;
    }
    __syncthreads();
    // 7. pass 2 over +-1: the same gate over every pixel's difference (mod.rs:534-598)
    for (int i = tid; i < VH * VW; i += 256) {
        const int y = i / VW + 1, x = i % VW + 1;  // in the pass-1 plane
        const int py = y + 1, px = x + 1;          // in the colour planes
        const float g = E.p[1][py * PW + px];
        const float inv = 1.0f / (0.1f * fmaxf(g, 0.0f) + 0.007f);
        float sw = 0.0f, sdr = 0.0f, sdb = 0.0f;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const float gk = E.p[1][(py + dy) * PW + px + dx];
                const float e = (gk - g) * inv;
                const float wt = 1.0f / (1.0f + e * e);
                const int q = (y + dy) * QW + x + dx;
                sw += wt;
                sdr += wt * E.r2.d1[0][q];
                sdb += wt * E.r2.d1[1][q];
            }
        const float dr = sdr / sw, db = sdb / sw;
        const uint32_t own = colour_at<P>(a, T, X0 - 1 + x - 1, Y0 - 1 + y - 1);
        const float p0 = E.p[0][py * PW + px], p2 = E.p[2][py * PW + px];
        // 6x6: green at a red or blue site moves towards the measured sample minus the cleaned
        // difference by xg = clamp(1 - 2 beta, 0, 1) (mod.rs:547-592)
        float gv = g;
        if (P == 6 && a.exg != 0.0f && own == 0) gv = g + a.exg * ((p0 - dr) - g);
        if (P == 6 && a.exg != 0.0f && own == 2) gv = g + a.exg * ((p2 - db) - g);
        E.r3.v[0][i] = own == 0 ? p0 : gv + dr;
        E.r3.v[1][i] = gv;
        E.r3.v[2][i] = own == 2 ? p2 : gv + db;
    }
    __syncthreads();
    // 8. gated smoothing (beta > 0, mod.rs:803-822), finish with unit balance, encode, store
    const int xe = min(X0 + ETW, ix1), ye = min(Y0 + ETH, iy1);
    for (int i = tid; i < ETW * ETH; i += 256) {
        const int ty = i / ETW, tx = i % ETW;
        const int fx = X0 + tx, fy = Y0 + ty;
        if (fx >= xe || fy >= ye) continue;
        const int c = (ty + 1) * VW + tx + 1;
        float r = E.r3.v[0][c], g = E.r3.v[1][c], b = E.r3.v[2][c];
        if (a.ebeta > 0.0f) {
            const float inv = 1.0f / (0.2f * fmaxf(g, 0.0f) + 0.015f);
            float sw = 0.0f, s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
            for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const int q = c + dy * VW + dx;
                    const float e = (E.r3.v[1][q] - g) * inv;
                    const float wt = 1.0f / (1.0f + e * e);
                    sw += wt;
                    s0 += wt * E.r3.v[0][q];
                    s1 += wt * E.r3.v[1][q];
                    s2 += wt * E.r3.v[2][q];
                }
            r = r + a.ebeta * (s0 / sw - r);
            const float g2 = g + a.ebeta * (s1 / sw - g);
            b = b + a.ebeta * (s2 / sw - b);
            g = g2;
        }
        float o0, o1, o2;
        finish(a, r, g, b, true, o0, o1, o2);
        store_both(a, s, (uint32_t)(fx - (int)a.crop_x), (uint32_t)(fy - (int)a.crop_y), o0, o1, o2);
    }
}

// ---- the kernel for tame inputs (edge_tame) ----
//
// Same stages and operations, arranged for throughput: every filter is a fully unrolled 7x7
// window held in registers with the coefficients read as scalars from the dense tables (a zero tap
// multiplies a finite value and adds +-0, which changes no sum), the direction weights are
// computed where the colour filters need them (no weight planes), and pass 1 reads the colours of
// the 3x3 neighbours from the phase, which is uniform per loop, so that the zero terms of its
// masked sums are skipped. The tile is 58 x 26 so that each phase of the 64 x 32 colour-plane
// region holds exactly 512 pixels: one per thread.
template <int TW, int TH>
struct EdgeGeom {
    static constexpr int MW = TW + 16, MH = TH + 16;  // balanced mosaic, origin (X0 - 8, Y0 - 8)
    static constexpr int SW = TW + 10, SH = TH + 10;  // sqrt(G0), origin - 5
    static constexpr int DW = TW + 8, DH = TH + 8;    // squared gradients, origin - 4
    static constexpr int PW = TW + 6, PH = TH + 6;    // colour planes, origin - 3
    static constexpr int QW = TW + 4, QH = TH + 4;    // pass-1 differences, origin - 2
    static constexpr int VW = TW + 2, VH = TH + 2;    // pass-2 colours, origin - 1
};

template <int TW, int TH>
struct EdgeLds2 {
    using G = EdgeGeom<TW, TH>;
    union {
        float m[G::MH * G::MW];
        float d1[2][G::QH * G::QW];
    } a;
    union {
        float sg[G::SH * G::SW];
        float p[3][G::PH * G::PW];
    } b;
    union {
        float dsq[4][G::DH * G::DW];
        float v[3][G::VH * G::VW];
    } c;
};

template <int TW, int TH, int NT>
__global__ __launch_bounds__(NT) void k_develop_edge2(DevArgs args, Border B, int tiles_x, uint32_t tiles) {
    const DevArgs &a = kernel_args(args);
    using G = EdgeGeom<TW, TH>;
    __shared__ Lds s;
    __shared__ EdgeLds2<TW, TH> E;
    load_lds(a, s);
    if (blockIdx.x >= tiles) {
        __syncthreads();
        const uint32_t i = (blockIdx.x - tiles) * blockDim.x + threadIdx.x;
        if (i < B.count) border_pixel(a, s, B, i);
        return;
    }
    const ctab_t *T = (const ctab_t *)(uintptr_t)a.etab;
    const int tid = threadIdx.x;
    const int ix0 = B.ix0 + (int)a.crop_x, ix1 = B.ix1 + (int)a.crop_x;
    const int iy0 = B.iy0 + (int)a.crop_y, iy1 = B.iy1 + (int)a.crop_y;
    const int X0 = ix0 + (int)(blockIdx.x % (uint32_t)tiles_x) * TW;
    const int Y0 = iy0 + (int)(blockIdx.x / (uint32_t)tiles_x) * TH;
    const int W = (int)a.width, H = (int)a.height;
    __syncthreads();
    // 1. balanced mosaic min(norm * wb[colour], 1), clamped reads beyond the frame
    for (int i = tid; i < G::MH * G::MW; i += NT) {
        const int y = min(max(Y0 - 8 + i / G::MW, 0), H - 1), x = min(max(X0 - 8 + i % G::MW, 0), W - 1);
        const uint32_t c = canon_of(a, edge_phase(a, x, y));
        const float wbv = c == 0 ? a.wb[0] : (c == 1 ? a.wb[1] : a.wb[2]);
        E.a.m[i] = fminf(norm_at(a, s, (uint32_t)x, (uint32_t)y) * wbv, 1.0f);
    }
    __syncthreads();
    // 2. sqrt(G0) over +-5, phase by phase: the dense 7x7 window, or the few taps of an
    //    identity-like filter
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        int fr, fc, nr, nc;
        phase_grid(a, k, X0 - 5, Y0 - 5, G::SW, G::SH, fr, fc, nr, nc);
        const int n = T->g0_n[k];
        if (n > 8) {
            for (int i = tid; i < nr * nc; i += NT) {
                const int ry = fr + 2 * (i / nc), rx = fc + 2 * (i % nc);
                const float *mp = &E.a.m[ry * G::MW + rx];
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < 49; t++) acc += T->g0d[k][t] * mp[(t / 7) * G::MW + t % 7];
                E.b.sg[ry * G::SW + rx] = sqrt_max0(acc);
            }
        } else {
            for (int i = tid; i < nr * nc; i += NT) {
                const int ry = fr + 2 * (i / nc), rx = fc + 2 * (i % nc);
                const float *mp = &E.a.m[ry * G::MW + rx];
                float acc = 0.0f;
                for (int t = 0; t < n; t++) {
                    const uint32_t idx = T->g0_idx[k][t];
                    acc += T->g0_c[k][t] * mp[(idx / 7u) * G::MW + idx % 7u];
                }
                E.b.sg[ry * G::SW + rx] = sqrt_max0(acc);
            }
        }
    }
    __syncthreads();
    // 3. squared gradients of sqrt(G0) over +-4
    const float r2 = 0.70710677f;
    for (int i = tid; i < G::DH * G::DW; i += NT) {
        const int y = i / G::DW + 1, x = i % G::DW + 1;
        const float *g = E.b.sg;
        const float dh = g[y * G::SW + x + 1] - g[y * G::SW + x - 1];
        const float dv = g[(y + 1) * G::SW + x] - g[(y - 1) * G::SW + x];
        const float e1 = (g[(y + 1) * G::SW + x + 1] - g[(y - 1) * G::SW + x - 1]) * r2;
        const float e2 = (g[(y + 1) * G::SW + x - 1] - g[(y - 1) * G::SW + x + 1]) * r2;
        E.c.dsq[0][i] = dh * dh;
        E.c.dsq[1][i] = dv * dv;
        E.c.dsq[2][i] = e1 * e1;
        E.c.dsq[3][i] = e2 * e2;
    }
    __syncthreads();
    // 4-5. per pixel of the +-3 region, phase by phase: the direction weights from the 3x3 sums of
    //      the squared gradients, then the two missing colours from the 49 samples of the window
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        const uint32_t own = canon_of(a, k);
        int fr, fc, nr, nc;
        phase_grid(a, k, X0 - 3, Y0 - 3, G::PW, G::PH, fr, fc, nr, nc);
        for (int i = tid; i < nr * nc; i += NT) {
            const int ry = fr + 2 * (i / nc), rx = fc + 2 * (i % nc);
            float w[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const float *q = &E.c.dsq[d][ry * G::DW + rx];
                const float h0 = q[0] + q[1] + q[2];
                const float h1 = q[G::DW] + q[G::DW + 1] + q[G::DW + 2];
                const float h2 = q[2 * G::DW] + q[2 * G::DW + 1] + q[2 * G::DW + 2];
                const float u = h0 + h1 + h2 + 1e-3f;
                w[d] = 1.0f / (u * u);
            }
            const float inv = 1.0f / (w[0] + w[1] + w[2] + w[3]);
            const float w0 = w[0] * inv, w1 = w[1] * inv, w2 = w[2] * inv, w3 = w[3] * inv;
            const float *mp = &E.a.m[(ry + 2) * G::MW + rx + 2];
            float win[49];
#pragma unroll
            for (int t = 0; t < 49; t++) win[t] = mp[(t / 7) * G::MW + t % 7];
            const int o = ry * G::PW + rx;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if ((uint32_t)c == own) {
                    E.b.p[c][o] = win[24];
                    continue;
                }
                float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
#pragma unroll
                for (int t = 0; t < 49; t++) {
                    acc0 += T->dird[k][c][t][0] * win[t];
                    acc1 += T->dird[k][c][t][1] * win[t];
                    acc2 += T->dird[k][c][t][2] * win[t];
                    acc3 += T->dird[k][c][t][3] * win[t];
                }
                E.b.p[c][o] = w0 * acc0 + w1 * acc1 + w2 * acc2 + w3 * acc3;
            }
        }
    }
    __syncthreads();
    // 6. pass 1 over +-2, phase by phase: the neighbours' colours are uniform, and a neighbour of
    //    the other colours adds exact zeros to the masked sums (mod.rs:476-530)
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        const uint32_t own = canon_of(a, k);
        int fr, fc, nr, nc;
        phase_grid(a, k, X0 - 2, Y0 - 2, G::QW, G::QH, fr, fc, nr, nc);
        for (int i = tid; i < nr * nc; i += NT) {
            const int qy = fr + 2 * (i / nc), qx = fc + 2 * (i % nc);  // in the pass-1 plane
            const int y = qy + 1, x = qx + 1;                          // in the colour planes
            const float g = E.b.p[1][y * G::PW + x];
            const float inv = 1.0f / (0.2f * fmaxf(g, 0.0f) + 0.015f);
            float sw0 = 0.0f, sd0 = 0.0f, su0 = 0.0f, n0 = 0.0f, sw1 = 0.0f, sd1 = 0.0f, su1 = 0.0f, n1 = 0.0f;
#pragma unroll
            for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const uint32_t col = canon_of(a, k ^ (((dy & 1) << 1) | (dx & 1)));
                    if (col == 1) continue;
                    const int q = (y + dy) * G::PW + x + dx;
                    const float gk = E.b.p[1][q];
                    const float e = (gk - g) * inv;
                    const float wt = 1.0f / (1.0f + e * e);
                    if (col == 0) {
                        const float d0 = E.b.p[0][q] - gk;
                        sw0 += wt;
                        sd0 += wt * d0;
                        su0 += d0;
                        n0 += 1.0f;
                    } else {
                        const float d2 = E.b.p[2][q] - gk;
                        sw1 += wt;
                        sd1 += wt * d2;
                        su1 += d2;
                        n1 += 1.0f;
                    }
                }
            const int o = qy * G::QW + qx;
            float r0, r1;
            if (own == 0)
                r0 = E.b.p[0][y * G::PW + x] - g;
            else
                r0 = sw0 > 1e-12f ? sd0 / sw0 : su0 / n0;
            if (own == 2)
                r1 = E.b.p[2][y * G::PW + x] - g;
            else
                r1 = sw1 > 1e-12f ? sd1 / sw1 : su1 / n1;
            E.a.d1[0][o] = r0;
            E.a.d1[1][o] = r1;
        }
    }
    __syncthreads();
    // 7. pass 2 over +-1 (mod.rs:534-598)
    for (int i = tid; i < G::VH * G::VW; i += NT) {
        const int y = i / G::VW + 1, x = i % G::VW + 1;  // in the pass-1 plane
        const int py = y + 1, px = x + 1;                // in the colour planes
        const float g = E.b.p[1][py * G::PW + px];
        const float inv = 1.0f / (0.1f * fmaxf(g, 0.0f) + 0.007f);
        float sw = 0.0f, sdr = 0.0f, sdb = 0.0f;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const float gk = E.b.p[1][(py + dy) * G::PW + px + dx];
                const float e = (gk - g) * inv;
                const float wt = 1.0f / (1.0f + e * e);
                const int q = (y + dy) * G::QW + x + dx;
                sw += wt;
                sdr += wt * E.a.d1[0][q];
                sdb += wt * E.a.d1[1][q];
            }
        const float dr = sdr / sw, db = sdb / sw;
        const uint32_t own = canon_of(a, edge_phase(a, X0 - 1 + x - 1, Y0 - 1 + y - 1));
        E.c.v[0][i] = own == 0 ? E.b.p[0][py * G::PW + px] : g + dr;
        E.c.v[1][i] = g;
        E.c.v[2][i] = own == 2 ? E.b.p[2][py * G::PW + px] : g + db;
    }
    __syncthreads();
    // 8. gated smoothing (beta > 0, mod.rs:803-822), finish with unit balance, encode, store
    const int xe = min(X0 + TW, ix1), ye = min(Y0 + TH, iy1);
    for (int i = tid; i < TW * TH; i += NT) {
        const int ty = i / TW, tx = i % TW;
        const int fx = X0 + tx, fy = Y0 + ty;
        if (fx >= xe || fy >= ye) continue;
        const int c = (ty + 1) * G::VW + tx + 1;
        float r = E.c.v[0][c], g = E.c.v[1][c], b = E.c.v[2][c];
        if (a.ebeta > 0.0f) {
            const float inv = 1.0f / (0.2f * fmaxf(g, 0.0f) + 0.015f);
            float sw = 0.0f, s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
            for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                for (int dx = -1; dx <= 1; dx++) {
                    const int q = c + dy * G::VW + dx;
                    const float e = (E.c.v[1][q] - g) * inv;
                    const float wt = 1.0f / (1.0f + e * e);
                    sw += wt;
                    s0 += wt * E.c.v[0][q];
                    s1 += wt * E.c.v[1][q];
                    s2 += wt * E.c.v[2][q];
                }
            r = r + a.ebeta * (s0 / sw - r);
            const float g2 = g + a.ebeta * (s1 / sw - g);
            b = b + a.ebeta * (s2 / sw - b);
            g = g2;
        }
        float o0, o1, o2;
        finish(a, r, g, b, true, o0, o1, o2);
        store_both(a, s, (uint32_t)(fx - (int)a.crop_x), (uint32_t)(fy - (int)a.crop_y), o0, o1, o2);
    }
}

constexpr int E2W = 58, E2H = 26, E2T = 512;
static_assert(((E2W + 6) / 2) * ((E2H + 6) / 2) == E2T, "one colour-plane pixel per thread and phase");

}  // namespace

hipError_t launch_develop_edge(const DevArgs &a, hipStream_t stream) {
    const Border B = make_border(a, (int)a.ahd_halo);
    const int iw = B.ix1 - B.ix0, ih = B.iy1 - B.iy0;
    const int tw = a.edge_tame ? E2W : ETW, th = a.edge_tame ? E2H : ETH, nt = a.edge_tame ? E2T : 256;
    const int tiles_x = iw > 0 ? (iw + tw - 1) / tw : 0;
    const uint32_t tiles = (iw > 0 && ih > 0) ? (uint32_t)(tiles_x * ((ih + th - 1) / th)) : 0u;
    const uint32_t border_blocks = (B.count + nt - 1) / nt;
    const uint32_t grid = tiles + border_blocks;
    if (grid == 0) return hipSuccess;
    if (a.edge_tame)
        hipLaunchKernelGGL((k_develop_edge2<E2W, E2H, E2T>), dim3(grid), dim3(nt), 0, stream, a, B, tiles_x, tiles);
    else if (a.eperiod == 6)
        hipLaunchKernelGGL(k_develop_edge<6>, dim3(grid), dim3(nt), 0, stream, a, B, tiles_x, tiles);
    else
        hipLaunchKernelGGL(k_develop_edge<2>, dim3(grid), dim3(nt), 0, stream, a, B, tiles_x, tiles);
    return hipGetLastError();
}

}  // namespace frh
