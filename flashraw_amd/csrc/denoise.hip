// CFA-plane DCT block denoiser for gfx950.
//
// The mosaic is split into its four 2x2 phase planes (W/2 x H/2 each, black subtracted, in DN).
// 8x8 blocks at stride `step` (positions 0, step, ..., and always the last one, as the
// reference's positions(): crates/flashraw/src/denoise/mod.rs:877-884) go through the
// orthonormal DCT-II (denoise/mod.rs:626-635); AC coefficients with |c| < thr * sigma are
// zeroed, sigma^2 = sig_a * max(block mean, 0) + sig_b; the blocks are transformed back and
// overlap-added with uniform weights. The output is the mosaic with black added back, rounded
// and clamped to u16, and/or the denoised planes in f32 (DN above black).
//
// Two implementations:
//
// * MFMA walker (f16 or bf16 operands, f32 accumulation; optionally split f16 operands, each
//   value as a rounded f16 plus its rounded remainder, three MFMAs per product), stride 4. One 16x16x16 MFMA tile
//   holds four overlapping blocks: rows {y..y+7} and {y+4..y+11} against columns {x..x+7} and
//   {x+4..x+11}, so a tile at (8a, 8b) owns every block position of its 8x8 cell. The forward
//   transform is P * Dc^T then Dr * T, the inverse C^T * Dr then W^T * Dc: each product's
//   result, still in the accumulator layout, is the next product's operand, with no shuffles.
//   The M (row) index of the tile is ordered so that the rows two blocks share, and the rows a
//   tile shares with the tile below it, land in the same lane and differ only in the register;
//   the overlap-add down a column is then register arithmetic. Columns shared inside a tile and
//   with the tile to the left are lanes 4 and 12 apart inside one 16-lane row (DPP row shifts).
//   A wave walks down a strip of tile columns (plus one tile column of halo on the left and one
//   tile row above) and writes each 8x8 cell once, for both planes of its mosaic row parity.
//   Pixel values enter the MFMA relative to a per-tile offset (the tile's first pixel), so the
//   16-bit operands carry local contrast, not the absolute level; the offset is added back in
//   f32 after the inverse transform.
//
// * f32 scalar variant, any stride 1..8: one thread per block performs the oracle's float
//   arithmetic in the oracle's order (per-block values equal the oracle's f32 statement), writes
//   its 64 reconstructed samples to a scratch buffer, and a gather pass sums each pixel's blocks
//   in ascending block order and divides by their count.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dn_kron.h"
#include "frh_internal.h"

namespace frh {
namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int DN_NB = 4;  // owned tile columns per wave (plus one of halo)
constexpr int DN_WAVES = 4;  // waves per workgroup

struct DnArgs {
    const uint16_t *src;
    uint16_t *dst;  // may be null
    float *dst_f;   // may be null
    uint32_t stride;  // samples per mosaic row (source and destinations)
    uint32_t pw, ph;  // plane size
    int plx, ply;     // last block position in x and y (pw - 8, ph - 8)
    uint32_t nt, nr;  // tile columns and rows: ceil(pw / 8), ceil(ph / 8)
    uint32_t strips, segs, seg_rows;
    float black[4];
    float sig_a, sig_b, thr;
    float d[64];  // the 8x8 DCT-II matrix D[k][i], row k = frequency
};

// Row order of the 16-row tile: M index i = 4g + r. r < 2: the upper block (pixel rows y..y+7),
// r >= 2: the lower block (rows y+4..y+11); the position inside the block is base[g] + (r & 1).
__device__ __forceinline__ int tile_row_in_block(int i) {
    const int g = i >> 2, r = i & 3;
    const int b1 = (0x2640 >> (4 * g)) & 15;  // {0, 4, 6, 2}
    const int b2 = (0x6204 >> (4 * g)) & 15;  // {4, 0, 2, 6}
    return (r < 2 ? b1 : b2) + (r & 1);
}
__device__ __forceinline__ bool tile_row_lower(int i) { return (i & 3) >= 2; }

template <typename V>
__device__ __forceinline__ V to_op(f4 x);
template <>
__device__ __forceinline__ half4 to_op<half4>(f4 x) {
    half4 h;
    h[0] = (_Float16)x[0];
    h[1] = (_Float16)x[1];
    h[2] = (_Float16)x[2];
    h[3] = (_Float16)x[3];
    return h;
}
template <>
__device__ __forceinline__ bf16x4 to_op<bf16x4>(f4 x) {
    bf16x4 h;
    h[0] = (__bf16)x[0];
    h[1] = (__bf16)x[1];
    h[2] = (__bf16)x[2];
    h[3] = (__bf16)x[3];
    return h;
}
template <typename V>
__device__ __forceinline__ f4 from_op(V h) {
    return f4{(float)h[0], (float)h[1], (float)h[2], (float)h[3]};
}
__device__ __forceinline__ f4 mfma(half4 a, half4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma(bf16x4 a, bf16x4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// An operand as one 16-bit value (SPLIT false) or as the sum of two (SPLIT true: the rounded
// value and the rounded remainder), which carries about 22 significant bits.
template <typename V, bool SPLIT>
struct Op {
    V h, l;
};
template <typename V, bool SPLIT>
__device__ __forceinline__ Op<V, SPLIT> make_op(f4 x) {
    Op<V, SPLIT> o;
    o.h = to_op<V>(x);
    if (SPLIT) o.l = to_op<V>(x - from_op<V>(o.h));
    return o;
}
// a * b with f32 accumulation: one MFMA, or three for split operands (the low x low term is
// below f32 rounding of the result and is left out).
template <typename V, bool SPLIT>
__device__ __forceinline__ f4 product(const Op<V, SPLIT> &a, const Op<V, SPLIT> &b) {
    const f4 z = {0.f, 0.f, 0.f, 0.f};
    if (!SPLIT) return mfma(a.h, b.h, z);
    return mfma(a.h, b.h, mfma(a.h, b.l, mfma(a.l, b.h, z)));
}

// DPP inside 16-lane rows; lanes whose source is outside the row read 0.
__device__ __forceinline__ float row_shr4(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xF, 0xF, true));
}
__device__ __forceinline__ float row_shl12(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x10C, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_value(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// 1 / (blocks covering plane coordinate p) along an axis whose last position is `last` (stride 4).
__device__ __forceinline__ float inv_cover(int p, int last) { return (p < 4 || p > last + 3) ? 1.0f : 0.5f; }

// Operands enter the products scaled by 1/16 (exact), which keeps the largest DC coefficient
// (8 x 16 x 15871 before scaling) inside the f16 range.
constexpr float DN_SCALE = 0.0625f;

template <typename V, bool SPLIT>
__global__ __launch_bounds__(DN_WAVES * 64) void k_denoise_mfma(DnArgs a) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * DN_WAVES + (threadIdx.x >> 6));
    if (wave >= a.strips * a.segs * 2) return;
    const int dy = (int)(wave & 1);
    const uint32_t rest = wave >> 1;
    const int strip = (int)(rest % a.strips), seg = (int)(rest / a.strips);
    const int b0 = strip * DN_NB;
    const int a0 = seg * (int)a.seg_rows;
    const int a1 = min(a0 + (int)a.seg_rows, (int)a.nr);
    const int n16 = lane & 15, G = lane >> 4;

    // Operand fragments of the four constant matrices (lane layout of the 16x16x16 MFMA:
    // A[l % 16][4 (l / 16) + r], B[4 (l / 16) + r][l % 16], D[4 (l / 16) + r][l % 16]).
    f4 f1, f2, f3, f4m;
    for (int r = 0; r < 4; r++) {
        const int kk = 4 * G + r;
        // step 1, B = Dc^T: Dc = diag(D, D) over the column index
        f1[r] = ((n16 >> 3) == (kk >> 3)) ? a.d[(n16 & 7) * 8 + (kk & 7)] : 0.0f;
        // step 2, A = Dr: coefficient row u = n16, tile row i = kk
        f2[r] = ((n16 >> 3) == (tile_row_lower(kk) ? 1 : 0)) ? a.d[(n16 & 7) * 8 + tile_row_in_block(kk)] : 0.0f;
        // step 3, B = Dr: coefficient row u = kk, tile row i = n16
        f3[r] = ((kk >> 3) == (tile_row_lower(n16) ? 1 : 0)) ? a.d[(kk & 7) * 8 + tile_row_in_block(n16)] : 0.0f;
        // step 4, B = Dc: coefficient column v = kk, tile column k = n16
        f4m[r] = ((kk >> 3) == (n16 >> 3)) ? a.d[(kk & 7) * 8 + (n16 & 7)] : 0.0f;
    }
    const Op<V, SPLIT> F1 = make_op<V, SPLIT>(f1), F2 = make_op<V, SPLIT>(f2), F3 = make_op<V, SPLIT>(f3),
                       F4 = make_op<V, SPLIT>(f4m);

    // This lane's load in a tile: tile row n16, four columns starting at {0, 4, 4, 8}[G].
    const int ld_row = (tile_row_lower(n16) ? 4 : 0) + tile_row_in_block(n16);
    const int ld_col = (0x8440 >> (4 * G)) & 15;
    const float bl0 = a.black[2 * dy], bl1 = a.black[2 * dy + 1];
    const bool edge_group = (G == 0 || G == 3);  // rows shared with the tile below: registers 2, 3
    const int out_row = ((0x2640 >> (4 * G)) & 15);  // owned rows: out_row + 0, 1
    const bool out_lane = (n16 < 4) || (n16 >= 8 && n16 < 12);
    const int out_col = n16 < 4 ? n16 : n16 - 4;
    const bool is_dc_lane = ((n16 & 7) == 0) && ((G & 1) == 0);
    const int dc_src = lane & 0x28;

    float carry[DN_NB][2][2];
    for (int t = 0; t < DN_NB; t++)
        for (int p = 0; p < 2; p++) carry[t][p][0] = carry[t][p][1] = 0.0f;

    for (int ta = (a0 > 0 ? a0 - 1 : a0); ta < a1; ta++) {
        const int y = 8 * ta;
        const bool vr1 = y <= a.ply, vr2 = y + 4 <= a.ply;
        // All tiles of the row first, without branches, so that their MFMA chains interleave;
        // tiles outside the plane compute on clamped loads and are masked to zero.
        f4 X[DN_NB + 1][2];
#pragma unroll
        for (int t = 0; t <= DN_NB; t++) {
            const int b = b0 - 1 + t;
            const int x = 8 * b;
            const bool vc1 = b >= 0 && x <= a.plx, vc2 = b >= 0 && x + 4 <= a.plx;
            const int prow = min(y + ld_row, (int)a.ph - 1);
            const int pcol = max(min(x + ld_col, (int)a.pw - 4), 0);
            const uint4 w = *reinterpret_cast<const uint4 *>(a.src + (size_t)(2 * prow + dy) * a.stride + 2 * pcol);
            const uint32_t wv[4] = {w.x, w.y, w.z, w.w};
            const bool blk_ok = ((G >= 2) ? vr2 : vr1) && ((n16 >= 8) ? vc2 : vc1);
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const float bl = p ? bl1 : bl0;
                f4 v;
                for (int r = 0; r < 4; r++) v[r] = (float)(p ? (wv[r] >> 16) : (wv[r] & 0xffffu)) - bl;
                const float c = lane_value(v[0], 0);
                const f4 T1 = product(make_op<V, SPLIT>((v - c) * DN_SCALE), F1);
                f4 C = product(F2, make_op<V, SPLIT>(T1));
                // per-lane block: upper/lower (G >= 2), left/right (n16 >= 8)
                const float dc0 = lane_value(C[0], 0), dc1 = lane_value(C[0], 8);
                const float dc2 = lane_value(C[0], 32), dc3 = lane_value(C[0], 40);
                const float dc = dc_src == 0 ? dc0 : dc_src == 8 ? dc1 : dc_src == 32 ? dc2 : dc3;
                const float mean = dc * (0.125f / DN_SCALE) + c;
                const float t_abs = DN_SCALE * a.thr * __builtin_sqrtf(a.sig_a * fmaxf(mean, 0.0f) + a.sig_b);
                for (int r = 0; r < 4; r++) {
                    const bool keep = (r == 0 && is_dc_lane) || fabsf(C[r]) >= t_abs;
                    C[r] = (keep && blk_ok) ? C[r] : 0.0f;
                }
                const f4 Wt = product(make_op<V, SPLIT>(C), F3);
                f4 R = product(make_op<V, SPLIT>(Wt), F4);
                for (int r = 0; r < 4; r++) {
                    const bool ok = ((r >= 2) ? vr2 : vr1) && ((n16 >= 8) ? vc2 : vc1);
                    R[r] = ok ? R[r] * (1.0f / DN_SCALE) + c : 0.0f;
                }
                X[t][p] = R;
            }
        }
        // Overlap-add across columns (DPP) and down the column (registers), then store.
#pragma unroll
        for (int t = 1; t <= DN_NB; t++) {
            const int b = b0 - 1 + t;
            const int x = 8 * b;
            float o[2][2];
#pragma unroll
            for (int p = 0; p < 2; p++) {
                f4 Y;
                for (int r = 0; r < 4; r++) Y[r] = X[t][p][r] + row_shr4(X[t][p][r]) + row_shl12(X[t - 1][p][r]);
                o[p][0] = Y[0] + (edge_group ? carry[t - 1][p][0] : Y[2]);
                o[p][1] = Y[1] + (edge_group ? carry[t - 1][p][1] : Y[3]);
                carry[t - 1][p][0] = Y[2];
                carry[t - 1][p][1] = Y[3];
            }
            if (ta >= a0 && b < (int)a.nt && out_lane) {
                const int pcol = x + out_col;
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const int prow = y + out_row + r;
                    if (prow < (int)a.ph && pcol < (int)a.pw) {
                        const float inv = inv_cover(prow, a.ply) * inv_cover(pcol, a.plx);
                        const float v0 = o[0][r] * inv, v1 = o[1][r] * inv;
                        const size_t off = (size_t)(2 * prow + dy) * a.stride + 2 * pcol;
                        if (a.dst) {
                            const float u0 = fminf(fmaxf(v0 + bl0 + 0.5f, 0.0f), 65535.0f);
                            const float u1 = fminf(fmaxf(v1 + bl1 + 0.5f, 0.0f), 65535.0f);
                            *reinterpret_cast<uint32_t *>(a.dst + off) = (uint32_t)u0 | ((uint32_t)u1 << 16);
                        }
                        if (a.dst_f) *reinterpret_cast<float2 *>(a.dst_f + off) = make_float2(v0, v1);
                    }
                }
            }
        }
    }
}


// ---- f32 scalar variant ----

struct DnScalarArgs {
    const uint16_t *src;
    float *blocks;  // [plane][by][bx][64]
    uint32_t stride, pw, ph, step, nx, ny, gx, gy;  // gx, gy: positions on the regular grid
    float black[4];
    float sig_a, sig_b, thr;
    float d[64];
};

__device__ __forceinline__ uint32_t dn_pos(uint32_t i, uint32_t n_grid, uint32_t step, uint32_t plane_n) {
    return i < n_grid ? i * step : plane_n - 8;
}

// One thread per (plane, block): the oracle's f32 statement (oracle.c fro_dct_denoise,
// precision 32) in the same order, without contraction.
__global__ __launch_bounds__(256) void k_denoise_blocks_f32(DnScalarArgs a) {
    const uint32_t per_plane = a.nx * a.ny;
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= 4 * per_plane) return;
    const uint32_t plane = id / per_plane, rem = id % per_plane;
    const uint32_t by = rem / a.nx, bx = rem % a.nx;
    const uint32_t dy = plane >> 1, dx = plane & 1;
    const uint32_t y0 = dn_pos(by, a.gy, a.step, a.ph), x0 = dn_pos(bx, a.gx, a.step, a.pw);
    const float bl = a.black[plane];
    float X[8][8], T[8][8], C[8][8];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++)
            X[i][j] = (float)a.src[(size_t)(2 * (y0 + i) + dy) * a.stride + 2 * (x0 + j) + dx] - bl;
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            float acc = 0.0f;
#pragma unroll
            for (int i = 0; i < 8; i++) acc += a.d[k * 8 + i] * X[i][j];
            T[k][j] = acc;
        }
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
        for (int l = 0; l < 8; l++) {
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < 8; j++) acc += T[k][j] * a.d[l * 8 + j];
            C[k][l] = acc;
        }
    const float mean = C[0][0] / 8.0f;
    const float sig = sqrtf(a.sig_a * (mean > 0.0f ? mean : 0.0f) + a.sig_b);
    const float t = a.thr * sig;
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
        for (int l = 0; l < 8; l++)
            if ((k | l) && !(fabsf(C[k][l]) >= t)) C[k][l] = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int l = 0; l < 8; l++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; k++) acc += a.d[k * 8 + i] * C[k][l];
            T[i][l] = acc;
        }
    float *o = a.blocks + (size_t)id * 64;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            float acc = 0.0f;
#pragma unroll
            for (int l = 0; l < 8; l++) acc += T[i][l] * a.d[l * 8 + j];
            o[i * 8 + j] = acc;
        }
}

// Blocks covering plane coordinate p: grid positions first..last (ascending), then the extra
// last position when it is off the grid.
__device__ __forceinline__ void dn_cover(uint32_t p, uint32_t step, uint32_t n_grid, uint32_t n_all, uint32_t plane_n,
                                         uint32_t &lo, uint32_t &hi, bool &extra) {
    lo = p >= 7 ? (p - 7 + step - 1) / step : 0;
    hi = min(p / step, n_grid - 1);  // inclusive; lo > hi means none
    extra = n_all > n_grid && p >= plane_n - 8;
}

struct DnGatherArgs {
    const float *blocks;
    uint16_t *dst;
    float *dst_f;
    uint32_t stride, pw, ph, step, nx, ny, gx, gy;
    float black[4];
};

__global__ __launch_bounds__(256) void k_denoise_gather(DnGatherArgs a) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    const uint32_t per_plane = a.pw * a.ph;
    if (id >= 4 * per_plane) return;
    const uint32_t plane = id / per_plane, rem = id % per_plane;
    const uint32_t py = rem / a.pw, px = rem % a.pw;
    const uint32_t dy = plane >> 1, dx = plane & 1;
    uint32_t ylo, yhi, xlo, xhi;
    bool yex, xex;
    dn_cover(py, a.step, a.gy, a.ny, a.ph, ylo, yhi, yex);
    dn_cover(px, a.step, a.gx, a.nx, a.pw, xlo, xhi, xex);
    float sum = 0.0f, cnt = 0.0f;
    const float *pb = a.blocks + (size_t)plane * a.nx * a.ny * 64;
    auto add_row = [&](uint32_t by) {
        const uint32_t y0 = dn_pos(by, a.gy, a.step, a.ph);
        for (uint32_t bx = xlo; bx <= xhi && bx < a.gx; bx++) {
            const uint32_t x0 = bx * a.step;
            sum += pb[((size_t)by * a.nx + bx) * 64 + (py - y0) * 8 + (px - x0)];
            cnt += 1.0f;
        }
        if (xex) {
            const uint32_t bx = a.nx - 1, x0 = a.pw - 8;
            sum += pb[((size_t)by * a.nx + bx) * 64 + (py - y0) * 8 + (px - x0)];
            cnt += 1.0f;
        }
    };
    for (uint32_t by = ylo; by <= yhi && by < a.gy; by++) add_row(by);
    if (yex) add_row(a.ny - 1);
    const float v = sum / cnt;
    const size_t o = (size_t)(2 * py + dy) * a.stride + 2 * px + dx;
    if (a.dst_f) a.dst_f[o] = v;
    if (a.dst) a.dst[o] = (uint16_t)fminf(fmaxf(v + a.black[plane] + 0.5f, 0.0f), 65535.0f);
}

void dct_matrix(float d[64]) {
    for (int k = 0; k < 8; k++) {
        const double s = k == 0 ? std::sqrt(1.0 / 8.0) : std::sqrt(2.0 / 8.0);
        for (int i = 0; i < 8; i++) d[k * 8 + i] = (float)(s * std::cos(M_PI * (2 * i + 1) * k / 16.0));
    }
}

uint32_t n_positions(uint32_t n, uint32_t step, uint32_t *grid) {
    const uint32_t g = (n - 8) / step + 1;
    *grid = g;
    return (g - 1) * step == n - 8 ? g : g + 1;
}

}  // namespace
}  // namespace frh

using namespace frh;

extern "C" int32_t frh_denoise_cfa(frh_context *ctx, const frh_dn_spec *spec, const void *src, void *dst,
                                   float *dst_f32, void *stream) {
    try {
        if (!ctx || !spec || !src) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        if (spec->size < sizeof(frh_dn_spec)) return fail(FRH_ERROR_INVALID_ARGUMENT, "spec size field too small");
        if (!dst && !dst_f32) return fail(FRH_ERROR_INVALID_ARGUMENT, "no output");
        const uint32_t W = spec->width, H = spec->height, stride = spec->stride ? spec->stride : W;
        if (W < 16 || H < 16 || (W & 1) || (H & 1) || stride < W)
            return fail(FRH_ERROR_INVALID_ARGUMENT, "mosaic must be at least 16x16 with even sides");
        if (spec->step < 1 || spec->step > 8) return fail(FRH_ERROR_INVALID_ARGUMENT, "step must be 1..8");
        if (!(spec->threshold >= 0.0f) || !(spec->sig_b >= 0.0f) || !(spec->sig_a >= 0.0f))
            return fail(FRH_ERROR_INVALID_ARGUMENT, "negative or NaN noise parameters");
        const uint32_t pw = W / 2, ph = H / 2;
        (void)hipSetDevice(ctx->device);
        hipStream_t st = (hipStream_t)stream;
        hipError_t e;
        if (spec->precision == FRH_DN_F16 || spec->precision == FRH_DN_BF16 || spec->precision == FRH_DN_F16X3) {
            if (spec->step != 4 || (pw - 8) % 4 || (ph - 8) % 4)
                return fail(FRH_ERROR_NOT_AVAILABLE,
                            "the MFMA denoiser needs stride 4 and a mosaic whose sides are multiples of 8");
            if ((reinterpret_cast<uintptr_t>(src) & 15) || (stride % 8) ||
                (dst && (reinterpret_cast<uintptr_t>(dst) & 3)) || (dst_f32 && (reinterpret_cast<uintptr_t>(dst_f32) & 7)))
                return fail(FRH_ERROR_INVALID_ARGUMENT, "source must be 16-byte aligned with a row stride that is a multiple of 8");
            const char *impl = getenv("FLASHRAW_DN_IMPL");  // "walker": the separable 16x16x16 kernel
            const bool walker = impl && strcmp(impl, "walker") == 0;
            if (spec->precision != FRH_DN_F16X3 && !walker) {
                // the kernel's stores address the outputs with 32-bit byte offsets
                const uint64_t samples = (uint64_t)H * stride;
                if ((dst && samples * 2 >= (1ull << 31)) || (dst_f32 && samples * 4 >= (1ull << 31)))
                    return fail(FRH_ERROR_NOT_AVAILABLE, "the MFMA denoiser's outputs must be below 2 GiB");
                // f16 / bf16: the Kronecker-form kernel (dn_kron.hip)
                const bool bf = spec->precision == FRH_DN_BF16;
                // the forward and inverse operand tables, one device allocation per operand type,
                // looked up by a short key (the tables depend only on the operand type)
                const uint32_t key[2] = {0x4b4e4444u /* "DDNK" */, bf ? 1u : 0u};
                int32_t rc = FRH_OK;
                const void *tabs = cached_table(ctx, 2, key, sizeof key, [&](std::vector<uint8_t> &dev) {
                    float d[64];
                    dct_matrix(d);
                    std::vector<uint16_t> kf, ki;
                    dn_kron_tables(d, bf, kf, ki);
                    dev.resize((kf.size() + ki.size()) * 2);
                    std::memcpy(dev.data(), kf.data(), kf.size() * 2);
                    std::memcpy(dev.data() + kf.size() * 2, ki.data(), ki.size() * 2);
                    return true;
                }, &rc);
                if (rc) return rc;
                DnkArgs k{};
                k.kf = static_cast<const uint16_t *>(tabs);
                k.ki = k.kf + 64 * 64;
                k.src = static_cast<const uint16_t *>(src);
                k.dst = static_cast<uint16_t *>(dst);
                k.dst_f = dst_f32;
                k.stride = stride;
                k.pw = (int)pw;
                k.ph = (int)ph;
                k.plx = (int)pw - 8;
                k.ply = (int)ph - 8;
                k.strips = (int)((pw + DNK_STRIP - 1) / DNK_STRIP);
                k.nsteps = (int)((ph + 7) / 8);
                k.seg_steps = spec->seg_rows ? (int)spec->seg_rows : dn_kron_seg_steps(k.strips, k.nsteps, ctx->num_cus);
                k.segs = (k.nsteps + k.seg_steps - 1) / k.seg_steps;
                for (int i = 0; i < 4; i++) k.black[i] = spec->black[i];
                const float t = spec->threshold / 16.0f;  // the forward's coefficients are scaled by 1/16
                k.t2a = t * t * spec->sig_a;
                k.t2b = t * t * spec->sig_b;
                if ((e = launch_dn_kron(k, bf, st)) != hipSuccess) return hip_fail(e, "denoise launch");
                return FRH_OK;
            }
            DnArgs a{};
            a.src = static_cast<const uint16_t *>(src);
            a.dst = static_cast<uint16_t *>(dst);
            a.dst_f = dst_f32;
            a.stride = stride;
            a.pw = pw;
            a.ph = ph;
            a.plx = (int)pw - 8;
            a.ply = (int)ph - 8;
            a.nt = (pw + 7) / 8;
            a.nr = (ph + 7) / 8;
            a.strips = (a.nt + DN_NB - 1) / DN_NB;
            a.seg_rows = spec->seg_rows ? spec->seg_rows : 12;
            a.segs = (a.nr + a.seg_rows - 1) / a.seg_rows;
            for (int i = 0; i < 4; i++) a.black[i] = spec->black[i];
            a.sig_a = spec->sig_a;
            a.sig_b = spec->sig_b;
            a.thr = spec->threshold;
            dct_matrix(a.d);
            const uint32_t waves = a.strips * a.segs * 2;
            const uint32_t grid = (waves + DN_WAVES - 1) / DN_WAVES;
            if (spec->precision == FRH_DN_F16)
                hipLaunchKernelGGL((k_denoise_mfma<half4, false>), dim3(grid), dim3(DN_WAVES * 64), 0, st, a);
            else if (spec->precision == FRH_DN_BF16)
                hipLaunchKernelGGL((k_denoise_mfma<bf16x4, false>), dim3(grid), dim3(DN_WAVES * 64), 0, st, a);
            else
                hipLaunchKernelGGL((k_denoise_mfma<half4, true>), dim3(grid), dim3(DN_WAVES * 64), 0, st, a);
            if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "denoise launch");
            return FRH_OK;
        }
        if (spec->precision != FRH_DN_F32) return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown precision");
        uint32_t gx, gy;
        const uint32_t nx = n_positions(pw, spec->step, &gx), ny = n_positions(ph, spec->step, &gy);
        const uint64_t blocks = 4ull * nx * ny;
        const uint64_t bytes = blocks * 64 * sizeof(float);
        float *scratch = static_cast<float *>(ctx->alloc(bytes, &e));
        if (!scratch) return hip_fail(e, "denoise scratch");
        DnScalarArgs s{};
        s.src = static_cast<const uint16_t *>(src);
        s.blocks = scratch;
        s.stride = stride;
        s.pw = pw;
        s.ph = ph;
        s.step = spec->step;
        s.nx = nx;
        s.ny = ny;
        s.gx = gx;
        s.gy = gy;
        for (int i = 0; i < 4; i++) s.black[i] = spec->black[i];
        s.sig_a = spec->sig_a;
        s.sig_b = spec->sig_b;
        s.thr = spec->threshold;
        dct_matrix(s.d);
        hipLaunchKernelGGL(k_denoise_blocks_f32, dim3((uint32_t)((blocks + 255) / 256)), dim3(256), 0, st, s);
        DnGatherArgs g{};
        g.blocks = scratch;
        g.dst = static_cast<uint16_t *>(dst);
        g.dst_f = dst_f32;
        g.stride = stride;
        g.pw = pw;
        g.ph = ph;
        g.step = spec->step;
        g.nx = nx;
        g.ny = ny;
        g.gx = gx;
        g.gy = gy;
        for (int i = 0; i < 4; i++) g.black[i] = spec->black[i];
        const uint64_t px = 4ull * pw * ph;
        hipLaunchKernelGGL(k_denoise_gather, dim3((uint32_t)((px + 255) / 256)), dim3(256), 0, st, g);
        e = hipGetLastError();
        // the scratch returns to the pool only once the launches that use it have finished
        const hipError_t e2 = hipStreamSynchronize(st);
        ctx->release(scratch, bytes);
        if (e != hipSuccess) return hip_fail(e, "denoise launch");
        if (e2 != hipSuccess) return hip_fail(e2, "denoise");
        return FRH_OK;
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}
