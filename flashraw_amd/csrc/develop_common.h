// Device helpers shared by the develop kernels: calibration, the finish (white balance, clip,
// matrix, levels), the encoders, orientation, and the bilinear pixel read straight from the
// frame. CPU statement: crates/flashraw/src/process.rs (Calib 246-299, finish_rgb 654-707,
// encoders 1593-1991, encode_developed 2018-2051, row_cols 967-1129).
#pragma once

#include "frh_device.h"

namespace frh {

struct Lds {
    float black[64];
    float scale[64];
    uint32_t lut[1025];  // 4097 sRGB codes, bytes packed four per word
};

// A kernel's DevArgs read in place from the kernel argument segment, where `args`, the kernel's
// first argument, starts at offset 0. Kernels whose code indexes the struct's arrays at run time
// take their arguments this way: naming the by-value argument itself there makes the compiler
// copy all of it (about 1 KB) to private memory, per lane, on every launch.
__device__ __forceinline__ const DevArgs &kernel_args(const DevArgs &args) {
    (void)args;
    return *(const DevArgs *)__builtin_amdgcn_kernarg_segment_ptr();
}

__device__ __forceinline__ void load_lds(const DevArgs &a, Lds &s) {
    for (int i = threadIdx.x; i < 64; i += blockDim.x) {
        s.black[i] = a.black[i];
        s.scale[i] = a.scale[i];
    }
    const uint32_t *l = reinterpret_cast<const uint32_t *>(a.lut8);
    for (int i = threadIdx.x; i < 1025; i += blockDim.x) s.lut[i] = l[i];
}

__device__ __forceinline__ uint32_t lut8(const Lds &s, uint32_t i) {
    return reinterpret_cast<const uint8_t *>(s.lut)[i];
}

// Channel (0 R, 1 G, 2 B) of frame sample (x, y) of a Bayer mosaic.
__device__ __forceinline__ uint32_t chan_at(const DevArgs &a, int x, int y) {
    const uint32_t i = (((uint32_t)y + a.oy) & 1u) * 2u + (((uint32_t)x + a.ox) & 1u);
    return (a.chan_bits >> (2u * i)) & 3u;
}

// White-balance multiplier of channel c (selects on uniform values, no indexed argument load).
__device__ __forceinline__ float wb_of(const DevArgs &a, uint32_t c) {
    return c == 0 ? a.wb[0] : (c == 1 ? a.wb[1] : a.wb[2]);
}

// Calib::norm (process.rs:288-298) of frame sample (x, y), in bounds.
__device__ __forceinline__ float norm_raw(const DevArgs &a, const Lds &s, float v, uint32_t x, uint32_t y);
// FULL = false: the u16 register walkers' form (no f32 samples, tone or output curve: frames with
// them take the other kernels).
template <bool FULL = true>
__device__ __forceinline__ float norm_at(const DevArgs &a, const Lds &s, uint32_t x, uint32_t y) {
    const size_t i = (size_t)y * a.stride + x;
    return norm_raw(a, s, (FULL && a.srcf) ? a.srcf[i] : (float)a.src[i], x, y);
}

// finish_rgb with Highlights::Clip (process.rs:654-707). `balanced` pixels come from the AHD
// mosaic, which is already white balanced (finish_run's unit balance, process.rs:1209-1213).
__device__ __forceinline__ void finish(const DevArgs &a, float r, float g, float b, bool balanced, float &o0,
                                       float &o1, float &o2) {
    if (!balanced) {
        r = r * a.wb[0];
        g = g * a.wb[1];
        b = b * a.wb[2];
    } else if (!a.unit_k) {
        r = r * a.kk[0];
        g = g * a.kk[1];
        b = b * a.kk[2];
    }
    r = rclamp01(r);
    g = rclamp01(g);
    b = rclamp01(b);
    o0 = a.m[0] * r + a.m[1] * g + a.m[2] * b;
    o1 = a.m[3] * r + a.m[4] * g + a.m[5] * b;
    o2 = a.m[6] * r + a.m[7] * g + a.m[8] * b;
    if (a.has_levels) {
        o0 = o0 * a.lscale - a.loffset;
        o1 = o1 * a.lscale - a.loffset;
        o2 = o2 * a.lscale - a.loffset;
    }
}

// f32_to_f16 (process.rs:1978-1991).
__device__ __forceinline__ uint32_t f16_bits(float x) {
    uint32_t b = __float_as_uint(x);
    uint32_t sign = (b >> 16) & 0x8000u;
    uint32_t a = b & 0x7fffffffu;
    uint32_t normal = (a + 0xc8000fffu + ((a >> 13) & 1u)) >> 13;
    uint32_t sub = __float_as_uint(__uint_as_float(a) + 0.5f) - 0x3f000000u;
    uint32_t h = a < 0x38800000u ? sub : normal;
    h = a >= 0x477ff000u ? (a > 0x7f800000u ? 0x7e00u : 0x7bffu) : h;
    return sign | h;
}

// srgb16_paired (process.rs:1960-1967).
__device__ __forceinline__ uint32_t srgb16(const DevArgs &a, float v) {
    float x = rclamp01(v) * 16384.0f;
    uint32_t i = trunc_u32(x);
    i = i < 16383u ? i : 16383u;
    float t = x - (float)i;
    float2 p = a.srgb16[i];
    return trunc_u32(p.x + p.y * t + 0.5f);
}

// linear16 (process.rs:1971-1973).
__device__ __forceinline__ uint32_t lin16(float v) { return trunc_u32(rclamp01(v) * 65535.0f + 0.5f); }

// ToneMap's max / min (crates/flashraw/src/tone.rs:98-108): plain comparisons (NaN -> the second).
__device__ __forceinline__ float tmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float tmin(float a, float b) { return a < b ? a : b; }

// ToneMap::position + read (tone.rs:148-162): the curve at linear x >= 0.
__device__ __forceinline__ float tone_curve(const DevArgs &a, float x) {
    const float t = sqrtf(tmin(x, a.tone_xmax)) * a.tone_scale;  // correctly rounded
    const float t1 = tmin(t, (float)(a.tone_n - 1u));
    const int i = (int)t1;
    const float2 p = a.tone_tab[i];
    return p.x + p.y * (t - (float)i);
}

// ToneMap::apply (tone.rs:170-207): one pixel of linear output RGB.
__device__ __forceinline__ void tone_px(const DevArgs &a, float &r, float &g, float &b) {
    const float *m = a.tone_m;
    const float c0 = tmax(m[0] * r + m[1] * g + m[2] * b, 0.0f);
    const float c1 = tmax(m[3] * r + m[4] * g + m[5] * b, 0.0f);
    const float c2 = tmax(m[6] * r + m[7] * g + m[8] * b, 0.0f);
    const float hi = tmax(tmax(c0, c1), c2), lo = tmin(tmin(c0, c1), c2);
    const float fh = tone_curve(a, hi), fl = tone_curve(a, lo);
    const float d = hi - lo;
    const float k = d > 0.0f ? (fh - fl) / d : 0.0f;
    const float base = fl - k * lo;
    r = base * a.tone_w[0] + k * r;
    g = base * a.tone_w[1] + k * g;
    b = base * a.tone_w[2] + k * b;
}

// An 8-bit code of linear v: the sRGB table in LDS, or the output curve's table (Lut8,
// process.rs:1679-1728: table[min(trunc(clamp(v, 0, 1) * n), n)]).
__device__ __forceinline__ uint32_t code8(const DevArgs &a, const Lds &s, float v) {
    if (!a.curve8) return lut8(s, lut8_index(v));
    const float x = v * (float)a.curve8_n;
    const uint32_t i = x >= 1.0f ? (uint32_t)fminf(x, (float)a.curve8_n) : 0u;
    return a.curve8[i];
}

// Destination of developed pixel (dx, dy) under EXIF orientation o (the inverse of
// encode_developed's source mapping, process.rs:2035-2044).
__device__ __forceinline__ void orient(int o, uint32_t dx, uint32_t dy, uint32_t w, uint32_t h, uint32_t &px,
                                       uint32_t &py) {
    switch (o) {
    case 2: px = w - 1 - dx; py = dy; break;
    case 3: px = w - 1 - dx; py = h - 1 - dy; break;
    case 4: px = dx; py = h - 1 - dy; break;
    case 5: px = dy; py = dx; break;
    case 6: px = h - 1 - dy; py = dx; break;
    case 7: px = h - 1 - dy; py = w - 1 - dx; break;
    case 8: px = dy; py = w - 1 - dx; break;
    default: px = dx; py = dy; break;
    }
}

// orient() as a linear map: px = X0 + Xx dx + Xy dy and py = Y0 + Yx dx + Yy dy, so the destination
// index py * dst_w + px of developed pixel (dx, dy) is c0 + cx dx + cy dy (modulo 2^32).
__device__ __forceinline__ void orient_linear(const DevArgs &a, uint32_t &c0, uint32_t &cx, uint32_t &cy) {
    const int w = (int)a.dev_w, h = (int)a.dev_h;
    int X0 = 0, Xx = 1, Xy = 0, Y0 = 0, Yx = 0, Yy = 1;
    switch (a.orientation) {
    case 2: X0 = w - 1; Xx = -1; break;
    case 3: X0 = w - 1; Xx = -1; Y0 = h - 1; Yy = -1; break;
    case 4: Y0 = h - 1; Yy = -1; break;
    case 5: Xx = 0; Xy = 1; Yx = 1; Yy = 0; break;
    case 6: X0 = h - 1; Xx = 0; Xy = -1; Yx = 1; Yy = 0; break;
    case 7: X0 = h - 1; Xx = 0; Xy = -1; Y0 = w - 1; Yx = -1; Yy = 0; break;
    case 8: Xx = 0; Xy = 1; Y0 = w - 1; Yx = -1; Yy = 0; break;
    default: break;
    }
    const uint32_t dw = a.dst_w;
    c0 = (uint32_t)Y0 * dw + (uint32_t)X0;
    cx = (uint32_t)Yx * dw + (uint32_t)Xx;
    cy = (uint32_t)Yy * dw + (uint32_t)Xy;
}

// Encode and store linear output pixel (r, g, b) of developed pixel (dx, dy) (any encoding and
// orientation).
template <bool FULL = true>
__device__ __forceinline__ void store_px(const DevArgs &a, const Lds &s, int mode, void *dst, uint32_t dx,
                                         uint32_t dy, float r, float g, float b) {
    if (FULL && a.tone && (mode == OUT_RGBA8 || mode == OUT_SRGB16)) tone_px(a, r, g, b);
    if (mode == OUT_RGBA8) {
        uint32_t px, py;
        orient(a.orientation, dx, dy, a.dev_w, a.dev_h, px, py);
        uint32_t v = FULL ? (code8(a, s, r) | (code8(a, s, g) << 8) | (code8(a, s, b) << 16) | 0xff000000u)
                          : (lut8(s, lut8_index(r)) | (lut8(s, lut8_index(g)) << 8) | (lut8(s, lut8_index(b)) << 16) |
                             0xff000000u);
        reinterpret_cast<uint32_t *>(dst)[(size_t)py * a.dst_w + px] = v;
    } else if (mode == OUT_LINEAR_F32) {
        float *p = reinterpret_cast<float *>(dst) + ((size_t)dy * a.dev_w + dx) * 3;
        p[0] = r;
        p[1] = g;
        p[2] = b;
    } else if (mode == OUT_SRGB16) {
        uint32_t px, py;
        orient(a.orientation, dx, dy, a.dev_w, a.dev_h, px, py);
        uint16_t *p = reinterpret_cast<uint16_t *>(dst) + ((size_t)py * a.dst_w + px) * 3;
        p[0] = (uint16_t)srgb16(a, r);
        p[1] = (uint16_t)srgb16(a, g);
        p[2] = (uint16_t)srgb16(a, b);
    } else if (mode == OUT_LINEAR16) {
        uint16_t *p = reinterpret_cast<uint16_t *>(dst) + ((size_t)dy * a.dev_w + dx) * 3;
        p[0] = (uint16_t)lin16(r);
        p[1] = (uint16_t)lin16(g);
        p[2] = (uint16_t)lin16(b);
    } else if (mode == OUT_LINEAR_F16) {
        uint16_t *p = reinterpret_cast<uint16_t *>(dst) + ((size_t)dy * a.dev_w + dx) * 3;
        p[0] = (uint16_t)f16_bits(r);
        p[1] = (uint16_t)f16_bits(g);
        p[2] = (uint16_t)f16_bits(b);
    }
}

template <bool FULL = true>
__device__ __forceinline__ void store_both(const DevArgs &a, const Lds &s, uint32_t dx, uint32_t dy, float r,
                                           float g, float b) {
    store_px<FULL>(a, s, a.out_mode, a.dst, dx, dy, r, g, b);
    if (a.out_mode2 != OUT_NONE) store_px<FULL>(a, s, a.out_mode2, a.dst2, dx, dy, r, g, b);
}

// The RGBA8 word of linear output (r, g, b) (process.rs:1712-1728; alpha 255, R in the low byte).
__device__ __forceinline__ uint32_t rgba8_word(const Lds &s, float r, float g, float b) {
    return lut8(s, lut8_index(r)) | (lut8(s, lut8_index(g)) << 8) | (lut8(s, lut8_index(b)) << 16) | 0xff000000u;
}

// The workgroups of one dispatch go to the device's 8 XCDs round-robin (workgroup b to XCD
// b mod 8), and each XCD has its own L2. xcd_block renumbers the n workgroups so that every XCD
// runs one contiguous range of logical blocks: neighbouring strips, which share cache lines at
// their edges and halo rows, then meet in one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    constexpr uint32_t X = 8;
    const uint32_t q = n / X, r = n % X, x = b % X, slot = b / X;
    return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// True when a launch stores RGBA8 only, unrotated, without levels, with bounded parameters: the
// kernels' inline fast path.
__host__ __device__ __forceinline__ bool fast_rgba8(const DevArgs &a) {
    return a.out_mode == OUT_RGBA8 && a.out_mode2 == OUT_NONE && a.orientation == 1 && a.bounded && !a.has_levels &&
           !a.tone && !a.curve8 && a.unit_k;
}

// True when a launch stores rotated RGBA8 (EXIF orientation 2..8) only, under the same conditions:
// the AHD walker's rotated fast path, for tame frames.
__host__ __device__ __forceinline__ bool fast_rgba8_rot(const DevArgs &a) {
    return a.out_mode == OUT_RGBA8 && a.out_mode2 == OUT_NONE && a.orientation >= 2 && a.orientation <= 8 &&
           a.bounded && !a.has_levels && !a.tone && !a.curve8 && a.unit_k;
}

// True when a launch stores RGBA8 and linear RGB16 (in either order) under the same conditions:
// the AHD walker's two-output fast path, for tame frames.
__host__ __device__ __forceinline__ bool fast_rgba8_lin16(const DevArgs &a) {
    const bool pair = (a.out_mode == OUT_RGBA8 && a.out_mode2 == OUT_LINEAR16) ||
                      (a.out_mode == OUT_LINEAR16 && a.out_mode2 == OUT_RGBA8);
    return pair && a.orientation == 1 && a.dst_w == a.dev_w && a.bounded && !a.has_levels && !a.tone && !a.curve8 &&
           a.unit_k;
}

__device__ __forceinline__ float med01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

// finish + RGBA8 encode of the fast path (finite values, no levels): the same operations as
// finish() and rgba8_word() with the clamps as v_med3_f32 (or the clamp output modifier).
__device__ __forceinline__ uint32_t finish_rgba8_fast(const DevArgs &a, const Lds &s, float r, float g, float b,
                                                      bool balanced) {
    if (!balanced) {
        r = r * a.wb[0];
        g = g * a.wb[1];
        b = b * a.wb[2];
    }
    r = med01(r);
    g = med01(g);
    b = med01(b);
    const float o0 = a.m[0] * r + a.m[1] * g + a.m[2] * b;
    const float o1 = a.m[3] * r + a.m[4] * g + a.m[5] * b;
    const float o2 = a.m[6] * r + a.m[7] * g + a.m[8] * b;
    const uint32_t i0 = (uint32_t)(med01(o0) * 4096.0f), i1 = (uint32_t)(med01(o1) * 4096.0f),
                   i2 = (uint32_t)(med01(o2) * 4096.0f);
    return lut8(s, i0) | (lut8(s, i1) << 8) | (lut8(s, i2) << 16) | 0xff000000u;
}

// finish_rgba8_fast of balanced pixels when the frame is AHD-tame (DevArgs::ahd_tame): the matrix
// rows pre-scaled by 4096 give exactly 4096 times each output (every product is 0 or normal), so
// clamping to [0, 4096] equals clamping to [0, 1] and scaling by 4096.
__device__ __forceinline__ uint32_t finish_rgba8_scaled(const DevArgs &a, const Lds &s, float r, float g, float b) {
    r = med01(r);
    g = med01(g);
    b = med01(b);
    const float o0 = a.m4k[0] * r + a.m4k[1] * g + a.m4k[2] * b;
    const float o1 = a.m4k[3] * r + a.m4k[4] * g + a.m4k[5] * b;
    const float o2 = a.m4k[6] * r + a.m4k[7] * g + a.m4k[8] * b;
    const uint32_t i0 = (uint32_t)__builtin_amdgcn_fmed3f(o0, 0.0f, 4096.0f),
                   i1 = (uint32_t)__builtin_amdgcn_fmed3f(o1, 0.0f, 4096.0f),
                   i2 = (uint32_t)__builtin_amdgcn_fmed3f(o2, 0.0f, 4096.0f);
    return lut8(s, i0) | (lut8(s, i1) << 8) | (lut8(s, i2) << 16) | 0xff000000u;
}

// finish_rgba8_scaled that also returns the three clamped outputs c = 4096 clamp(o, 0, 1) (exact),
// for the linear 16-bit code of the same pixel: linear16(o) = trunc(clamp(o, 0, 1) * 65535 + 0.5)
// (process.rs:1971-1973) is trunc(c * (65535 / 4096) + 0.5), since 65535 / 4096 is a float and
// c * (65535 / 4096) is the same real number as clamp(o, 0, 1) * 65535, rounded once.
__device__ __forceinline__ uint32_t finish_rgba8_scaled_c(const DevArgs &a, const Lds &s, float r, float g, float b,
                                                          float c[3]) {
    r = med01(r);
    g = med01(g);
    b = med01(b);
    const float o0 = a.m4k[0] * r + a.m4k[1] * g + a.m4k[2] * b;
    const float o1 = a.m4k[3] * r + a.m4k[4] * g + a.m4k[5] * b;
    const float o2 = a.m4k[6] * r + a.m4k[7] * g + a.m4k[8] * b;
    c[0] = __builtin_amdgcn_fmed3f(o0, 0.0f, 4096.0f);
    c[1] = __builtin_amdgcn_fmed3f(o1, 0.0f, 4096.0f);
    c[2] = __builtin_amdgcn_fmed3f(o2, 0.0f, 4096.0f);
    return lut8(s, (uint32_t)c[0]) | (lut8(s, (uint32_t)c[1]) << 8) | (lut8(s, (uint32_t)c[2]) << 16) | 0xff000000u;
}
__device__ __forceinline__ uint32_t lin16_of_4k(float c) { return (uint32_t)(c * (65535.0f / 4096.0f) + 0.5f); }

// Raw samples of frame columns x0, x0 + 1 of row y as one word (x0 in the low half). Addresses
// are clamped into the frame, so the load never faults and needs no branch; the caller masks
// samples outside the frame. HOT: the pair is one aligned 32-bit word inside the frame.
// A raw buffer resource over `bytes` bytes at `base`: loads and stores then take a scalar
// (wave-uniform) byte offset plus a 32-bit per-lane offset, with no 64-bit address arithmetic.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
    const uint64_t p = (uint64_t)(uintptr_t)base;
    const uint64_t q = ((uint64_t)uniform((uint32_t)(p >> 32)) << 32) | uniform((uint32_t)p);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)q, (short)0, (int)uniform(bytes), 0x00020000);
}

// Bytes from `a.src` to the end of the frame's last sample.
__device__ __forceinline__ uint32_t src_bytes(const DevArgs &a) {
    return ((a.height - 1u) * a.stride + a.width) * 2u;
}

template <bool HOT>
__device__ __forceinline__ uint32_t fetch_pair(const DevArgs &a, __amdgpu_buffer_rsrc_t rs, int x0, int y) {
    const int yc = min(max(y, 0), (int)a.height - 1);
    const int soff = (int)uniform((uint32_t)yc * a.stride * 2u);
    if (HOT) return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, x0 * 2, soff, 0);
    const int c0 = min(max(x0, 0), (int)a.width - 1), c1 = min(max(x0 + 1, 0), (int)a.width - 1);
    const uint32_t v0 = __builtin_amdgcn_raw_buffer_load_b16(rs, c0 * 2, soff, 0);
    const uint32_t v1 = __builtin_amdgcn_raw_buffer_load_b16(rs, c1 * 2, soff, 0);
    return v0 | (v1 << 16);
}

// Calib::norm (process.rs:288-298) of raw sample v at frame (x, y), in the frame.
__device__ __forceinline__ float norm_raw(const DevArgs &a, const Lds &s, float v, uint32_t x, uint32_t y) {
    if (a.simple_cal) {
        // at most a 2x2 pattern and no deltas: the per-parity values (the same numbers as the table)
        const uint32_t i = (y & 1u) * 2u + (x & 1u);
        const float b = i == 0 ? a.qb[0] : i == 1 ? a.qb[1] : i == 2 ? a.qb[2] : a.qb[3];
        const float k = i == 0 ? a.qs[0] : i == 1 ? a.qs[1] : i == 2 ? a.qs[2] : a.qs[3];
        return (v - b) * k;
    }
    uint32_t X = x + a.ox, Y = y + a.oy;
    uint32_t i = a.cal_pow2 ? (Y & (a.cal_ph - 1u)) * a.cal_pw + (X & (a.cal_pw - 1u))
                            : (Y % a.cal_ph) * a.cal_pw + X % a.cal_pw;
    float b = s.black[i];
    if (a.n_rowd) b += Y < a.n_rowd ? a.rowd[Y] : 0.0f;
    if (a.n_cold) b += X < a.n_cold ? a.cold[X] : 0.0f;
    return (v - b) * s.scale[i];
}

// Normalised samples of a fetched pair at row y (parity PY); 0 outside the frame (`in0`, `in1`:
// the lane's columns are inside, `row_in`: the row is). HOT: every column of the wave is inside
// the frame and the calibration is the scalar one; a row outside the frame is then normalised
// from its clamped fetch like any other. No stored pixel depends on such a row (the walkers store
// only pixels whose whole neighbourhood, 1 row for bilinear and 5 for AHD, lies in the frame;
// frame-edge pixels come from the border blocks), and without the wave-uniform select, which
// the compiler makes a branch, a walker's trip stays one block of straight-line code.
template <int PY, bool HOT>
__device__ __forceinline__ void norm_pair(const DevArgs &a, const Lds &s, uint32_t raw, int x0, int y, bool row_in,
                                          bool in0, bool in1, float &n0, float &n1) {
    const float v0 = (float)(raw & 0xffffu), v1 = (float)(raw >> 16);
    if (HOT) {
        n0 = (v0 - a.qb[PY * 2 + 0]) * a.qs[PY * 2 + 0];
        n1 = (v1 - a.qb[PY * 2 + 1]) * a.qs[PY * 2 + 1];
        return;
    }
    if (a.simple_cal) {
        n0 = (v0 - a.qb[PY * 2 + 0]) * a.qs[PY * 2 + 0];
        n1 = (v1 - a.qb[PY * 2 + 1]) * a.qs[PY * 2 + 1];
    } else {
        const uint32_t yc = (uint32_t)min(max(y, 0), (int)a.height - 1);
        n0 = norm_raw(a, s, v0, (uint32_t)min(max(x0, 0), (int)a.width - 1), yc);
        n1 = norm_raw(a, s, v1, (uint32_t)min(max(x0 + 1, 0), (int)a.width - 1), yc);
    }
    n0 = row_in && in0 ? n0 : 0.0f;
    n1 = row_in && in1 ? n1 : 0.0f;
}

// Store the RGBA8 words of developed pixels (dx, dy) and (dx + 1, dy) on the fast path; `n` is 1
// or 2 pixels, 8-byte stores when aligned.
__device__ __forceinline__ void store_rgba8_pair(const DevArgs &a, __amdgpu_buffer_rsrc_t rd, uint32_t dx,
                                                 uint32_t dy, uint32_t w0, uint32_t w1, int n) {
    const int soff = (int)uniform(dy * a.dst_w * 4u);
    const int voff = (int)(dx * 4u);
    if (n == 2 && ((dy * a.dst_w + dx) & 1u) == 0) {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        u32x2 v;
        v.x = w0;
        v.y = w1;
        __builtin_amdgcn_raw_buffer_store_b64(v, rd, voff, soff, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(w0, rd, voff, soff, 0);
        if (n == 2) __builtin_amdgcn_raw_buffer_store_b32(w1, rd, voff + 4, soff, 0);
    }
}

// Kernel::row_cols for one pixel of a Bayer mosaic, read straight from the frame
// (process.rs:1076-1126): the interior formulas, or at the frame border the mean of the
// in-frame same-colour 3x3 samples in row-major order.
template <bool FULL = true>
__device__ __forceinline__ void bilinear_global(const DevArgs &a, const Lds &s, int x, int y, float rgb[3]) {
    const int w = (int)a.width, h = (int)a.height;
    uint32_t c0 = chan_at(a, x, y);
    if (x > 0 && y > 0 && x + 1 < w && y + 1 < h) {
        float own = norm_at<FULL>(a, s, x, y);
        float l = norm_at<FULL>(a, s, x - 1, y), r = norm_at<FULL>(a, s, x + 1, y);
        float u = norm_at<FULL>(a, s, x, y - 1), d = norm_at<FULL>(a, s, x, y + 1);
        if (c0 == 1) {
            uint32_t ch = chan_at(a, x + 1, y);
            float hv = (l + r) * 0.5f, vv = (u + d) * 0.5f;
            rgb[1] = own;
            rgb[0] = ch == 0 ? hv : vv;
            rgb[2] = ch == 0 ? vv : hv;
        } else {
            float orth = (l + r + u + d) * 0.25f;
            float diag = (norm_at<FULL>(a, s, x - 1, y - 1) + norm_at<FULL>(a, s, x + 1, y - 1) + norm_at<FULL>(a, s, x - 1, y + 1) +
                          norm_at<FULL>(a, s, x + 1, y + 1)) *
                         0.25f;
            rgb[0] = c0 == 0 ? own : diag;
            rgb[1] = orth;
            rgb[2] = c0 == 0 ? diag : own;
        }
        return;
    }
    for (uint32_t c = 0; c < 3; c++) {
        float sum = 0.0f, n = 0.0f;
        if (c == c0) {
            sum = norm_at<FULL>(a, s, x, y);
            n = 1.0f;
        } else {
            for (int dy = -1; dy <= 1; dy++) {
                for (int dx = -1; dx <= 1; dx++) {
                    int xx = x + dx, yy = y + dy;
                    if (chan_at(a, xx, yy) != c) continue;
                    if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
                    sum += norm_at<FULL>(a, s, xx, yy);
                    n += 1.0f;
                }
            }
        }
        rgb[c] = n > 0.0f ? sum / n : 0.0f;
    }
}

// Pixels of a band outside a demosaic's interior [halo, w - halo) x [halo, h - halo), developed
// by the bilinear formulas (strip_to, process.rs:1183-1224), enumerated row-major: the rows
// above the interior, the side columns of the interior rows, the rows below.
struct Border {
    int ix0, ix1, iy0, iy1;  // interior in developed coordinates (iy0 = iy1 = band end: no interior)
    uint32_t count;
};

__host__ inline Border make_border(const DevArgs &a, int halo) {
    Border b;
    const int cx0 = (int)a.crop_x, cy0 = (int)a.crop_y;
    const int x_first = halo > cx0 ? halo : cx0;
    const int x_end = (int)a.width - halo < cx0 + (int)a.dev_w ? (int)a.width - halo : cx0 + (int)a.dev_w;
    const int y_first = halo > cy0 + (int)a.band_y0 ? halo : cy0 + (int)a.band_y0;
    const int y_end = (int)a.height - halo < cy0 + (int)a.band_y1 ? (int)a.height - halo : cy0 + (int)a.band_y1;
    const uint64_t band_px = (uint64_t)a.dev_w * (a.band_y1 - a.band_y0);
    if (x_end > x_first && y_end > y_first) {
        b.ix0 = x_first - cx0;
        b.ix1 = x_end - cx0;
        b.iy0 = y_first - cy0;
        b.iy1 = y_end - cy0;
        b.count = (uint32_t)(band_px - (uint64_t)(b.ix1 - b.ix0) * (b.iy1 - b.iy0));
    } else {
        b.ix0 = b.ix1 = 0;
        b.iy0 = b.iy1 = (int)a.band_y1;
        b.count = (uint32_t)band_px;
    }
    return b;
}

// Developed coordinates of border pixel i of a band: the rows above the interior, the columns
// left and right of it, the rows below.
__device__ __forceinline__ void border_xy(const DevArgs &a, const Border &B, uint32_t i, int &dx, int &dy) {
    const int dw = (int)a.dev_w;
    const int b0 = (int)a.band_y0;
    const int top = B.iy0 - b0;
    const int mid = B.iy1 - B.iy0;
    const int side = B.ix0 + (dw - B.ix1);
    const uint32_t ntop = (uint32_t)top * dw, nmid = (uint32_t)mid * side;
    if (i < ntop) {
        dy = b0 + (int)(i / dw);
        dx = (int)(i % dw);
    } else if (i < ntop + nmid) {
        const uint32_t j = i - ntop;
        dy = B.iy0 + (int)(j / side);
        const int c = (int)(j % side);
        dx = c < B.ix0 ? c : B.ix1 + (c - B.ix0);
    } else {
        const uint32_t j = i - ntop - nmid;
        dy = B.iy1 + (int)(j / dw);
        dx = (int)(j % dw);
    }
}

template <bool FULL = true>
__device__ __forceinline__ void border_pixel(const DevArgs &a, const Lds &s, const Border &B, uint32_t i) {
    int dx, dy;
    border_xy(a, B, i, dx, dy);
    float rgb[3];
    bilinear_global<FULL>(a, s, (int)a.crop_x + dx, (int)a.crop_y + dy, rgb);
    float o0, o1, o2;
    finish(a, rgb[0], rgb[1], rgb[2], false, o0, o1, o2);
    store_both<FULL>(a, s, (uint32_t)dx, (uint32_t)dy, o0, o1, o2);
}

// Calib::norm (process.rs:288-298) of component c of window sample (x, y) from the generic table
// gcal (black[cpp][pattern] then scale[cpp][pattern]).
__device__ __forceinline__ float gen_norm(const DevArgs &a, uint32_t x, uint32_t y, uint32_t c) {
    const size_t at = (size_t)y * a.stride + (size_t)x * a.cpp + c;
    const float v = a.srcf ? a.srcf[at] : (float)a.src[at];
    const uint32_t X = x + a.ox, Y = y + a.oy;
    const uint32_t P = a.cal_pw * a.cal_ph;
    const uint32_t i = c * P + (Y % a.cal_ph) * a.cal_pw + X % a.cal_pw;
    float b = a.gcal[i];
    if (a.n_rowd) b += Y < a.n_rowd ? a.rowd[Y] : 0.0f;
    if (a.n_cold) b += X < a.n_cold ? a.cold[X] : 0.0f;
    return (v - b) * a.gcal[a.cpp * P + i];
}

// Site descriptor of window sample (x, y): own colour in bits 0..1, then per colour R, G, B the
// 9-bit mask of its 3x3 neighbours of that colour.
__device__ __forceinline__ uint32_t site_at(const DevArgs &a, uint32_t x, uint32_t y) {
    return a.sites[((y + a.oy) % a.cfa_h) * a.cfa_w + (x + a.ox) % a.cfa_w];
}

// Kernel::row_cols for a single-component CFA other than Bayer (process.rs:1098-1126): each colour
// is the mean of the same-colour samples of the 3x3 neighbourhood, sum * (1 / n) inside the frame
// and sum / n over the neighbours inside at the border, neighbours in row-major order.
__device__ __forceinline__ void cfa_mean_rgb(const DevArgs &a, uint32_t x, uint32_t y, float &r, float &g, float &b) {
    const uint32_t site = site_at(a, x, y);
    const bool inside = x > 0 && x + 1 < a.width && y > 0 && y + 1 < a.height;
    float v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const uint32_t m = (site >> (2 + 9 * c)) & 0x1ffu;
        float sum = 0.0f, n = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; t++) {
            if (!((m >> t) & 1u)) continue;
            const int xx = (int)x + t % 3 - 1, yy = (int)y + t / 3 - 1;
            if (!inside && (xx < 0 || xx >= (int)a.width || yy < 0 || yy >= (int)a.height)) continue;
            sum += gen_norm(a, (uint32_t)xx, (uint32_t)yy, 0);
            n += 1.0f;
        }
        if (inside) {
            const uint32_t cnt = __builtin_popcount(m);
            v[c] = sum * (cnt ? 1.0f / (float)cnt : 0.0f);
        } else {
            v[c] = n > 0.0f ? sum / n : 0.0f;
        }
    }
    r = v[0];
    g = v[1];
    b = v[2];
}

// Red's index in the frame-parity colour table fc[(y & 1) * 2 + (x & 1)] (the walkers' template
// parameter), or -1.
__host__ inline int red_position(const DevArgs &a) {
    for (int i = 0; i < 4; i++) {
        const uint32_t y = (uint32_t)(i >> 1), x = (uint32_t)(i & 1);
        if (a.chan[((y + a.oy) & 1u) * 2u + ((x + a.ox) & 1u)] == 0) return i;
    }
    return -1;
}

// Correctly rounded square root of x = max(v, 0) (IEEE sqrt, the reference's f32::sqrt), for x = +-0
// and x >= 2^-96: the hardware reciprocal square root estimate y, g = x y, then one refinement
// g + (x - g^2) y / 2 with both steps fused (the residual x - g^2 is exact in an fma). On gfx950 this
// equals the correctly rounded square root for every float in [2^-96, 2^128), checked exhaustively
// against the compiler's correctly rounded sqrtf (frh_check_sqrt; a GPU test repeats the check).
// At x = +-0 the estimate of 2^-96 keeps g = x and the result x. Smaller positive inputs (never
// produced by normalised samples, whose non-zero magnitudes are far larger) take the library path
// for the whole wave (sqrt_max0), and so does +inf (where g would be inf * 0).
__device__ __forceinline__ float sqrt_nonneg(float x) {
    const float y = __builtin_amdgcn_rsqf(fmaxf(x, 0x1p-96f));
    const float g = x * y;
    const float r = __builtin_fmaf(-g, g, x);
    return __builtin_fmaf(r, 0.5f * y, g);
}
__device__ __forceinline__ float sqrt_max0(float v) {
    const float x = fmaxf(v, 0.0f);
    float r = sqrt_nonneg(x);
    const bool rare = (x > 0.0f && x < 0x1p-96f) || x == __builtin_inff();
    if (__builtin_amdgcn_ballot_w64(rare)) {
        asm volatile("" ::: "memory");  // keep this rare path a branch, not speculated into every row
        r = rare ? sqrtf(x) : r;
    }
    return r;
}
// max / min; NN: neither operand is a NaN, so v_med3_f32 with an infinite third operand gives the
// same value and, unlike v_max_f32 / v_min_f32 in IEEE mode, needs no quieting of its operands.
// `inf` is +inf held in a register the compiler cannot see through (opaque_inf), or the median
// would be folded back into a max / min.
template <bool NN>
__device__ __forceinline__ float fmax_nn(float a, float b, float inf) {
    return NN ? __builtin_amdgcn_fmed3f(a, b, inf) : fmaxf(a, b);
}
template <bool NN>
__device__ __forceinline__ float fmin_nn(float a, float b, float inf) {
    return NN ? __builtin_amdgcn_fmed3f(a, b, -inf) : fminf(a, b);
}
__device__ __forceinline__ float opaque_inf() {
    float r;
    asm volatile("s_mov_b32 %0, 0x7f800000" : "=s"(r));
    return r;
}
// sqrt(max(v, 0)); TAME: v is known to be 0 or at least 2^-96 and finite (DevArgs::ahd_tame), so
// max(max(v, 0), 2^-96) = max(v, 2^-96) feeds the estimate directly.
template <bool TAME>
__device__ __forceinline__ float sqrt_m0(float v, float inf) {
    if (!TAME) return sqrt_max0(v);
    const float x = fmax_nn<true>(v, 0.0f, inf);
    const float y = __builtin_amdgcn_rsqf(fmax_nn<true>(v, 0x1p-96f, inf));
    const float g = x * y;
    const float r = __builtin_fmaf(-g, g, x);
    return __builtin_fmaf(r, 0.5f * y, g);
}

// Value of the lane to the left (lane - 1) / right (lane + 1) of this one (DPP wave shifts; the
// edge lanes read 0 and are halo lanes). Every lane must be active when these run.
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ uint32_t from_left_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ uint32_t from_right_u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, true);
}

// Colour (0 R, 1 G, 2 B) of a sample with frame row parity py and column parity px, when red
// sits at index RP of the frame-parity table and blue opposite it.
template <int RP>
__device__ __forceinline__ constexpr int site(int py, int px) {
    return (py * 2 + px) == RP ? 0 : ((py * 2 + px) == 3 - RP ? 2 : 1);
}

}  // namespace frh
