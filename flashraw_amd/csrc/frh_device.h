// Device-side parameter block and scalar helpers shared by the develop kernels (gfx950).
//
// Arithmetic rules (they make the kernels bit-identical to the reference's CPU develop,
// crates/flashraw/src/process.rs and ahd.rs): every file is compiled with -ffp-contract=off, so
// no multiply-add is fused; f32 division and sqrt are HIP's correctly rounded defaults; sums keep
// the reference's left-to-right order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frh {

// Output encodings (frh_develop_output).
enum OutMode : int32_t {
    OUT_RGBA8 = 0,
    OUT_LINEAR_F32 = 1,
    OUT_SRGB16 = 2,
    OUT_LINEAR16 = 3,
    OUT_LINEAR_F16 = 4,
    OUT_NONE = -1
};

// Tap lists of the edge-aware filters (edgeaware/mod.rs:130-160) for a Bayer pattern: per
// canonical position, the taps of G0 whose coefficient is not zero and, per colour, the taps where
// some direction is not zero (with all four directions' coefficients), in tap order. The dense
// copies (zero taps included) serve the kernel that multiplies every tap, which gives the same
// bits when every value it multiplies is finite (a zero tap then adds +-0 to a sum that is never
// -0).
struct EdgeTab {
    float g0d[4][49];           // dense Bayer taps (the tame kernel)
    float dird[4][3][49][4];
    uint32_t canon[36];         // colour of each canonical position
    uint32_t g0_n[36];          // per canonical position: G0's non-zero taps
    uint32_t dir_n[36][3];      // per position and colour: taps non-zero in some direction
    uint32_t g0_idx[36][49];
    float g0_c[36][49];
    uint32_t dir_idx[36][3][49];
    float dir_c[36][3][49][4];
};

// Everything one develop launch reads, passed by value as the kernel argument.
struct DevArgs {
    // input: `src` points at the window's sample (0, 0); `stride` samples per raster row
    const uint16_t *src;
    uint32_t stride;
    uint32_t width, height;  // frame (window) size
    uint32_t ox, oy;         // sensor position of frame sample (0, 0)
    // calibration pattern (Calib, process.rs:246-299): cal_pw x cal_ph <= 64 entries
    uint32_t cal_pw, cal_ph;
    uint32_t cal_pow2;       // 1 when cal_pw and cal_ph are powers of two (mask instead of modulo)
    float black[64];
    float scale[64];
    const float *rowd;
    uint32_t n_rowd;
    const float *cold;
    uint32_t n_cold;
    // Bayer channel of sensor parity ((Y & 1) * 2 + (X & 1))
    uint32_t chan[4];
    uint32_t chan_bits;      // chan[i] in bits 2i..2i+1 (register-only lookups in the kernels)
    // The calibration at frame parity (y & 1) * 2 + (x & 1) when the pattern is at most 2x2 and
    // there are no row/column deltas (`simple_cal`): the walkers read these as scalars.
    uint32_t simple_cal;
    float qb[4], qs[4];
    // All parameters are bounded (|value| <= 1e6), so no infinity or NaN can arise and the RGBA8
    // fast path may clamp with v_med3_f32 (identical to f32::clamp on finite values; the sign of a
    // zero cannot change an 8-bit code).
    uint32_t bounded;
    // The calibration, white balance and matrix bound the AHD walker's values (see ahd_tame in
    // api.cpp): every value it takes a square root of is 0 or at least 2^-96 and finite, and every
    // product of the RGBA8 finish is a normal float, so the square root needs no test for tiny or
    // infinite inputs and the finish may take the matrix pre-scaled by 4096 (`m4k`).
    uint32_t ahd_tame;
    float m4k[9];
    // The AHD and edge-aware finish multiplies the balanced pixel by wb / wb per channel
    // (process.rs:1209): 1 for every multiplier except 0, whose channel becomes NaN. unit_k: all
    // three are 1 and the multiply is skipped.
    uint32_t unit_k;
    float kk[3];
    // develop region and band
    uint32_t crop_x, crop_y, dev_w, dev_h;
    uint32_t band_y0, band_y1;  // developed rows [band_y0, band_y1) of this launch
    float wb[4];
    float m[9];
    int32_t has_levels;
    float lscale, loffset;
    int32_t orientation;  // EXIF 1..8 for the display encodings
    uint32_t dst_w;       // destination width of the display encodings (oriented)
    int32_t out_mode;
    void *dst;
    int32_t out_mode2;
    void *dst2;
    const uint8_t *lut8;     // 4097 entries (+ padding to 4100)
    const float2 *srgb16;    // 16384 (value, step) pairs
    const float *wtab;       // 73 x 73 AHD blend weights a / (a + b)
    uint32_t ahd_halo;       // 6: pixels nearer the border than this are bilinear (ahd.rs:33)
    uint32_t ahd_impl;       // 0: register row walker, 1: LDS-tiled (both give the same bytes)
    uint32_t ahd_rows;       // rows per walker strip (even)
    uint32_t ahd_wpb;        // waves per block of the AHD walker (1, 2, 4 or 8; 0: 4)
    uint32_t bil_impl;       // bilinear: 0 register row walker, 1 LDS-tiled (the same bytes)
    uint32_t bil_rows;       // rows per bilinear walker strip (even)
    uint32_t bil_wpb;        // waves per block of the bilinear walker (1, 2, 4, 8 or 16; 0: 4)
    uint32_t bil_cpl;        // frame columns per lane of the bilinear walker: 4 (walk4_ok launches) or 2
    // edge-aware demosaic: tap lists, canonical position of frame sample (0, 0), smoothing weight
    // (the border width is ahd_halo)
    uint32_t edge_aware;
    const EdgeTab *etab;
    uint32_t ecanon_bits;  // colour of canonical position i in bits 2i..2i+1
    uint32_t ecx, ecy;
    float ebeta;
    uint32_t eperiod;  // 2 (Bayer) or 6 (X-Trans)
    float exg;         // X-Trans green re-estimate weight clamp(1 - 2 beta, 0, 1)
    // The inputs bound every intermediate value of the edge-aware demosaic far from overflow
    // (see edge_tables in api.cpp): dense taps and uniform colour masks give the same bits.
    uint32_t edge_tame;
    // Frames beyond a u16 Bayer mosaic. srcf (when not NULL) holds f32 raw values in place of src;
    // cpp components per pixel (stride counts samples). The generic kernel develops every frame
    // that is not a single-component Bayer mosaic at full size: mode 0 grey from component 0,
    // 1 a single-component CFA of cfa_w x cfa_h (site descriptors `sites`), 2 RGB from
    // components 0..2; `half` averages 2x2 blocks. gcal holds black[cpp][cal_pw * cal_ph] then
    // scale[cpp][cal_pw * cal_ph].
    const float *srcf;
    uint32_t cpp;
    uint32_t generic;
    uint32_t gmode;
    uint32_t half;
    const float *gcal;
    const uint32_t *sites;
    uint32_t cfa_w, cfa_h;
    // Display outputs (RGBA8, sRGB16): the standard tone (GpuTone: `tone_n` (value, step) pairs
    // read at sqrt(min(x, xmax)) * scale, output RGB -> ProPhoto `tone_m`, ProPhoto white
    // `tone_w`) and an RGBA8 output curve (`curve8_n` segments; NULL: the sRGB table). A curve's
    // 16-bit table replaces `srgb16`.
    uint32_t tone;
    const float2 *tone_tab;
    uint32_t tone_n;
    float tone_xmax, tone_scale;
    float tone_m[9], tone_w[3];
    const uint8_t *curve8;
    uint32_t curve8_n;
};

// Waves per block of a single-frame walker launch, a tile-shape parameter the autotuner chooses with
// the strip height: a power of two up to `max` (8 for AHD, whose registers allow 3 waves per SIMD
// and so no 16-wave block; 16 for bilinear), anything else 4.
constexpr int AHD_MAX_WAVES = 8, BIL_MAX_WAVES = 16;
__host__ __device__ inline uint32_t walk_wpb(uint32_t w, uint32_t max) {
    return (w >= 1 && w <= max && (w & (w - 1)) == 0) ? w : 4u;
}

// Window of a raster for the input kernels (unpack, linearise, normalise): packed payload bytes
// (kind 2) or u16 / f32 samples (kinds 0 / 1), the window's first sample (x0 counts samples,
// x * cpp) and row, and `n` samples per window row over `rows` rows.
struct InArgs {
    const void *src;
    uint64_t len;        // packed: payload bytes
    uint64_t stride;     // packed: bytes per raster row; u16 / f32: samples per raster row
    uint32_t row_bytes;  // packed: bytes of sample data per raster row (bytes past it read 0)
    uint32_t bits;       // packed: 1..16
    uint32_t kind;
    uint32_t x0, y0;
    uint32_t n, rows;
    const uint16_t *lut;  // linearisation of integer samples, NULL for none
    uint32_t lut_len;
    // normalise: origin of the window, component count and calibration (as DevArgs)
    uint32_t ox, oy, cpp;
    uint32_t cal_pw, cal_ph;
    const float *gcal;
    const float *rowd, *cold;
    uint32_t n_rowd, n_cold;
};

// Rust `x as u32` for x below 2^32: truncation, with NaN and everything below 1 giving 0.
__device__ __forceinline__ uint32_t trunc_u32(float x) { return x >= 1.0f ? (uint32_t)x : 0u; }

// The sRGB8 table index of linear v: `min(trunc(clamp(v, 0, 1) * 4096), 4096)`
// (process.rs:1712-1728). Scaling by 4096 is exact, so clamping after the scale is the same.
__device__ __forceinline__ uint32_t lut8_index(float v) {
    float x = v * 4096.0f;
    return x >= 1.0f ? (uint32_t)fminf(x, 4096.0f) : 0u;
}

// f32::clamp(x, 0, 1): `if x < 0 {0}; if x > 1 {1}; x` (NaN and signed zero pass through).
__device__ __forceinline__ float rclamp01(float x) {
    x = x < 0.0f ? 0.0f : x;
    return x > 1.0f ? 1.0f : x;
}

}  // namespace frh
