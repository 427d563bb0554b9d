/*
 * CPU oracle for the fused RAW develop path and the CFA-plane DCT block denoiser.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, operation for operation, the reference's
 * CPU develop path (crates/flashraw/src/process.rs, crates/flashraw/src/ahd.rs) so that the
 * HIP kernels can be checked against it, and it doubles as the CPU baseline that bench.py
 * times. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (flashraw_amd/lib/libflashraw_hip.so) never links or calls it.
 */
#ifndef FLASHRAW_ORACLE_H
#define FLASHRAW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A single-component CFA mosaic with its calibration (the subset of the reference's RawImage +
 * FrameInfo the develop path reads: crates/flashraw-core/src/decoder.rs:25-74). */
typedef struct fro_frame {
    const uint16_t *u16;      /* width*height samples, row-major */
    uint32_t width, height;
    uint32_t origin_x, origin_y; /* sensor position of sample (0,0): anchors CFA and black */
    uint32_t cfa_w, cfa_h;    /* 1..8 */
    uint8_t cfa[64];          /* CfaColor codes (0 R, 1 G, 2 B, others undeveloped), [y*cfa_w+x] */
    uint32_t black_rows, black_cols, black_samples;
    const float *black_values;
    uint32_t n_black_values;
    const float *row_deltas;
    uint32_t n_row_deltas;
    const float *col_deltas;
    uint32_t n_col_deltas;
    const uint32_t *white_values;
    uint32_t n_white;
    /* f32 samples (raw values) in place of u16 when not NULL (RawImage PixelData::F32) */
    const float *f32;
    /* components per pixel (0 means 1); cfa_w == 0 means no CFA */
    uint32_t cpp;
} fro_frame;

enum fro_demosaic { FRO_BILINEAR = 0, FRO_AHD = 1, FRO_EDGE = 2 };

/* Tables of the edge-aware demosaic (crates/flashraw/src/edgeaware/mod.rs:57-160) of a pattern
 * of `period` 2 (Bayer) or 6 (X-Trans): g0[P*P][49] (the non-directional green filter per
 * canonical position) and dir[P*P][3][4][49] (per canonical position, colour and direction),
 * taps row-major over 7x7; canon[P*P] the colours of the canonical pattern; (cx, cy) the
 * canonical position of frame sample (0, 0); beta the weight of the gated smoothing. */
typedef struct fro_edge {
    const float *g0;
    const float *dir;
    uint8_t canon[36];
    uint32_t cx, cy;
    float beta;
    uint32_t period; /* 0 means 2 */
} fro_edge;

/* Encodings (crates/flashraw/src/process.rs:770-781), plus 8-bit RGB without alpha. */
enum fro_encoding {
    FRO_RGBA8 = 0,
    FRO_SRGB16 = 1,
    FRO_LINEAR16 = 2,
    FRO_LINEAR_F16 = 3,
    FRO_LINEAR_F32 = 4,
    FRO_RGB8 = 5
};

/* The standard tone of the display outputs (crates/flashraw/src/tone.rs:81-207, GpuTone): in
 * linear ProPhoto (m: output RGB -> ProPhoto) the largest and smallest channel go through the
 * curve table[i] = (value, step) read at t = sqrt(min(x, xmax)) * scale, and the result is
 * (fl - k lo) white + k p. */
typedef struct fro_tone {
    const float *table; /* n pairs */
    uint32_t n;
    float xmax, scale;
    float m[9];
    float white[3];
} fro_tone;

typedef struct fro_params {
    int32_t demosaic;
    uint32_t crop_x, crop_y, crop_w, crop_h;
    float wb[4];
    float m[9];              /* camera -> linear sRGB, row-major */
    int32_t has_levels;
    float levels_scale, levels_offset;
    int32_t orientation;     /* EXIF 1..8, applied to RGBA8 / SRGB16 / RGB8 only */
    int32_t encoding;
    const fro_edge *edge;    /* FRO_EDGE only */
    int32_t half_size;       /* Demosaic::HalfSize: 2x2 blocks averaged per colour */
    const fro_tone *tone;    /* display outputs (RGBA8, RGB8, SRGB16) only; NULL for none */
    /* output curve (CurveTables, crates/flashraw/src/color.rs:343-367): 8-bit codes at
     * curve8_len - 1 segments, 16-bit codes at 16385 points; NULL for the sRGB curve */
    const uint8_t *curve8;
    uint32_t curve8_len;
    const float *curve16;
    uint32_t curve16_len;
} fro_params;

/* Fused develop of `f` with `p` into `out` (out_w*out_h*channels samples of the encoding's
 * type). Returns 0, or -1 for an invalid request. `threads` <= 0 uses all OpenMP threads. */
int fro_develop(const fro_frame *f, const fro_params *p, void *out, uint64_t out_bytes, int threads,
                uint32_t *out_w, uint32_t *out_h);

/* The 4097-entry sRGB 8-bit table (process.rs:1903-1915). */
void fro_srgb8_table(uint8_t out[4097]);
/* The 16385-entry sRGB 16-bit table (process.rs:1922-1934). */
void fro_srgb16_table(float out[16385]);
/* IEEE half bits of x (process.rs:1978-1991). */
uint16_t fro_f32_to_f16(float x);

/* The normalised frame (Kernel::normalize_row, process.rs:881-945, for every row) into
 * out[height][width * cpp]. Returns 0, or -1 for an invalid frame. */
int fro_normalize(const fro_frame *f, float *out);
/* process::normalize (process.rs:302-320): as fro_normalize, each value then max(v, 0). */
int fro_normalize_clamped(const fro_frame *f, float *out);

/* PackOrder (crates/flashraw-core/src/decoder.rs:203-212). */
enum fro_pack_order { FRO_MSB_BYTES = 0, FRO_LSB_BYTES = 1, FRO_MSB16 = 2, FRO_MSB32 = 3 };
/* unpack_row_scalar (crates/flashraw-codecs/src/packed.rs:137-180) over `h` rows of `w` samples
 * of `bits` (1..16) bits: row y is src[y * stride .. min(y * stride + row_bytes, len)] with
 * row_bytes = ceil(w * bits / 8), zeros beyond (the row semantics of the DNG decoders and of
 * crates/flashraw-gpu/tests/equivalence.rs:52-63). Returns 0, or -1 for invalid arguments. */
int fro_unpack(const uint8_t *src, uint64_t len, uint32_t w, uint32_t h, uint32_t bits, int order, uint64_t stride,
               uint16_t *out);
/* The linearisation curve of integer samples: v -> lut[min(v, len - 1)] (run.rs:38-39). */
void fro_linearize(uint16_t *v, uint64_t n, const uint16_t *lut, uint32_t lut_len);
/* Plan::set_noise_from (crates/flashraw/src/edgeaware/mod.rs:369-409) over the normalised frame
 * `norm` (w x h): the smoothing weight of the edge-aware demosaic. canon[period^2] are the
 * colours of the canonical pattern, (cx, cy) the canonical position of frame sample (0, 0). */
float fro_edge_beta(const float *norm, uint32_t w, uint32_t h, const uint8_t *canon, uint32_t period, uint32_t cx,
                    uint32_t cy, float g_gain);

/* ---- CFA-plane DCT block denoiser (build-defined workload, no reference counterpart) ----
 * The mosaic is split into its four 2x2 phase planes (W/2 x H/2 each, black subtracted, in DN);
 * 8x8 blocks at stride `step` (positions include the last one) are transformed with the
 * orthonormal DCT-II, AC coefficients with |c| < thr * sigma are zeroed, where
 * sigma^2 = sig_a * max(DC / 8, 0) + sig_b, the blocks are transformed back and overlap-added
 * with uniform weights. `out_f32` (may be NULL) receives the denoised planes re-interleaved as a
 * mosaic in DN above black; `out_u16` (may be NULL) receives the mosaic with black added back,
 * rounded and clamped to [0, 65535]. `precision` 64 computes in double, 32 in float. */
int fro_dct_denoise(const uint16_t *mosaic, uint32_t width, uint32_t height, const float black[4],
                    float sig_a, float sig_b, float thr, uint32_t step, int precision, float *out_f32,
                    uint16_t *out_u16, int threads);


/* ---- luma / chroma noise reduction of developed linear RGB (crates/flashraw/src/denoise/mod.rs,
 * nr.c) ---- */

/* denoise::apply on interleaved RGB (3 floats per pixel), in place. */
int fro_nr_apply(float *rgb, uint32_t w, uint32_t h, float luma, float chroma);
/* The same with the guided filter's box means as direct window sums in the order of the
 * reference's GPU kernels (box_direct 1) instead of the CPU path's running sums (0). */
int fro_nr_apply_order(float *rgb, uint32_t w, uint32_t h, float luma, float chroma, int box_direct);
/* Pyramid levels (level_count) and the noise estimate's block step (estimate_step, 0: none). */
uint32_t fro_nr_level_count(uint32_t w, uint32_t h, uint32_t max);
uint32_t fro_nr_estimate_step(uint32_t w, uint32_t h);
/* The measured blocks of a plane (per block: guide mean, then 63 squared coefficients; entry
 * l * 8 + k is horizontal frequency l, vertical k). Returns the count; out NULL only counts. */
uint32_t fro_nr_stats(const float *p, const float *guide, uint32_t w, uint32_t h, float *out);
/* Model::from_stats: bins (at most 12) means and 64 deviations per bin; -1 if too few blocks. */
int fro_nr_model_from_stats(const float *stats, uint32_t count, uint32_t *bins, float *means, float *sigmas);
/* Pieces: the 2-D DCT of an 8x8 block and its inverse, box_mean, down, add_up, one DCT pass. */
void fro_nr_dct_block(const float *p, uint32_t w, uint32_t x0, uint32_t y0, float *out);
void fro_nr_inverse_block(const float *coef, float *out);
int fro_nr_box_mean(const float *in, uint32_t w, uint32_t h, uint32_t r, float *out);
int fro_nr_box_mean_order(const float *in, uint32_t w, uint32_t h, uint32_t r, float *out, int direct);
int fro_nr_down(const float *in, uint32_t w, uint32_t h, float *out);
int fro_nr_add_up(float *d, uint32_t w, uint32_t h, const float *coarse);
int fro_nr_dct_pass(const float *src, const float *guide, const float *pilot, uint32_t w, uint32_t h,
                    uint32_t bins, const float *means, const float *sigmas, float gain, float hard, uint32_t step,
                    float *out);

#ifdef __cplusplus
}
#endif
#endif
