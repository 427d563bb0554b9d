/*
 * flashraw_hip.h: C ABI of the MI355X (gfx950) develop path.
 *
 * Drop-in boundary for the reference's GPU develop entry points (crates/flashraw-gpu). Every
 * declaration cites the reference interface it replaces. Conventions follow the reference's C
 * API (crates/flashraw-capi/src/lib.rs:1-16): `frh_` prefix, int32 enums, structs that start with
 * their own `size`, a status code per call and a thread-local last error message. No call
 * panics or throws across the boundary.
 */
#ifndef FLASHRAW_HIP_H
#define FLASHRAW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the numbering of the reference's `flashraw_status`
 * (crates/flashraw-capi/include/flashraw.h:44-114) for the kinds this path can return. */
enum frh_status {
    FRH_OK = 0,
    FRH_ERROR_GPU_UNAVAILABLE = 6,
    FRH_ERROR_OUT_OF_MEMORY = 7,
    FRH_ERROR_LIMIT_EXCEEDED = 8,
    FRH_ERROR_NOT_AVAILABLE = 9,
    FRH_ERROR_INVALID_ARGUMENT = 10,
    FRH_ERROR_INTERNAL = 12,
    FRH_ERROR_CANCELLED = 14
};

typedef struct frh_context frh_context;
typedef struct frh_image frh_image;

/* ---- cancellation (flashraw_cancel_token, crates/flashraw-capi/include/flashraw.h:818-824 and
 * 1247-1275; CancelToken, crates/flashraw-core/src/cancel.rs:26-76) ----
 * A shared flag that asks a running develop or noise reduction to stop. The entry points that take
 * one (frh_run_develop, frh_develop_into, frh_run_denoise, frh_denoise_into, through the spec's
 * `cancel` field) check it between their stages and launches, as the reference's GPU facade does
 * between its stages (crates/flashraw/src/gpu.rs:292-299): work already queued on the device runs
 * to its end, nothing further is queued, and the call returns FRH_ERROR_CANCELLED with no result
 * (the caller's output buffers then hold unspecified bytes). Cancel and is_cancelled may be called
 * from any thread, also while a call uses the token; a token must not be freed while a call holds
 * it. A cancelled token stays cancelled. */
typedef struct frh_cancel_token frh_cancel_token;
/* flashraw_cancel_token_new: a token that is not cancelled. */
frh_cancel_token *frh_cancel_token_new(void);
/* CancelToken::with_poll (cancel.rs:40-45): a token that also asks poll(user) on every check
 * while it is not cancelled, and is cancelled for good once poll answers non-zero. poll runs on the
 * thread that checks and must be cheap. */
frh_cancel_token *frh_cancel_token_with_poll(int32_t (*poll)(void *user), void *user);
/* flashraw_cancel_token_cancel: does nothing for NULL. */
void frh_cancel_token_cancel(const frh_cancel_token *t);
/* flashraw_cancel_token_is_cancelled: non-zero once cancelled (0 for NULL). */
int32_t frh_cancel_token_is_cancelled(const frh_cancel_token *t);
/* flashraw_cancel_token_free: does nothing for NULL. */
void frh_cancel_token_free(frh_cancel_token *t);

/* flashraw_core::Rect */
typedef struct frh_rect {
    uint32_t x, y, width, height;
} frh_rect;

/* GpuInput (crates/flashraw-gpu/src/run.rs:11-24): decoded u16 or f32 samples, or a uniformly
 * bit-packed payload unpacked on the device. */
enum frh_input_kind { FRH_INPUT_U16 = 0, FRH_INPUT_F32 = 1, FRH_INPUT_PACKED = 2 };

/* PackOrder (crates/flashraw-core/src/decoder.rs:203-212). */
enum frh_pack_order { FRH_PACK_MSB_BYTES = 0, FRH_PACK_LSB_BYTES = 1, FRH_PACK_MSB16 = 2, FRH_PACK_MSB32 = 3 };

/* GpuFrame (run.rs:27-40). u16 / f32 input: `data` holds raster_width * cpp samples per row and
 * `data_len` counts samples. Packed input: `data` holds the payload from its first row,
 * `data_len` counts bytes (bytes past it read as 0), rows start `row_stride` bytes apart and hold
 * ceil(raster_width * cpp * bits / 8) bytes of samples of `bits` (1..16) bits in PackOrder
 * `order`. When `data_on_device` is 1, `data` is a device pointer and nothing is uploaded
 * (resident input). `linearization` (NULL or empty for none) maps integer samples
 * v -> table[min(v, len - 1)]. */
typedef struct frh_frame {
    uint32_t size;
    int32_t input;
    const void *data;
    uint64_t data_len;
    int32_t data_on_device;
    uint32_t raster_width, raster_height, cpp;
    frh_rect window;
    uint32_t origin_x, origin_y;
    const uint16_t *linearization;
    uint32_t linearization_len;
    uint32_t bits;
    int32_t order;
    uint64_t row_stride;
} frh_frame;

/* The FrameInfo fields the develop path reads (crates/flashraw-core/src/decoder.rs:25-74;
 * crates/flashraw-gpu/src/kernels.rs:275-345): CFA, black level (BlackLevel, types.rs:357-436)
 * and white level (WhiteLevel, types.rs:438-467). CFA colours are CfaColor codes
 * (R 0, G 1, B 2, types.rs:84-102), row-major, at most 8x8. */
typedef struct frh_frame_info {
    uint32_t size;
    uint32_t cfa_width, cfa_height; /* 0: no CFA */
    uint8_t cfa[64];
    uint32_t black_rows, black_cols, black_samples;
    const float *black_values;
    uint32_t black_values_len;
    const float *black_row_deltas;
    uint32_t black_row_deltas_len;
    const float *black_col_deltas;
    uint32_t black_col_deltas_len;
    const uint32_t *white_values;
    uint32_t white_values_len;
} frh_frame_info;

/* DevelopOutput (run.rs:96-111). 16-bit outputs are stored as 3 samples per pixel. */
enum frh_develop_output {
    FRH_OUTPUT_RGBA8 = 0,
    FRH_OUTPUT_LINEAR_RGB_F32 = 1,
    FRH_OUTPUT_RGB16 = 2,
    FRH_OUTPUT_LINEAR_RGB16 = 3,
    FRH_OUTPUT_LINEAR_RGB_F16 = 4,
    FRH_OUTPUT_NONE = -1
};

/* EdgeSpec (run.rs:141-160). The AHD form (`ahd` = 1, beta 0, period 2) uses the pattern and the
 * halo; its filter tables may be NULL. The edge-aware form (`ahd` = 0) takes Bayer (period 2) and
 * 6x6 (period 6, X-Trans) mosaics: with P = period it needs canon[P*P] (colour of each canonical
 * position), g0[P*P*49] (7x7 filter of the first green estimate per canonical position) and
 * dir[P*P*3*4*49] (7x7 filters per position, colour and direction), beta in [0, 1] and a halo of
 * at least 8; frame sample (x, y) sits at canonical position ((y + cy) % P) * P + (x + cx) % P.
 * The device copy of its tap lists is built once per distinct table contents and kept by the
 * context. Other periods return FRH_ERROR_NOT_AVAILABLE. */
typedef struct frh_edge_spec {
    uint32_t size;
    uint32_t period, cx, cy;
    float beta;
    uint32_t halo;
    const uint8_t *canon;
    uint32_t canon_len;
    const float *g0;
    uint32_t g0_len;
    const float *dir;
    uint32_t dir_len;
    int32_t ahd;
} frh_edge_spec;

/* GpuCurve (run.rs:181-188): the encoder tables of a non-sRGB output curve (CurveTables,
 * crates/flashraw/src/color.rs:343-367): `lut8` 8-bit codes at lut8_len - 1 segments of [0, 1]
 * (RGBA8 output, used when lut8_len > 1) and `lut16` 16-bit codes as f32 at 16385 points (sRGB16
 * output, used when lut16_len is 16385; the sRGB curve otherwise). */
typedef struct frh_curve {
    uint32_t size;
    const uint8_t *lut8;
    uint32_t lut8_len;
    const float *lut16;
    uint32_t lut16_len;
} frh_curve;

/* GpuTone (run.rs:190-205; ToneMap, crates/flashraw/src/tone.rs:81-207): the standard tone of
 * the display outputs (RGBA8, sRGB16). `table` holds table_len (value, step) pairs of the curve
 * at evenly spaced sqrt(x / xmax); in linear ProPhoto (`to_prophoto`, row-major, from linear
 * output RGB) the largest and smallest channel go through the curve and the pixel becomes
 * (fl - k lo) white + k p. table_len 0 means no tone. */
typedef struct frh_tone {
    uint32_t size;
    const float *table;
    uint32_t table_len;
    float xmax, scale;
    float to_prophoto[9];
    float white[3];
} frh_tone;

/* DevelopSpec (run.rs:115-135). `output2` (FRH_OUTPUT_NONE for none) asks the same pass to
 * store a second encoding, e.g. linear 16-bit and RGBA8 together. `curve` and `tone` (NULL for
 * the sRGB curve and no tone) apply to the display encodings only. `cancel` (NULL: the develop
 * cannot be stopped) is DecodeOptions::cancel (crates/flashraw/src/options.rs:501-508), read by
 * frh_run_develop and frh_develop_into; plans, batches and the serving loop ignore it. A spec
 * whose `size` ends before `cancel` (the layout without it) is accepted and has no token. */
typedef struct frh_develop_spec {
    uint32_t size;
    frh_rect crop;
    int32_t half_size;
    float wb[4];
    float cam_to_rgb[9];
    int32_t orientation; /* EXIF 1..8 */
    int32_t output;
    int32_t output2;
    int32_t has_levels;
    float levels[2]; /* scale, offset */
    const frh_edge_spec *edge;
    const frh_curve *curve;
    const frh_tone *tone;
    const frh_cancel_token *cancel;
} frh_develop_spec;

/* GpuTimings (run.rs:229-239). */
typedef struct frh_timings {
    double upload_ms;
    double compute_ms;
    double kernel_ms;
    int32_t has_kernel_ms;
    uint32_t bands;
} frh_timings;

/* ---- context (crates/flashraw-gpu/src/context.rs:775-842) ---- */

/* flashraw_gpu::available(). 0 when the environment variable FLASHRAW_GPU is 0, off, false, none
 * or cpu (any case; disabled_by_env, context.rs:734-747), without touching the device. */
int32_t frh_available(void);
/* flashraw_gpu::init() / context(): a context on HIP device `device`. FRH_ERROR_GPU_UNAVAILABLE
 * ("disabled by FLASHRAW_GPU") under the same environment switch (GpuContext::new,
 * context.rs:203-206). */
int32_t frh_context_create(int32_t device, frh_context **out);
void frh_context_destroy(frh_context *ctx);
/* GpuContext::info() / adapter_info(): a description of the device into `buf`. */
int32_t frh_context_info(const frh_context *ctx, char *buf, size_t len);
/* GpuContext::override_limits (context.rs:704): caps the bytes one band may bind, forcing
 * banded runs (0 restores the default, which never bands on this device). */
int32_t frh_set_limits(frh_context *ctx, uint64_t max_binding_bytes);
/* The message of the calling thread's last error (empty when none). */
const char *frh_last_error_message(void);

/* ---- input operations (run.rs:426 GpuContext::run with GpuOp::Unpack / GpuOp::Normalize,
 * run.rs:503-517) ---- */

/* GpuOp::Unpack: the frame's window as u16 samples (window.width * cpp per row), packed
 * payloads unpacked and integer samples passed through the linearisation curve; an image of
 * format U16 with `cpp` channels. f32 input is FRH_ERROR_INVALID_ARGUMENT. */
int32_t frh_run_unpack(frh_context *ctx, const frh_frame *frame, frh_image **out, frh_timings *timings);
/* GpuOp::Normalize: max((v - black) * scale, 0) of every sample of the window (Calib of each
 * component, row and column deltas; normalize.wgsl), an image of format F32 with `cpp`
 * channels. */
int32_t frh_run_normalize(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info, frh_image **out,
                          frh_timings *timings);
/* The same on `stream` for device-resident input into the caller's device buffer `dst` (4-byte
 * aligned for unpack). Normalising packed or linearised input unpacks into pooled scratch and
 * synchronises `stream` before returning it. */
int32_t frh_unpack_into(frh_context *ctx, const frh_frame *frame, void *dst, uint64_t dst_bytes, void *stream);
int32_t frh_normalize_into(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info, void *dst,
                           uint64_t dst_bytes, void *stream);

/* ---- develop (run.rs:426 GpuContext::run with GpuOp::Develop; run_blocking run.rs:1076) ---- */

/* Upload `frame` (unless resident), develop it and return the device-resident image. */
int32_t frh_run_develop(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info,
                        const frh_develop_spec *spec, frh_image **out, frh_image **out2, frh_timings *timings);

/* The same develop into caller-owned device buffers on `stream` (a hipStream_t; NULL is the
 * null stream), without synchronising: the resident form of GpuContext::run. `frame->data`
 * must be a device pointer. */
int32_t frh_develop_into(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info,
                         const frh_develop_spec *spec, void *dst, uint64_t dst_bytes, void *dst2,
                         uint64_t dst2_bytes, void *stream);

/* A validated develop (the descriptor work of run.rs:439-760 done once) that is launched many
 * times on resident buffers: repeated frames of one shape, batching and serving. The frame's
 * `data` is only used for its geometry; each launch names its own device input. */
typedef struct frh_plan frh_plan;
int32_t frh_plan_create(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info,
                        const frh_develop_spec *spec, frh_plan **out);
/* Develop device input `src` (the raster the frame describes) into `dst` (and `dst2` for
 * output2) on `stream` (NULL: the null stream), without synchronising. */
int32_t frh_plan_launch(const frh_plan *plan, const void *src, void *dst, void *dst2, void *stream);
/* Autotune the plan's tile shape: the strip height (frame rows one wave develops), the waves per
 * block (1, 2, 4, 8, and 16 for bilinear) and, for the bilinear RGBA8 develop of 4-aligned frames,
 * the frame columns per lane (2 or 4) of the row-walker kernels. Each candidate is timed with HIP
 * events on `stream` over `iters` launches on the given resident buffers; the fastest is kept in the
 * plan and remembered by the context for develops of the same geometry. *best (optional) receives
 * the strip height (frh_plan_tile_shape gives all three). No reference counterpart: the reference
 * fixes 16x16 workgroups (crates/flashraw-gpu/shaders/develop.wgsl:185). */
int32_t frh_plan_autotune(frh_plan *plan, const void *src, void *dst, void *dst2, void *stream, int32_t iters,
                          uint32_t *best);
/* The plan's tile shape: strip height, waves per block and frame columns per lane of its row-walker
 * launch (all 0 for the kernels without one: edge-aware, generic layouts, the LDS-tiled
 * alternates). Any pointer may be NULL. No reference counterpart (see frh_plan_autotune). */
int32_t frh_plan_tile_shape(const frh_plan *plan, uint32_t *rows, uint32_t *waves_per_block,
                            uint32_t *cols_per_lane);
/* Bytes of the plan's first and second outputs. */
uint64_t frh_plan_output_bytes(const frh_plan *plan, int32_t which);
/* No reference counterpart (the reference picks one WGSL pipeline per demosaic, run.rs:545-600):
 * which kernel family the plan's launches take. The edge-aware demosaic has two kernels that give
 * the same bytes: one for inputs that bound every intermediate value (dense 7x7 taps in registers)
 * and the general one (tap lists); FLASHRAW_EDGE_IMPL=general forces the latter, as
 * FLASHRAW_AHD_IMPL=tiled and FLASHRAW_BILINEAR_IMPL=tiled force the LDS-tiled demosaics. */
enum frh_kernel {
    FRH_KERNEL_BILINEAR_WALK = 0,
    FRH_KERNEL_BILINEAR_TILED = 1,
    FRH_KERNEL_AHD_WALK = 2,
    FRH_KERNEL_AHD_TILED = 3,
    FRH_KERNEL_EDGE_TAME = 4,
    FRH_KERNEL_EDGE_GENERAL = 5,
    FRH_KERNEL_GENERIC = 6
};
int32_t frh_plan_kernel(const frh_plan *plan);
void frh_plan_destroy(frh_plan *plan);
/* The context's own stream (a hipStream_t). */
void *frh_context_stream(const frh_context *ctx);

/* Output geometry of a develop: width, height (after orientation), channels and bytes. */
int32_t frh_develop_output_size(const frh_frame *frame, const frh_develop_spec *spec, int32_t output,
                                uint32_t *width, uint32_t *height, uint32_t *channels, uint64_t *bytes);

/* ---- batched develop ----
 * No reference counterpart (the reference develops one frame per GpuContext::run, run.rs:426):
 * many device-resident frames of any size, Bayer phase, bit depth, white balance and colour
 * matrix developed by one persistent launch per demosaic kind (AHD frames, bilinear frames),
 * which can be captured once into a HIP graph and replayed. Every frame's bytes are those of its
 * own single-frame develop. Frames use a black level of at most a 2x2 pattern without row or
 * column deltas, and one output each. */
typedef struct frh_batch frh_batch;
/* Validate n develops (each as frh_develop_into would) into the device buffers dsts[i]. */
int32_t frh_batch_create(frh_context *ctx, uint32_t n, const frh_frame *frames, const frh_frame_info *infos,
                         const frh_develop_spec *specs, void *const *dsts, frh_batch **out);
/* Enqueue the batch on `stream` (NULL: the null stream): the captured graph when there is one. */
int32_t frh_batch_launch(frh_batch *batch, void *stream);
/* Capture the batch's launch sequence on `stream` (not NULL) into a HIP graph that later
 * frh_batch_launch calls replay. */
int32_t frh_batch_capture(frh_batch *batch, void *stream);
/* Time candidate strip heights (the walkers' tile shape) on the batch itself, `iters` launches
 * each, and keep the fastest per demosaic kind (re-capturing a captured graph). */
int32_t frh_batch_autotune(frh_batch *batch, void *stream, int32_t iters, uint32_t *best_ahd_rows,
                           uint32_t *best_bilinear_rows);
/* Fix the strip heights (0: each frame's default). */
int32_t frh_batch_set_strip_rows(frh_batch *batch, uint32_t ahd_rows, uint32_t bilinear_rows);
/* Work items (strips and border groups) in the batch's lists. */
uint64_t frh_batch_items(const frh_batch *batch);
void frh_batch_destroy(frh_batch *batch);
/* No counterpart in the reference. For callers that cannot allocate device memory (code being
 * captured into a HIP graph, tight serving loops, harnesses that forbid hipMalloc): the batched
 * develop of frh_batch_create + frh_batch_launch in one call that allocates nothing. Descriptors,
 * work lists and list counters go to a fixed device array of the library (about 290 KB per GPU,
 * held until the process ends), filled by copy launches on `stream` (NULL: the null stream)
 * ahead of the list kernels; the call does not synchronise. At most 256 frames per call
 * (FRH_ERROR_LIMIT_EXCEEDED beyond). Frames are device-resident u16 Bayer mosaics at full size
 * with a black level of at most a 2x2 pattern, no deltas, linearisation, tone, curve or second
 * output, bilinear or AHD (FRH_ERROR_NOT_AVAILABLE otherwise, refused before anything reaches the
 * device). Strip heights ahd_rows / bilinear_rows (0: each frame's default). Every call on a GPU
 * shares that GPU's array, which keeps the last call's contents (the counters are cleared by each
 * call): calls that overlap on one GPU (different streams or threads) race; order them on one
 * stream. Each frame's bytes equal its frh_batch_launch / single-frame develop. */
int32_t frh_batch_develop_noalloc(frh_context *ctx, uint32_t n, const frh_frame *frames, const frh_frame_info *infos,
                                  const frh_develop_spec *specs, void *const *dsts, uint32_t ahd_rows,
                                  uint32_t bilinear_rows, void *stream);

/* ---- CFA-plane DCT block denoiser ----
 * No reference counterpart as an entry point: the reference denoises developed linear RGB
 * (GpuContext::denoise, crates/flashraw-gpu/src/denoise.rs:817). This one works on the mosaic's
 * four 2x2 phase planes with the reference's block transform (the orthonormal DCT-II of
 * crates/flashraw/src/denoise/mod.rs:621-635), hard threshold (dct_pass, mod.rs:921-974) and
 * block positions (positions(), mod.rs:877-884): 8x8 blocks at stride `step` plus the last
 * position, AC coefficients with |c| < threshold * sigma zeroed, sigma^2 = sig_a * max(block
 * mean, 0) + sig_b (DN), uniform overlap-add. */
enum frh_dn_precision {
    FRH_DN_F16 = 0,  /* MFMA, f16 operands, f32 accumulation (stride 4, sides multiples of 8) */
    FRH_DN_BF16 = 1, /* MFMA, bf16 operands, f32 accumulation (same limits) */
    FRH_DN_F32 = 2,  /* scalar f32, any stride 1..8, even sides; synchronises `stream` */
    FRH_DN_F16X3 = 3 /* MFMA, split f16 operands (value + rounded remainder, 3 products), f32 acc */
};
typedef struct frh_dn_spec {
    uint32_t size;
    uint32_t width, height, stride; /* mosaic size; stride in samples (0: width) */
    float black[4];                 /* black level of the planes, (y & 1) * 2 + (x & 1) */
    float sig_a, sig_b, threshold;  /* threshold: 2.7 in the reference's hard pass (mod.rs:47) */
    uint32_t step;
    int32_t precision;
    uint32_t seg_rows; /* MFMA walker: 8-row tile rows per wave (0: 12) */
} frh_dn_spec;
/* Denoise the device u16 mosaic `src` into `dst` (u16 mosaic, black added back, rounded and
 * clamped) and/or `dst_f32` (f32 mosaic in DN above black); either may be NULL, not both. The
 * output must not overlap the input. Asynchronous on `stream` (NULL: the null stream) for the
 * MFMA precisions. The f16 and bf16 precisions form their operands as a sample minus one sample
 * of the same 8-row step as 16-bit integers: exact for any mosaic of 15 bits or fewer, while
 * larger differences saturate at 32767; split f16 and f32 take any u16 samples. */
int32_t frh_denoise_cfa(frh_context *ctx, const frh_dn_spec *spec, const void *src, void *dst, float *dst_f32,
                        void *stream);

/* ---- luma / chroma noise reduction of developed linear RGB (GpuContext::denoise,
 * crates/flashraw-gpu/src/denoise.rs:817; the method of crates/flashraw/src/denoise/mod.rs) ---- */

/* PlaneSpec (denoise.rs:44-57): how one plane is denoised. */
typedef struct frh_nr_plane {
    uint32_t levels;    /* pyramid levels, the full image counting as one (1..8) */
    float gain;         /* multiplier of the noise deviations (thresholds and Wiener noise) */
    uint32_t steps[2];  /* block step of the two finest levels (2 beyond); 1, 2, 4 or 8 */
    int32_t hard_only0; /* the full-resolution level gets the hard-threshold pass only */
    float keep;         /* fraction of the difference to the input put back */
} frh_nr_plane;

/* DenoiseSpec (denoise.rs:59-80). The noise models the reference takes from the caller
 * (`models`, flashraw::denoise::gpu_models, mod.rs:196-222) are computed by the library from the
 * blocks it measures: frh_nr_model_from_stats, each coarse level capped by the finer one. */
typedef struct frh_nr_spec {
    uint32_t size;
    int32_t has_luma, has_chroma; /* 0: that plane unchanged */
    frh_nr_plane luma, chroma;
    uint32_t chroma_radius; /* guided filter window radius at half resolution */
    float hard;             /* hard threshold of the first pass, in noise deviations */
    uint32_t max_blocks;    /* most blocks the noise estimate measures per level */
    int32_t output;         /* frh_output */
    int32_t orientation;    /* EXIF 1..8, display outputs */
    const frh_tone *tone;
    const frh_curve *curve;
    const frh_cancel_token *cancel; /* NULL: cannot be stopped; checked between the stages */
} frh_nr_spec;

/* flashraw::denoise::gpu_spec (mod.rs:155-193): the spec of strengths `luma` and `chroma`
 * (clamped to [0, 1]; 0 leaves that part unchanged) for a w x h image. */
int32_t frh_nr_spec_for(uint32_t w, uint32_t h, float luma, float chroma, int32_t output, frh_nr_spec *spec);
/* GpuContext::denoise: denoise the linear RGB image `input` (format F32, 3 channels, as a develop
 * to FRH_OUTPUT_LINEAR_RGB_F32 returns it) and encode it to spec->output. */
int32_t frh_run_denoise(frh_context *ctx, const frh_image *input, const frh_nr_spec *spec, frh_image **out,
                        frh_timings *timings);
/* The same on device buffers: `rgb` (w * h * 3 floats) into `dst` (the output's bytes), on
 * `stream` (NULL: the context's); synchronises `stream` once to read the measured blocks back. */
int32_t frh_denoise_into(frh_context *ctx, const float *rgb, uint32_t w, uint32_t h, const frh_nr_spec *spec,
                         void *dst, uint64_t dst_bytes, void *stream);
/* Model::from_stats (mod.rs:803-848): the noise model of `count` measured blocks (64 floats each:
 * the luma mean, then the squared coefficients l * 8 + k) into at most 12 bins (`means`,
 * `sigmas` of 12 * 64 floats). FRH_ERROR_NOT_AVAILABLE when fewer than 32 blocks are finite. */
int32_t frh_nr_model_from_stats(const float *stats, uint32_t count, uint32_t *bins, float *means, float *sigmas);
/* The same model fitted on the context's device, as frh_run_denoise fits every level (the host
 * statement above when a bin would hold more than 4608 blocks). */
int32_t frh_nr_model_fit(frh_context *ctx, const float *stats, uint32_t count, uint32_t *bins, float *means,
                         float *sigmas);

/* ---- serving loop ----
 * No reference counterpart: a request loop over one context. Jobs (a host frame, its info and
 * spec, an optional CFA-plane denoise, a host output buffer) are admitted in arrival order while
 * their device buffers (input, denoised input, output) fit the device budget; uploads overlap
 * running batches; every resident job joins the next batched develop (continuous batching);
 * outputs are copied back and the buffers reused. Frames use a black level of at most a 2x2
 * pattern without deltas and one output; the frame's host data and the output buffer must stay
 * valid until the job is done (pinned host memory lets the copies run asynchronously). */
typedef struct frh_server frh_server;
typedef struct frh_server_config {
    uint32_t size;
    uint64_t device_budget_bytes;  /* bound on every device buffer the server holds (0: 64 GiB): the
                                    * buffers of admitted jobs, counted at their allocated sizes (size
                                    * classes at most 12.5 % above the request), and the idle ones kept
                                    * for reuse */
    uint32_t max_batch_frames;     /* frames per batched launch (0: 64) */
    uint32_t max_inflight_batches; /* batches queued on the device at once (0: 2) */
} frh_server_config;
typedef struct frh_job_result {
    uint32_t size;
    int32_t status;
    double queue_ms;      /* submitted -> admitted (device memory granted) */
    double upload_ms;     /* admitted -> input resident */
    double batch_wait_ms; /* resident -> its batch enqueued */
    double latency_ms;    /* submitted -> output in host memory */
    uint32_t batch_frames; /* frames in the job's batch */
} frh_job_result;
typedef struct frh_server_stats {
    uint32_t size;
    uint64_t jobs_submitted, jobs_done, jobs_failed, batches, batched_frames;
    uint64_t peak_device_bytes, device_budget_bytes; /* peak: allocated bytes of admitted jobs */
} frh_server_stats;
int32_t frh_server_create(frh_context *ctx, const frh_server_config *cfg, frh_server **out);
/* Queue a job; `denoise` NULL for none (otherwise its width/height/stride are the frame's). */
int32_t frh_server_submit(frh_server *server, const frh_frame *frame, const frh_frame_info *info,
                          const frh_develop_spec *spec, const frh_dn_spec *denoise, void *host_out,
                          uint64_t host_out_bytes, uint64_t *ticket);
/* Wait up to timeout_ms (< 0: forever) for a job; *done = 1 and `result` filled when finished
 * (the ticket is then released). */
int32_t frh_server_wait(frh_server *server, uint64_t ticket, int32_t timeout_ms, frh_job_result *result,
                        int32_t *done);
int32_t frh_server_get_stats(frh_server *server, frh_server_stats *stats);
/* Finish every submitted job, then release the server. */
void frh_server_destroy(frh_server *server);
/* Pinned host memory for frames and outputs (hipHostMalloc). */
void *frh_host_alloc(uint64_t bytes);
void frh_host_free(void *p);

/* ---- images (crates/flashraw-gpu/src/image.rs:52-306) ---- */

enum frh_pixel_format { FRH_FORMAT_RGBA8 = 0, FRH_FORMAT_F32 = 1, FRH_FORMAT_U16 = 2 };

uint32_t frh_image_width(const frh_image *img);
uint32_t frh_image_height(const frh_image *img);
uint32_t frh_image_channels(const frh_image *img);
int32_t frh_image_format(const frh_image *img);
const void *frh_image_device_ptr(const frh_image *img);
uint64_t frh_image_bytes(const frh_image *img);
/* GpuImage::download (image.rs:171-214): copy the pixels to host memory. */
int32_t frh_image_download(const frh_image *img, void *host, uint64_t bytes);
void frh_image_free(frh_image *img);

/* ---- tables shared with every develop (kernels.rs:348-371) ---- */

/* The 4097-entry sRGB 8-bit table the kernels encode with (process.rs:1903-1915). */
int32_t frh_srgb8_table(uint8_t *out, size_t len);

/* ---- diagnostics ----
 * No reference counterpart. The kernels' f32 square root (the AHD Lab and the edge-aware gradients,
 * f32::sqrt in the reference) is the hardware reciprocal square root refined by one fma step, which
 * is correctly rounded on gfx950 for every input the kernels give it. This runs every float in
 * [2^-96, 2^128), and +-0, through it on the context's device and reports in *mismatches how many
 * differ from the correctly rounded square root (0 on a device where the develop is exact). */
int32_t frh_check_sqrt(frh_context *ctx, uint64_t *mismatches);
/* No reference counterpart. The context's cache of per-frame device tables (calibration, CFA
 * sites, black-level deltas, linearisation, tone and curve tables, edge-aware tap lists): entries
 * and device bytes. Hits compare the full table contents; the cache keeps at most 64 MiB and 4096
 * entries, freeing the least recently used tables no prepared plan holds (after synchronising the
 * device, so that no queued launch still reads them) when a new table would not fit. */
int32_t frh_context_table_cache(const frh_context *ctx, uint64_t *entries, uint64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
