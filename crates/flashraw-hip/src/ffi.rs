//! Raw declarations of `libflashraw_hip.so` (the C ABI in `include/flashraw_hip.h` of the
//! MI355X build). Field order, types and sizes are those of the header; the CPU test
//! `tests/test_rust_crate.py` of that build checks every `#[repr(C)]` struct here against the
//! layout gcc gives the header, and every function against the header and the library's exports.
#![allow(non_camel_case_types, dead_code)]

use std::os::raw::{c_char, c_void};

pub const FRH_OK: i32 = 0;
pub const FRH_ERROR_GPU_UNAVAILABLE: i32 = 6;
pub const FRH_ERROR_OUT_OF_MEMORY: i32 = 7;
pub const FRH_ERROR_LIMIT_EXCEEDED: i32 = 8;
pub const FRH_ERROR_NOT_AVAILABLE: i32 = 9;
pub const FRH_ERROR_INVALID_ARGUMENT: i32 = 10;
pub const FRH_ERROR_INTERNAL: i32 = 12;
pub const FRH_ERROR_CANCELLED: i32 = 14;

pub const FRH_INPUT_U16: i32 = 0;
pub const FRH_INPUT_F32: i32 = 1;
pub const FRH_INPUT_PACKED: i32 = 2;

pub const FRH_PACK_MSB_BYTES: i32 = 0;
pub const FRH_PACK_LSB_BYTES: i32 = 1;
pub const FRH_PACK_MSB16: i32 = 2;
pub const FRH_PACK_MSB32: i32 = 3;

pub const FRH_OUTPUT_RGBA8: i32 = 0;
pub const FRH_OUTPUT_LINEAR_RGB_F32: i32 = 1;
pub const FRH_OUTPUT_RGB16: i32 = 2;
pub const FRH_OUTPUT_LINEAR_RGB16: i32 = 3;
pub const FRH_OUTPUT_LINEAR_RGB_F16: i32 = 4;
pub const FRH_OUTPUT_NONE: i32 = -1;

pub const FRH_FORMAT_RGBA8: i32 = 0;
pub const FRH_FORMAT_F32: i32 = 1;
pub const FRH_FORMAT_U16: i32 = 2;

pub const FRH_DN_F16: i32 = 0;
pub const FRH_DN_BF16: i32 = 1;
pub const FRH_DN_F32: i32 = 2;
pub const FRH_DN_F16X3: i32 = 3;

pub const FRH_KERNEL_BILINEAR_WALK: i32 = 0;
pub const FRH_KERNEL_BILINEAR_TILED: i32 = 1;
pub const FRH_KERNEL_AHD_WALK: i32 = 2;
pub const FRH_KERNEL_AHD_TILED: i32 = 3;
pub const FRH_KERNEL_EDGE_TAME: i32 = 4;
pub const FRH_KERNEL_EDGE_GENERAL: i32 = 5;
pub const FRH_KERNEL_GENERIC: i32 = 6;

macro_rules! opaque {
    ($($name:ident),*) => {
        $(#[repr(C)]
        pub struct $name {
            _private: [u8; 0],
        })*
    };
}
opaque!(frh_context, frh_image, frh_cancel_token, frh_plan, frh_batch, frh_server);

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct frh_rect {
    pub x: u32,
    pub y: u32,
    pub width: u32,
    pub height: u32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct frh_frame {
    pub size: u32,
    pub input: i32,
    pub data: *const c_void,
    pub data_len: u64,
    pub data_on_device: i32,
    pub raster_width: u32,
    pub raster_height: u32,
    pub cpp: u32,
    pub window: frh_rect,
    pub origin_x: u32,
    pub origin_y: u32,
    pub linearization: *const u16,
    pub linearization_len: u32,
    pub bits: u32,
    pub order: i32,
    pub row_stride: u64,
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct frh_frame_info {
    pub size: u32,
    pub cfa_width: u32,
    pub cfa_height: u32,
    pub cfa: [u8; 64],
    pub black_rows: u32,
    pub black_cols: u32,
    pub black_samples: u32,
    pub black_values: *const f32,
    pub black_values_len: u32,
    pub black_row_deltas: *const f32,
    pub black_row_deltas_len: u32,
    pub black_col_deltas: *const f32,
    pub black_col_deltas_len: u32,
    pub white_values: *const u32,
    pub white_values_len: u32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct frh_edge_spec {
    pub size: u32,
    pub period: u32,
    pub cx: u32,
    pub cy: u32,
    pub beta: f32,
    pub halo: u32,
    pub canon: *const u8,
    pub canon_len: u32,
    pub g0: *const f32,
    pub g0_len: u32,
    pub dir: *const f32,
    pub dir_len: u32,
    pub ahd: i32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct frh_curve {
    pub size: u32,
    pub lut8: *const u8,
    pub lut8_len: u32,
    pub lut16: *const f32,
    pub lut16_len: u32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct frh_tone {
    pub size: u32,
    pub table: *const f32,
    pub table_len: u32,
    pub xmax: f32,
    pub scale: f32,
    pub to_prophoto: [f32; 9],
    pub white: [f32; 3],
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct frh_develop_spec {
    pub size: u32,
    pub crop: frh_rect,
    pub half_size: i32,
    pub wb: [f32; 4],
    pub cam_to_rgb: [f32; 9],
    pub orientation: i32,
    pub output: i32,
    pub output2: i32,
    pub has_levels: i32,
    pub levels: [f32; 2],
    pub edge: *const frh_edge_spec,
    pub curve: *const frh_curve,
    pub tone: *const frh_tone,
    pub cancel: *const frh_cancel_token,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct frh_timings {
    pub upload_ms: f64,
    pub compute_ms: f64,
    pub kernel_ms: f64,
    pub has_kernel_ms: i32,
    pub bands: u32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct frh_dn_spec {
    pub size: u32,
    pub width: u32,
    pub height: u32,
    pub stride: u32,
    pub black: [f32; 4],
    pub sig_a: f32,
    pub sig_b: f32,
    pub threshold: f32,
    pub step: u32,
    pub precision: i32,
    pub seg_rows: u32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct frh_nr_plane {
    pub levels: u32,
    pub gain: f32,
    pub steps: [u32; 2],
    pub hard_only0: i32,
    pub keep: f32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy)]
pub struct frh_nr_spec {
    pub size: u32,
    pub has_luma: i32,
    pub has_chroma: i32,
    pub luma: frh_nr_plane,
    pub chroma: frh_nr_plane,
    pub chroma_radius: u32,
    pub hard: f32,
    pub max_blocks: u32,
    pub output: i32,
    pub orientation: i32,
    pub tone: *const frh_tone,
    pub curve: *const frh_curve,
    pub cancel: *const frh_cancel_token,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct frh_server_config {
    pub size: u32,
    pub device_budget_bytes: u64,
    pub max_batch_frames: u32,
    pub max_inflight_batches: u32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct frh_job_result {
    pub size: u32,
    pub status: i32,
    pub queue_ms: f64,
    pub upload_ms: f64,
    pub batch_wait_ms: f64,
    pub latency_ms: f64,
    pub batch_frames: u32,
}

#[repr(C)]
#[derive(Debug, Clone, Copy, Default)]
pub struct frh_server_stats {
    pub size: u32,
    pub jobs_submitted: u64,
    pub jobs_done: u64,
    pub jobs_failed: u64,
    pub batches: u64,
    pub batched_frames: u64,
    pub peak_device_bytes: u64,
    pub device_budget_bytes: u64,
}

// linked by build.rs (cargo:rustc-link-lib=dylib=flashraw_hip)
unsafe extern "C" {
    pub fn frh_cancel_token_new() -> *mut frh_cancel_token;
    pub fn frh_cancel_token_with_poll(poll: Option<unsafe extern "C" fn(user: *mut c_void) -> i32>, user: *mut c_void) -> *mut frh_cancel_token;
    pub fn frh_cancel_token_cancel(t: *const frh_cancel_token);
    pub fn frh_cancel_token_is_cancelled(t: *const frh_cancel_token) -> i32;
    pub fn frh_cancel_token_free(t: *mut frh_cancel_token);
    pub fn frh_available() -> i32;
    pub fn frh_context_create(device: i32, out: *mut *mut frh_context) -> i32;
    pub fn frh_context_destroy(ctx: *mut frh_context);
    pub fn frh_context_info(ctx: *const frh_context, buf: *mut c_char, len: usize) -> i32;
    pub fn frh_set_limits(ctx: *mut frh_context, max_binding_bytes: u64) -> i32;
    pub fn frh_last_error_message() -> *const c_char;
    pub fn frh_run_unpack(ctx: *mut frh_context, frame: *const frh_frame, out: *mut *mut frh_image, timings: *mut frh_timings) -> i32;
    pub fn frh_run_normalize(ctx: *mut frh_context, frame: *const frh_frame, info: *const frh_frame_info, out: *mut *mut frh_image, timings: *mut frh_timings) -> i32;
    pub fn frh_unpack_into(ctx: *mut frh_context, frame: *const frh_frame, dst: *mut c_void, dst_bytes: u64, stream: *mut c_void) -> i32;
    pub fn frh_normalize_into(ctx: *mut frh_context, frame: *const frh_frame, info: *const frh_frame_info, dst: *mut c_void, dst_bytes: u64, stream: *mut c_void) -> i32;
    pub fn frh_run_develop(ctx: *mut frh_context, frame: *const frh_frame, info: *const frh_frame_info, spec: *const frh_develop_spec, out: *mut *mut frh_image, out2: *mut *mut frh_image, timings: *mut frh_timings) -> i32;
    pub fn frh_develop_into(ctx: *mut frh_context, frame: *const frh_frame, info: *const frh_frame_info, spec: *const frh_develop_spec, dst: *mut c_void, dst_bytes: u64, dst2: *mut c_void, dst2_bytes: u64, stream: *mut c_void) -> i32;
    pub fn frh_plan_create(ctx: *mut frh_context, frame: *const frh_frame, info: *const frh_frame_info, spec: *const frh_develop_spec, out: *mut *mut frh_plan) -> i32;
    pub fn frh_plan_launch(plan: *const frh_plan, src: *const c_void, dst: *mut c_void, dst2: *mut c_void, stream: *mut c_void) -> i32;
    pub fn frh_plan_autotune(plan: *mut frh_plan, src: *const c_void, dst: *mut c_void, dst2: *mut c_void, stream: *mut c_void, iters: i32, best: *mut u32) -> i32;
    pub fn frh_plan_tile_shape(plan: *const frh_plan, rows: *mut u32, waves_per_block: *mut u32, cols_per_lane: *mut u32) -> i32;
    pub fn frh_plan_output_bytes(plan: *const frh_plan, which: i32) -> u64;
    pub fn frh_plan_kernel(plan: *const frh_plan) -> i32;
    pub fn frh_plan_destroy(plan: *mut frh_plan);
    pub fn frh_context_stream(ctx: *const frh_context) -> *mut c_void;
    pub fn frh_develop_output_size(frame: *const frh_frame, spec: *const frh_develop_spec, output: i32, width: *mut u32, height: *mut u32, channels: *mut u32, bytes: *mut u64) -> i32;
    pub fn frh_batch_create(ctx: *mut frh_context, n: u32, frames: *const frh_frame, infos: *const frh_frame_info, specs: *const frh_develop_spec, dsts: *const *mut c_void, out: *mut *mut frh_batch) -> i32;
    pub fn frh_batch_launch(batch: *mut frh_batch, stream: *mut c_void) -> i32;
    pub fn frh_batch_capture(batch: *mut frh_batch, stream: *mut c_void) -> i32;
    pub fn frh_batch_autotune(batch: *mut frh_batch, stream: *mut c_void, iters: i32, best_ahd_rows: *mut u32, best_bilinear_rows: *mut u32) -> i32;
    pub fn frh_batch_set_strip_rows(batch: *mut frh_batch, ahd_rows: u32, bilinear_rows: u32) -> i32;
    pub fn frh_batch_items(batch: *const frh_batch) -> u64;
    pub fn frh_batch_destroy(batch: *mut frh_batch);
    pub fn frh_batch_develop_noalloc(ctx: *mut frh_context, n: u32, frames: *const frh_frame, infos: *const frh_frame_info, specs: *const frh_develop_spec, dsts: *const *mut c_void, ahd_rows: u32, bilinear_rows: u32, stream: *mut c_void) -> i32;
    pub fn frh_denoise_cfa(ctx: *mut frh_context, spec: *const frh_dn_spec, src: *const c_void, dst: *mut c_void, dst_f32: *mut f32, stream: *mut c_void) -> i32;
    pub fn frh_nr_spec_for(w: u32, h: u32, luma: f32, chroma: f32, output: i32, spec: *mut frh_nr_spec) -> i32;
    pub fn frh_run_denoise(ctx: *mut frh_context, input: *const frh_image, spec: *const frh_nr_spec, out: *mut *mut frh_image, timings: *mut frh_timings) -> i32;
    pub fn frh_denoise_into(ctx: *mut frh_context, rgb: *const f32, w: u32, h: u32, spec: *const frh_nr_spec, dst: *mut c_void, dst_bytes: u64, stream: *mut c_void) -> i32;
    pub fn frh_nr_model_from_stats(stats: *const f32, count: u32, bins: *mut u32, means: *mut f32, sigmas: *mut f32) -> i32;
    pub fn frh_nr_model_fit(ctx: *mut frh_context, stats: *const f32, count: u32, bins: *mut u32, means: *mut f32, sigmas: *mut f32) -> i32;
    pub fn frh_server_create(ctx: *mut frh_context, cfg: *const frh_server_config, out: *mut *mut frh_server) -> i32;
    pub fn frh_server_submit(server: *mut frh_server, frame: *const frh_frame, info: *const frh_frame_info, spec: *const frh_develop_spec, denoise: *const frh_dn_spec, host_out: *mut c_void, host_out_bytes: u64, ticket: *mut u64) -> i32;
    pub fn frh_server_wait(server: *mut frh_server, ticket: u64, timeout_ms: i32, result: *mut frh_job_result, done: *mut i32) -> i32;
    pub fn frh_server_get_stats(server: *mut frh_server, stats: *mut frh_server_stats) -> i32;
    pub fn frh_server_destroy(server: *mut frh_server);
    pub fn frh_host_alloc(bytes: u64) -> *mut c_void;
    pub fn frh_host_free(p: *mut c_void);
    pub fn frh_image_width(img: *const frh_image) -> u32;
    pub fn frh_image_height(img: *const frh_image) -> u32;
    pub fn frh_image_channels(img: *const frh_image) -> u32;
    pub fn frh_image_format(img: *const frh_image) -> i32;
    pub fn frh_image_device_ptr(img: *const frh_image) -> *const c_void;
    pub fn frh_image_bytes(img: *const frh_image) -> u64;
    pub fn frh_image_download(img: *const frh_image, host: *mut c_void, bytes: u64) -> i32;
    pub fn frh_image_free(img: *mut frh_image);
    pub fn frh_srgb8_table(out: *mut u8, len: usize) -> i32;
    pub fn frh_check_sqrt(ctx: *mut frh_context, mismatches: *mut u64) -> i32;
    pub fn frh_context_table_cache(ctx: *const frh_context, entries: *mut u64, bytes: *mut u64) -> i32;
}
