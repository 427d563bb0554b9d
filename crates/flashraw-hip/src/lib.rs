//! # flashraw-hip
//!
//! The MI355X (gfx950, ROCm/HIP) backend of FlashRaw's GPU develop path, over
//! `libflashraw_hip.so`. It offers the public surface of `flashraw-gpu`
//! (crates/flashraw-gpu/src/lib.rs:43-51): [`context`], [`init`], [`available`],
//! [`GpuContext::run`] / [`GpuContext::run_blocking`] with [`GpuOp`] (unpack, normalise,
//! develop), [`GpuContext::denoise`] / [`GpuContext::denoise_blocking`], [`GpuImage`] and its
//! download, [`GpuTimings`], with the same types, field names and error kinds, so the
//! `flashraw` facade can select it where it selects `flashraw-gpu` (crates/flashraw/src/gpu.rs
//! and policy.rs). Images live in HIP device memory; [`GpuImage::device_ptr`] takes the place
//! of the wgpu buffer accessors.
//!
//! What differs, by construction of the device: images are never tiled (288 GB of HBM holds
//! any frame whole), the noise models of [`DenoiseSpec`] are fitted by the library on the
//! device (the same statistics as `Model::from_stats`, bit for bit) instead of by the
//! `models` callback, and calls complete before they return (the `async` forms are ready
//! futures).
//!
//! ```no_run
//! # fn main() -> flashraw_core::Result<()> {
//! use flashraw_hip::{GpuFrame, GpuOp};
//! let ctx = flashraw_hip::context()?;
//! println!("{}", ctx.info());
//! # let (data, info) = (vec![0u16; 16], flashraw_core::FrameInfo::new(4, 4, 12));
//! let run = ctx.run_blocking(&GpuFrame::from_u16(&data, 4, 4, 1), &info, &GpuOp::Normalize)?;
//! let (pixels, _ms) = run.image.download()?;
//! # Ok(()) }
//! ```

pub mod ffi;

use flashraw_core::{CancelToken, Error, FrameInfo, Orientation, PackOrder, PackedPayload, Rect, Result};
use std::ffi::CStr;
use std::os::raw::c_void;
use std::sync::{Arc, Mutex};

// ---- errors ----

/// The calling thread's last library error as the reference's error kind for `status`.
fn error(status: i32) -> Error {
    // SAFETY: the library returns a NUL-terminated thread-local string (never NULL).
    let msg = unsafe { CStr::from_ptr(ffi::frh_last_error_message()) }.to_string_lossy().into_owned();
    match status {
        ffi::FRH_ERROR_OUT_OF_MEMORY => Error::OutOfMemory(msg),
        ffi::FRH_ERROR_LIMIT_EXCEEDED => Error::LimitExceeded(msg),
        ffi::FRH_ERROR_NOT_AVAILABLE => Error::NotAvailable(msg),
        ffi::FRH_ERROR_INVALID_ARGUMENT => Error::InvalidArgument(msg),
        ffi::FRH_ERROR_CANCELLED => Error::Cancelled,
        // FRH_ERROR_GPU_UNAVAILABLE, and device failures (FRH_ERROR_INTERNAL)
        _ => Error::GpuUnavailable(msg),
    }
}

fn check(status: i32) -> Result<()> {
    if status == ffi::FRH_OK { Ok(()) } else { Err(error(status)) }
}

// ---- adapter and context (crates/flashraw-gpu/src/context.rs) ----

/// Coarse adapter class used by backend selection (context.rs:13-24).
#[derive(Debug, Clone, Copy, PartialEq, Eq, Hash)]
pub enum AdapterClass {
    Discrete,
    Integrated,
    Virtual,
    Software,
    Unknown,
}

impl AdapterClass {
    pub fn name(self) -> &'static str {
        match self {
            AdapterClass::Discrete => "discrete",
            AdapterClass::Integrated => "integrated",
            AdapterClass::Virtual => "virtual",
            AdapterClass::Software => "software",
            AdapterClass::Unknown => "unknown",
        }
    }
}

/// Description of the device in use (context.rs:38-48). An MI355X is a discrete adapter whose
/// kernels are timed on the device.
#[derive(Debug, Clone, PartialEq)]
pub struct GpuAdapterInfo {
    pub name: String,
    pub backend: String,
    pub driver: String,
    pub class: AdapterClass,
    pub timestamps: bool,
    pub max_buffer_size: u64,
    pub max_storage_binding: u64,
}

impl std::fmt::Display for GpuAdapterInfo {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "{} ({}, {})", self.name, self.backend, self.class.name())
    }
}

/// A HIP device context: tables, pools and streams of one device (context.rs:115-120).
pub struct GpuContext {
    raw: *mut ffi::frh_context,
    info: GpuAdapterInfo,
    max_binding: Mutex<Option<u64>>,
}

// SAFETY: every frh_* entry point on a context is thread-safe (the library serialises its
// shared state with the context's mutex).
unsafe impl Send for GpuContext {}
unsafe impl Sync for GpuContext {}

impl Drop for GpuContext {
    fn drop(&mut self) {
        // SAFETY: raw came from frh_context_create and is released once.
        unsafe { ffi::frh_context_destroy(self.raw) }
    }
}

impl GpuContext {
    /// A context on HIP device 0 (`HIP_VISIBLE_DEVICES` picks the card), as
    /// `GpuContext::new` (context.rs:203-246). `FLASHRAW_GPU=0|off|false|none|cpu` makes it
    /// `Error::GpuUnavailable`, as there.
    pub async fn new() -> Result<Self> {
        Self::with_device(0)
    }

    /// A context on HIP device `device`.
    pub fn with_device(device: i32) -> Result<Self> {
        let mut raw = std::ptr::null_mut();
        // SAFETY: out-pointer to a local.
        check(unsafe { ffi::frh_context_create(device, &mut raw) })?;
        let mut buf = vec![0u8; 256];
        // SAFETY: buf holds buf.len() bytes.
        unsafe { ffi::frh_context_info(raw, buf.as_mut_ptr().cast(), buf.len()) };
        let name = CStr::from_bytes_until_nul(&buf).map(|c| c.to_string_lossy().into_owned()).unwrap_or_default();
        let info = GpuAdapterInfo {
            name,
            backend: "HIP".into(),
            driver: "ROCm".into(),
            class: AdapterClass::Discrete,
            timestamps: true,
            max_buffer_size: u64::MAX,
            max_storage_binding: u64::MAX,
        };
        Ok(GpuContext { raw, info, max_binding: Mutex::new(None) })
    }

    pub fn info(&self) -> &GpuAdapterInfo {
        &self.info
    }

    /// Device memory is not shared with the host.
    pub fn unified_memory(&self) -> bool {
        false
    }

    /// The library's pools and table cache are bounded; nothing to release early.
    pub fn trim(&self) {}

    /// Every call completes before it returns: nothing is pending.
    pub async fn wait_idle(&self) -> Result<()> {
        Ok(())
    }

    /// Caps the bytes one band may bind, forcing banded runs (context.rs:704-709; the buffer
    /// cap has no counterpart: device buffers are not limited below the memory size).
    pub fn override_limits(&self, max_binding: Option<u64>, _max_buffer: Option<u64>) {
        *self.max_binding.lock().unwrap() = max_binding;
        // SAFETY: raw is a live context.
        unsafe { ffi::frh_set_limits(self.raw, max_binding.unwrap_or(0)) };
    }

    pub fn max_binding_bytes(&self) -> u64 {
        self.max_binding.lock().unwrap().unwrap_or(self.info.max_storage_binding)
    }

    pub fn max_buffer_bytes(&self) -> u64 {
        self.info.max_buffer_size
    }

    /// The raw context, for the `ffi` entry points (plans, batches, the serving loop).
    pub fn as_raw(&self) -> *mut ffi::frh_context {
        self.raw
    }

    /// Run `op` on `frame` (run.rs:426 `GpuContext::run`): upload, compute, and a
    /// device-resident image. Completes before returning.
    pub async fn run(self: &Arc<Self>, frame: &GpuFrame<'_>, info: &FrameInfo, op: &GpuOp) -> Result<GpuRun> {
        self.run_with(frame, info, op, None)
    }

    /// Blocking [`GpuContext::run`] (run.rs:1076).
    pub fn run_blocking(self: &Arc<Self>, frame: &GpuFrame<'_>, info: &FrameInfo, op: &GpuOp) -> Result<GpuRun> {
        self.run_with(frame, info, op, None)
    }

    /// [`GpuContext::run_blocking`] that checks `cancel` between its stages and returns
    /// `Error::Cancelled` once it is cancelled (the checks the facade makes around a GPU develop,
    /// crates/flashraw/src/gpu.rs:292-299, made inside the call).
    pub fn run_cancellable(
        self: &Arc<Self>,
        frame: &GpuFrame<'_>,
        info: &FrameInfo,
        op: &GpuOp,
        cancel: &CancelToken,
    ) -> Result<GpuRun> {
        let token = Token::bridge(cancel)?;
        self.run_with(frame, info, op, Some(&*token))
    }

    fn run_with(self: &Arc<Self>, frame: &GpuFrame<'_>, info: &FrameInfo, op: &GpuOp, cancel: Option<&Token>) -> Result<GpuRun> {
        let f = frame.to_ffi();
        let fi = InfoFfi::new(info);
        let mut out = std::ptr::null_mut();
        let mut t = ffi::frh_timings::default();
        // SAFETY: every pointer passed is to a local (or into `frame`/`info`, which outlive the
        // call); the library copies what it keeps.
        let status = unsafe {
            match op {
                GpuOp::Unpack => ffi::frh_run_unpack(self.raw, &f, &mut out, &mut t),
                GpuOp::Normalize => ffi::frh_run_normalize(self.raw, &f, &fi.raw, &mut out, &mut t),
                GpuOp::Develop(spec) => {
                    let s = SpecFfi::new(spec, cancel);
                    ffi::frh_run_develop(self.raw, &f, &fi.raw, s.raw(), &mut out, std::ptr::null_mut(), &mut t)
                }
            }
        };
        check(status)?;
        Ok(GpuRun { image: GpuImage::wrap(self.clone(), out), timings: t.into() })
    }

    /// Denoise the linear RGB image `input` (a develop to [`DevelopOutput::LinearRgbF32`]) and
    /// encode it to `spec.output` (denoise.rs:817 `GpuContext::denoise`).
    pub async fn denoise(self: &Arc<Self>, input: &GpuImage, spec: &DenoiseSpec<'_>) -> Result<GpuRun> {
        self.denoise_with(input, spec, None)
    }

    /// Blocking [`GpuContext::denoise`] (denoise.rs:1180).
    pub fn denoise_blocking(self: &Arc<Self>, input: &GpuImage, spec: &DenoiseSpec<'_>) -> Result<GpuRun> {
        self.denoise_with(input, spec, None)
    }

    /// [`GpuContext::denoise_blocking`] that checks `cancel` between its stages.
    pub fn denoise_cancellable(self: &Arc<Self>, input: &GpuImage, spec: &DenoiseSpec<'_>, cancel: &CancelToken) -> Result<GpuRun> {
        let token = Token::bridge(cancel)?;
        self.denoise_with(input, spec, Some(&*token))
    }

    fn denoise_with(self: &Arc<Self>, input: &GpuImage, spec: &DenoiseSpec<'_>, cancel: Option<&Token>) -> Result<GpuRun> {
        let tone = spec.tone.map(|t| tone_ffi(&t));
        let curve = spec.curve.map(|c| curve_ffi(&c));
        let plane = |p: &Option<PlaneSpec>| {
            p.map(|p| ffi::frh_nr_plane {
                levels: p.levels,
                gain: p.gain,
                steps: p.steps,
                hard_only0: p.hard_only0 as i32,
                keep: p.keep,
            })
        };
        let s = ffi::frh_nr_spec {
            size: std::mem::size_of::<ffi::frh_nr_spec>() as u32,
            has_luma: spec.luma.is_some() as i32,
            has_chroma: spec.chroma.is_some() as i32,
            luma: plane(&spec.luma).unwrap_or_default(),
            chroma: plane(&spec.chroma).unwrap_or_default(),
            chroma_radius: spec.chroma_radius,
            hard: spec.hard,
            max_blocks: spec.max_blocks,
            output: spec.output.code(),
            orientation: spec.orientation as i32,
            tone: tone.as_ref().map_or(std::ptr::null(), |t| t as *const _),
            curve: curve.as_ref().map_or(std::ptr::null(), |c| c as *const _),
            cancel: cancel.map_or(std::ptr::null(), |t| t.raw as *const _),
        };
        let mut out = std::ptr::null_mut();
        let mut t = ffi::frh_timings::default();
        // SAFETY: pointers to locals and to a live image of this context's device.
        check(unsafe { ffi::frh_run_denoise(self.raw, input.raw, &s, &mut out, &mut t) })?;
        Ok(GpuRun { image: GpuImage::wrap(self.clone(), out), timings: t.into() })
    }
}

/// A library cancel token that reads a `CancelToken`'s flag on every check.
struct Token {
    raw: *mut ffi::frh_cancel_token,
    _keep: CancelToken,
}

unsafe extern "C" fn poll_flag(user: *mut c_void) -> i32 {
    // SAFETY: user is the CancelToken held by the Token for the life of the library token.
    let t = unsafe { &*(user as *const CancelToken) };
    t.is_cancelled() as i32
}

impl Token {
    fn bridge(cancel: &CancelToken) -> Result<Box<Token>> {
        let mut b = Box::new(Token { raw: std::ptr::null_mut(), _keep: cancel.clone() });
        let user = &b._keep as *const CancelToken as *mut c_void;
        // SAFETY: `user` points into the box, which outlives the library token (freed in Drop).
        b.raw = unsafe { ffi::frh_cancel_token_with_poll(Some(poll_flag), user) };
        if b.raw.is_null() {
            return Err(Error::OutOfMemory("cancel token".into()));
        }
        if cancel.is_cancelled() {
            // SAFETY: live token.
            unsafe { ffi::frh_cancel_token_cancel(b.raw) };
        }
        Ok(b)
    }
}

impl Drop for Token {
    fn drop(&mut self) {
        // SAFETY: created by frh_cancel_token_with_poll, freed once, no call holds it any more.
        unsafe { ffi::frh_cancel_token_free(self.raw) }
    }
}

// ---- the process-wide context (context.rs:775-842) ----

static CONTEXT: Mutex<Option<Arc<GpuContext>>> = Mutex::new(None);

/// The shared context, created on first use.
pub async fn init() -> Result<Arc<GpuContext>> {
    context()
}

/// Blocking [`init`].
pub fn context() -> Result<Arc<GpuContext>> {
    let mut g = CONTEXT.lock().unwrap();
    if let Some(c) = g.as_ref() {
        return Ok(c.clone());
    }
    let c = Arc::new(GpuContext::with_device(0)?);
    *g = Some(c.clone());
    Ok(c)
}

/// The shared context if one has been created.
pub fn try_context() -> Option<Arc<GpuContext>> {
    CONTEXT.lock().unwrap().clone()
}

/// Install `ctx` as the shared context.
pub fn set_context(ctx: Arc<GpuContext>) {
    *CONTEXT.lock().unwrap() = Some(ctx);
}

/// Whether a device can be used (false under the `FLASHRAW_GPU` off switch), without creating a
/// context.
pub fn available() -> bool {
    // SAFETY: no arguments.
    unsafe { ffi::frh_available() != 0 }
}

/// The shared context's adapter description, when it exists or can be created.
pub fn adapter_info() -> Option<GpuAdapterInfo> {
    context().ok().map(|c| c.info().clone())
}

/// Block on a future whose work is already done (the `async` forms here complete before they
/// return); kept for callers of `flashraw_gpu::block_on`.
pub fn block_on<F: std::future::Future>(f: F) -> F::Output {
    use std::task::{Context, Poll, RawWaker, RawWakerVTable, Waker};
    fn noop(_: *const ()) {}
    fn clone(p: *const ()) -> RawWaker {
        RawWaker::new(p, &VTABLE)
    }
    static VTABLE: RawWakerVTable = RawWakerVTable::new(clone, noop, noop, noop);
    // SAFETY: the vtable's functions do nothing with the data pointer.
    let waker = unsafe { Waker::from_raw(RawWaker::new(std::ptr::null(), &VTABLE)) };
    let mut cx = Context::from_waker(&waker);
    let mut f = std::pin::pin!(f);
    loop {
        if let Poll::Ready(v) = f.as_mut().poll(&mut cx) {
            return v;
        }
        std::thread::yield_now();
    }
}

// ---- frames and specs (crates/flashraw-gpu/src/run.rs) ----

/// Input raster (run.rs:11-24).
#[derive(Debug, Clone, Copy)]
pub enum GpuInput<'a> {
    Packed { data: &'a [u8], bits: u32, order: PackOrder, row_stride: usize },
    U16(&'a [u16]),
    F32(&'a [f32]),
}

/// A frame to process: an input raster plus the window and coordinate anchors (run.rs:27-40).
#[derive(Debug, Clone, Copy)]
pub struct GpuFrame<'a> {
    pub input: GpuInput<'a>,
    pub raster_width: u32,
    pub raster_height: u32,
    pub cpp: u32,
    pub window: Rect,
    pub origin: (u32, u32),
    pub linearization: Option<&'a [u16]>,
}

impl<'a> GpuFrame<'a> {
    pub fn from_packed(p: &PackedPayload<'a>) -> Self {
        GpuFrame {
            input: GpuInput::Packed { data: p.data, bits: p.bits, order: p.order, row_stride: p.row_stride },
            raster_width: p.width,
            raster_height: p.height,
            cpp: 1,
            window: Rect::full(p.width, p.height),
            origin: (0, 0),
            linearization: None,
        }
    }
    pub fn from_u16(data: &'a [u16], width: u32, height: u32, cpp: u32) -> Self {
        Self::plain(GpuInput::U16(data), width, height, cpp)
    }
    pub fn from_f32(data: &'a [f32], width: u32, height: u32, cpp: u32) -> Self {
        Self::plain(GpuInput::F32(data), width, height, cpp)
    }
    fn plain(input: GpuInput<'a>, width: u32, height: u32, cpp: u32) -> Self {
        GpuFrame {
            input,
            raster_width: width,
            raster_height: height,
            cpp,
            window: Rect::full(width, height),
            origin: (0, 0),
            linearization: None,
        }
    }
    /// Process only `r` (raster coordinates); the origin moves with it.
    pub fn with_window(mut self, r: Rect) -> Self {
        self.origin = (self.origin.0 + r.x - self.window.x, self.origin.1 + r.y - self.window.y);
        self.window = r;
        self
    }
    pub fn with_origin(mut self, origin: (u32, u32)) -> Self {
        self.origin = origin;
        self
    }
    pub fn with_linearization(mut self, lut: Option<&'a [u16]>) -> Self {
        self.linearization = lut.filter(|l| !l.is_empty());
        self
    }
    /// Bytes uploaded for this frame.
    pub fn upload_bytes(&self) -> usize {
        match self.input {
            GpuInput::Packed { data, .. } => data.len(),
            GpuInput::U16(d) => d.len() * 2,
            GpuInput::F32(d) => d.len() * 4,
        }
    }

    fn to_ffi(&self) -> ffi::frh_frame {
        let (input, data, data_len, bits, order, row_stride) = match self.input {
            GpuInput::U16(d) => (ffi::FRH_INPUT_U16, d.as_ptr() as *const c_void, d.len() as u64, 0, 0, 0),
            GpuInput::F32(d) => (ffi::FRH_INPUT_F32, d.as_ptr() as *const c_void, d.len() as u64, 0, 0, 0),
            GpuInput::Packed { data, bits, order, row_stride } => {
                let o = match order {
                    PackOrder::MsbBytes => ffi::FRH_PACK_MSB_BYTES,
                    PackOrder::LsbBytes => ffi::FRH_PACK_LSB_BYTES,
                    PackOrder::Msb16 => ffi::FRH_PACK_MSB16,
                    PackOrder::Msb32 => ffi::FRH_PACK_MSB32,
                };
                (ffi::FRH_INPUT_PACKED, data.as_ptr() as *const c_void, data.len() as u64, bits, o, row_stride as u64)
            }
        };
        let lin = self.linearization.unwrap_or(&[]);
        ffi::frh_frame {
            size: std::mem::size_of::<ffi::frh_frame>() as u32,
            input,
            data,
            data_len,
            data_on_device: 0,
            raster_width: self.raster_width,
            raster_height: self.raster_height,
            cpp: self.cpp,
            window: rect(self.window),
            origin_x: self.origin.0,
            origin_y: self.origin.1,
            linearization: if lin.is_empty() { std::ptr::null() } else { lin.as_ptr() },
            linearization_len: lin.len() as u32,
            bits,
            order,
            row_stride,
        }
    }
}

fn rect(r: Rect) -> ffi::frh_rect {
    ffi::frh_rect { x: r.x, y: r.y, width: r.width, height: r.height }
}

/// The FrameInfo fields the develop reads, with the arrays they point into.
struct InfoFfi {
    raw: ffi::frh_frame_info,
}

impl InfoFfi {
    fn new(info: &FrameInfo) -> Self {
        let mut cfa = [0u8; 64];
        let (cw, ch) = match &info.cfa {
            Some(p) => {
                for y in 0..p.height() {
                    for x in 0..p.width() {
                        cfa[y * p.width() + x] = p.color_at(x, y) as u8;
                    }
                }
                (p.width() as u32, p.height() as u32)
            }
            None => (0, 0),
        };
        let b = &info.black_level;
        let ptr = |v: &[f32]| if v.is_empty() { std::ptr::null() } else { v.as_ptr() };
        InfoFfi {
            raw: ffi::frh_frame_info {
                size: std::mem::size_of::<ffi::frh_frame_info>() as u32,
                cfa_width: cw,
                cfa_height: ch,
                cfa,
                black_rows: b.rows as u32,
                black_cols: b.cols as u32,
                black_samples: b.samples as u32,
                black_values: ptr(&b.values),
                black_values_len: b.values.len() as u32,
                black_row_deltas: ptr(&b.row_deltas),
                black_row_deltas_len: b.row_deltas.len() as u32,
                black_col_deltas: ptr(&b.col_deltas),
                black_col_deltas_len: b.col_deltas.len() as u32,
                white_values: info.white_level.values.as_ptr(),
                white_values_len: info.white_level.values.len() as u32,
            },
        }
    }
}

/// Output of the develop kernel (run.rs:96-111).
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum DevelopOutput {
    Rgba8,
    LinearRgbF32,
    Rgb16,
    LinearRgb16,
    LinearRgbF16,
}

impl DevelopOutput {
    fn code(self) -> i32 {
        match self {
            DevelopOutput::Rgba8 => ffi::FRH_OUTPUT_RGBA8,
            DevelopOutput::LinearRgbF32 => ffi::FRH_OUTPUT_LINEAR_RGB_F32,
            DevelopOutput::Rgb16 => ffi::FRH_OUTPUT_RGB16,
            DevelopOutput::LinearRgb16 => ffi::FRH_OUTPUT_LINEAR_RGB16,
            DevelopOutput::LinearRgbF16 => ffi::FRH_OUTPUT_LINEAR_RGB_F16,
        }
    }
}

/// Parameters of the fused develop pass (run.rs:113-135).
#[derive(Debug, Clone, Copy)]
pub struct DevelopSpec {
    pub crop: Rect,
    pub half_size: bool,
    pub wb: [f32; 4],
    pub cam_to_rgb: [[f32; 3]; 3],
    pub orientation: Orientation,
    pub output: DevelopOutput,
    pub levels: Option<[f32; 2]>,
    pub curve: Option<GpuCurve>,
    pub tone: Option<GpuTone>,
    pub edge: Option<EdgeSpec>,
}

/// The edge-aware (or AHD) demosaic of one frame (run.rs:137-160).
#[derive(Debug, Clone, Copy)]
pub struct EdgeSpec {
    pub period: u32,
    pub cx: u32,
    pub cy: u32,
    pub beta: f32,
    pub halo: u32,
    pub canon: &'static [u8],
    pub g0: &'static [f32],
    pub dir: &'static [f32],
    pub ahd: bool,
}

/// Encoder tables of a non-sRGB output curve (run.rs:181-188).
#[derive(Debug, Clone, Copy)]
pub struct GpuCurve {
    pub lut8: &'static [u8],
    pub lut16: &'static [f32],
}

/// The standard tone of the display outputs (run.rs:190-205).
#[derive(Debug, Clone, Copy)]
pub struct GpuTone {
    pub table: &'static [[f32; 2]],
    pub xmax: f32,
    pub scale: f32,
    pub to_prophoto: [[f32; 3]; 3],
    pub white: [f32; 3],
}

fn curve_ffi(c: &GpuCurve) -> ffi::frh_curve {
    ffi::frh_curve {
        size: std::mem::size_of::<ffi::frh_curve>() as u32,
        lut8: c.lut8.as_ptr(),
        lut8_len: c.lut8.len() as u32,
        lut16: c.lut16.as_ptr(),
        lut16_len: c.lut16.len() as u32,
    }
}

fn tone_ffi(t: &GpuTone) -> ffi::frh_tone {
    let m = t.to_prophoto;
    ffi::frh_tone {
        size: std::mem::size_of::<ffi::frh_tone>() as u32,
        table: t.table.as_ptr() as *const f32,
        table_len: t.table.len() as u32,
        xmax: t.xmax,
        scale: t.scale,
        to_prophoto: [m[0][0], m[0][1], m[0][2], m[1][0], m[1][1], m[1][2], m[2][0], m[2][1], m[2][2]],
        white: t.white,
    }
}

/// A develop spec with the structs its pointers name.
struct SpecFfi {
    raw: ffi::frh_develop_spec,
    edge: Option<Box<ffi::frh_edge_spec>>,
    curve: Option<Box<ffi::frh_curve>>,
    tone: Option<Box<ffi::frh_tone>>,
}

impl SpecFfi {
    fn new(s: &DevelopSpec, cancel: Option<&Token>) -> Self {
        let m = s.cam_to_rgb;
        let edge = s.edge.map(|e| {
            Box::new(ffi::frh_edge_spec {
                size: std::mem::size_of::<ffi::frh_edge_spec>() as u32,
                period: e.period,
                cx: e.cx,
                cy: e.cy,
                beta: e.beta,
                halo: e.halo,
                canon: e.canon.as_ptr(),
                canon_len: e.canon.len() as u32,
                g0: if e.g0.is_empty() { std::ptr::null() } else { e.g0.as_ptr() },
                g0_len: e.g0.len() as u32,
                dir: if e.dir.is_empty() { std::ptr::null() } else { e.dir.as_ptr() },
                dir_len: e.dir.len() as u32,
                ahd: e.ahd as i32,
            })
        });
        let curve = s.curve.map(|c| Box::new(curve_ffi(&c)));
        let tone = s.tone.map(|t| Box::new(tone_ffi(&t)));
        let mut out = SpecFfi {
            raw: ffi::frh_develop_spec {
                size: std::mem::size_of::<ffi::frh_develop_spec>() as u32,
                crop: rect(s.crop),
                half_size: s.half_size as i32,
                wb: s.wb,
                cam_to_rgb: [m[0][0], m[0][1], m[0][2], m[1][0], m[1][1], m[1][2], m[2][0], m[2][1], m[2][2]],
                orientation: s.orientation as i32,
                output: s.output.code(),
                output2: ffi::FRH_OUTPUT_NONE,
                has_levels: s.levels.is_some() as i32,
                levels: s.levels.unwrap_or([1.0, 0.0]),
                edge: std::ptr::null(),
                curve: std::ptr::null(),
                tone: std::ptr::null(),
                cancel: cancel.map_or(std::ptr::null(), |t| t.raw as *const _),
            },
            edge,
            curve,
            tone,
        };
        out.raw.edge = out.edge.as_deref().map_or(std::ptr::null(), |e| e as *const _);
        out.raw.curve = out.curve.as_deref().map_or(std::ptr::null(), |c| c as *const _);
        out.raw.tone = out.tone.as_deref().map_or(std::ptr::null(), |t| t as *const _);
        out
    }

    fn raw(&self) -> &ffi::frh_develop_spec {
        &self.raw
    }
}

/// What to compute (run.rs:217-227).
#[allow(clippy::large_enum_variant)]
#[derive(Debug, Clone, Copy)]
pub enum GpuOp {
    Unpack,
    Normalize,
    Develop(DevelopSpec),
}

/// Per-stage timings of one run, in milliseconds (run.rs:228-239).
#[derive(Debug, Clone, Copy, Default, PartialEq)]
pub struct GpuTimings {
    pub upload_ms: f64,
    pub compute_ms: f64,
    pub kernel_ms: Option<f64>,
    pub bands: u32,
}

impl From<ffi::frh_timings> for GpuTimings {
    fn from(t: ffi::frh_timings) -> Self {
        GpuTimings {
            upload_ms: t.upload_ms,
            compute_ms: t.compute_ms,
            kernel_ms: (t.has_kernel_ms != 0).then_some(t.kernel_ms),
            bands: t.bands,
        }
    }
}

/// Result of [`GpuContext::run`]: a device-resident image plus timings (run.rs:241-246).
#[derive(Debug)]
pub struct GpuRun {
    pub image: GpuImage,
    pub timings: GpuTimings,
}

// ---- images (crates/flashraw-gpu/src/image.rs) ----

/// Element format of a [`GpuImage`] (image.rs:8-16).
#[derive(Debug, Clone, Copy, PartialEq, Eq, Hash)]
pub enum GpuPixelFormat {
    U16,
    F32,
    Rgba8,
}

impl GpuPixelFormat {
    pub fn bytes_per_sample(self) -> usize {
        match self {
            GpuPixelFormat::U16 => 2,
            GpuPixelFormat::F32 => 4,
            GpuPixelFormat::Rgba8 => 1,
        }
    }
}

/// One band of a tiled image (image.rs:28-34). Images here are never tiled; the type exists
/// for callers that match on [`GpuImage::tiles`].
#[derive(Debug)]
pub struct GpuTile {
    pub rect: Rect,
    pub device_ptr: *const c_void,
}

/// Host copy of a downloaded [`GpuImage`] (image.rs:136-142).
#[derive(Debug, Clone, PartialEq)]
pub enum HostPixels {
    U8(Vec<u8>),
    U16(Vec<u16>),
    F32(Vec<f32>),
}

/// An image in HIP device memory (image.rs:52-61), freed when dropped.
pub struct GpuImage {
    ctx: Arc<GpuContext>,
    raw: *mut ffi::frh_image,
    width: u32,
    height: u32,
    channels: u32,
    format: GpuPixelFormat,
}

// SAFETY: the image is a device allocation with immutable metadata; the library's download
// is thread-safe.
unsafe impl Send for GpuImage {}
unsafe impl Sync for GpuImage {}

impl std::fmt::Debug for GpuImage {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "GpuImage({}x{}x{} {:?})", self.width, self.height, self.channels, self.format)
    }
}

impl Drop for GpuImage {
    fn drop(&mut self) {
        // SAFETY: raw came from a frh_run_* call and is freed once.
        unsafe { ffi::frh_image_free(self.raw) }
    }
}

impl GpuImage {
    fn wrap(ctx: Arc<GpuContext>, raw: *mut ffi::frh_image) -> Self {
        // SAFETY: raw is a live image.
        let (width, height, channels, format) = unsafe {
            (
                ffi::frh_image_width(raw),
                ffi::frh_image_height(raw),
                ffi::frh_image_channels(raw),
                ffi::frh_image_format(raw),
            )
        };
        let format = match format {
            ffi::FRH_FORMAT_RGBA8 => GpuPixelFormat::Rgba8,
            ffi::FRH_FORMAT_F32 => GpuPixelFormat::F32,
            _ => GpuPixelFormat::U16,
        };
        GpuImage { ctx, raw, width, height, channels, format }
    }

    pub fn width(&self) -> u32 {
        self.width
    }
    pub fn height(&self) -> u32 {
        self.height
    }
    /// Samples per pixel (1 for CFA/mono, 3 for RGB, 4 for RGBA8).
    pub fn channels(&self) -> u32 {
        self.channels
    }
    pub fn format(&self) -> GpuPixelFormat {
        self.format
    }
    /// Number of samples (`width * height * channels`; bytes for RGBA8).
    pub fn len(&self) -> usize {
        self.width as usize * self.height as usize * self.channels as usize
    }
    pub fn is_empty(&self) -> bool {
        self.len() == 0
    }
    /// Bytes of image data.
    pub fn byte_len(&self) -> usize {
        self.len() * self.format.bytes_per_sample()
    }
    /// The HIP device address of the pixels (row-major, from byte 0): the counterpart of
    /// `GpuImage::buffer` for HIP kernels and libraries of the caller.
    pub fn device_ptr(&self) -> *const c_void {
        // SAFETY: live image.
        unsafe { ffi::frh_image_device_ptr(self.raw) }
    }
    pub fn is_tiled(&self) -> bool {
        false
    }
    pub fn tiles(&self) -> &[GpuTile] {
        &[]
    }
    pub fn context(&self) -> &Arc<GpuContext> {
        &self.ctx
    }

    /// Copy the pixels to host memory; the second value is the time taken in milliseconds
    /// (image.rs:207-214).
    pub async fn download_async(&self) -> Result<(HostPixels, f64)> {
        self.download()
    }

    /// Blocking download (image.rs:297-302).
    pub fn download(&self) -> Result<(HostPixels, f64)> {
        let t0 = std::time::Instant::now();
        let n = self.len();
        let bytes = self.byte_len() as u64;
        let px = match self.format {
            GpuPixelFormat::Rgba8 => {
                let mut v = vec![0u8; n];
                // SAFETY: v holds `bytes` bytes.
                check(unsafe { ffi::frh_image_download(self.raw, v.as_mut_ptr().cast(), bytes) })?;
                HostPixels::U8(v)
            }
            GpuPixelFormat::U16 => {
                let mut v = vec![0u16; n];
                // SAFETY: as above.
                check(unsafe { ffi::frh_image_download(self.raw, v.as_mut_ptr().cast(), bytes) })?;
                HostPixels::U16(v)
            }
            GpuPixelFormat::F32 => {
                let mut v = vec![0f32; n];
                // SAFETY: as above.
                check(unsafe { ffi::frh_image_download(self.raw, v.as_mut_ptr().cast(), bytes) })?;
                HostPixels::F32(v)
            }
        };
        Ok((px, t0.elapsed().as_secs_f64() * 1e3))
    }

    /// [`GpuImage::download`] with 4-channel 16-bit images reduced to RGB (image.rs:214-225);
    /// 16-bit outputs here are stored with 3 samples per pixel already.
    pub fn download_rgb(&self) -> Result<(HostPixels, f64)> {
        let (px, ms) = self.download()?;
        Ok(match (px, self.channels) {
            (HostPixels::U16(v), 4) => (HostPixels::U16(v.chunks_exact(4).flat_map(|p| [p[0], p[1], p[2]]).collect()), ms),
            (px, _) => (px, ms),
        })
    }

    pub async fn download_rgb_async(&self) -> Result<(HostPixels, f64)> {
        self.download_rgb()
    }
}

// ---- luma / chroma noise reduction (crates/flashraw-gpu/src/denoise.rs) ----

/// The noise model of one pyramid level of a plane (denoise.rs:30-40).
#[derive(Debug, Clone, PartialEq, Default)]
pub struct NoiseModel {
    pub means: Vec<f32>,
    pub sigmas: Vec<[f32; 64]>,
    pub pixel_var: Vec<f32>,
}

/// The CPU side of the noise estimate in `flashraw-gpu` (denoise.rs:42-43). Not called here:
/// the library fits the models on the device from the blocks it measures.
pub type ModelsFn = dyn Fn(&[Vec<Option<&[[f32; 64]]>>]) -> Vec<Vec<Option<NoiseModel>>>;

/// How one plane is denoised (denoise.rs:44-57).
#[derive(Debug, Clone, Copy, PartialEq)]
pub struct PlaneSpec {
    pub levels: u32,
    pub gain: f32,
    pub steps: [u32; 2],
    pub hard_only0: bool,
    pub keep: f32,
}

/// Everything [`GpuContext::denoise`] needs (denoise.rs:59-80).
pub struct DenoiseSpec<'a> {
    pub luma: Option<PlaneSpec>,
    pub chroma: Option<PlaneSpec>,
    pub chroma_radius: u32,
    pub hard: f32,
    pub max_blocks: u32,
    /// Kept for the `flashraw-gpu` signature; the models come from the library's own fit.
    pub models: &'a ModelsFn,
    pub output: DevelopOutput,
    pub orientation: Orientation,
    pub tone: Option<GpuTone>,
    pub curve: Option<GpuCurve>,
}

// ---- tables ----

/// The 4097-entry sRGB 8-bit table the kernels encode with (kernels.rs:348-371).
pub fn srgb8_table() -> Vec<u8> {
    let mut t = vec![0u8; 4097];
    // SAFETY: t holds 4097 bytes.
    unsafe { ffi::frh_srgb8_table(t.as_mut_ptr(), t.len()) };
    t
}

#[cfg(test)]
mod layout {
    use super::ffi::*;
    use std::mem::size_of;

    // sizeof of the header's structs as gcc lays them out on x86-64 (tests/abi_layout.c)
    #[test]
    fn struct_sizes_match_the_header() {
        assert_eq!(size_of::<frh_rect>(), 16);
        assert_eq!(size_of::<frh_frame>(), 96);
        assert_eq!(size_of::<frh_frame_info>(), 152);
        assert_eq!(size_of::<frh_edge_spec>(), 72);
        assert_eq!(size_of::<frh_curve>(), 40);
        assert_eq!(size_of::<frh_tone>(), 80);
        assert_eq!(size_of::<frh_develop_spec>(), 136);
        assert_eq!(size_of::<frh_timings>(), 32);
        assert_eq!(size_of::<frh_dn_spec>(), 56);
        assert_eq!(size_of::<frh_nr_plane>(), 24);
        assert_eq!(size_of::<frh_nr_spec>(), 104);
        assert_eq!(size_of::<frh_server_config>(), 24);
        assert_eq!(size_of::<frh_job_result>(), 48);
        assert_eq!(size_of::<frh_server_stats>(), 64);
    }
}
