// C ABI of the input operations: GpuOp::Unpack and GpuOp::Normalize (crates/flashraw-gpu/src/run.rs:
// 503-517, GpuContext::run 426) - the window of a packed payload or u16 raster as u16 samples
// through the linearisation curve, or normalised f32 samples max((v - black) * scale, 0).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "flashraw_hip.h"
#include "frh_internal.h"
#include "frh_launch.h"

using namespace frh;

namespace {

double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// The frame's window as input-kernel arguments: host input is uploaded (the window's rows) into a
// pooled buffer owned by the caller (*d_in, *in_bytes).
int32_t input_args(frh_context *ctx, const frh_frame *f, const InputDesc &in, hipStream_t st, InArgs &a,
                   void **d_in, uint64_t *in_bytes) {
    memset(&a, 0, sizeof(a));
    const uint8_t *base = static_cast<const uint8_t *>(f->data);
    uint32_t first_row = f->window.y;
    hipError_t e;
    if (!f->data_on_device) {
        *in_bytes = in.upload_bytes ? in.upload_bytes : 4;
        *d_in = ctx->alloc(*in_bytes, &e);
        if (!*d_in) return hip_fail(e, "input allocation");
        if (in.upload_bytes &&
            (e = hipMemcpyAsync(*d_in, base + in.first_byte, in.upload_bytes, hipMemcpyHostToDevice, st)) != hipSuccess)
            return hip_fail(e, "upload");
        base = static_cast<const uint8_t *>(*d_in);
        first_row = 0;
    }
    a.kind = (uint32_t)in.kind;
    a.src = base;
    a.len = f->data_on_device ? in.len : in.upload_bytes;
    a.stride = in.stride;
    a.row_bytes = (uint32_t)in.row_bytes;
    a.bits = in.bits;
    a.x0 = f->window.x * in.cpp;
    a.y0 = first_row;
    a.n = in.w * in.cpp;
    a.rows = in.h;
    a.cpp = in.cpp;
    a.ox = f->origin_x;
    a.oy = f->origin_y;
    if (in.lut) {
        int32_t rc;
        a.lut = static_cast<const uint16_t *>(device_table(ctx, f->linearization, f->linearization_len * 2ull, &rc));
        if (rc) return rc;
        a.lut_len = f->linearization_len;
    }
    return FRH_OK;
}

// Calibration of the normalisation (Calib::new for every component, deltas) into `a`.
int32_t calib_args(frh_context *ctx, const frh_frame_info *info, InArgs &a) {
    std::vector<float> gcal;
    int32_t rc = build_calib(info, a.cpp, gcal, a.cal_pw, a.cal_ph);
    if (rc) return rc;
    a.gcal = static_cast<const float *>(device_table(ctx, gcal.data(), gcal.size() * 4ull, &rc));
    if (rc) return rc;
    if (info->black_row_deltas_len) {
        if (!info->black_row_deltas) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL row deltas");
        a.n_rowd = info->black_row_deltas_len;
        a.rowd = static_cast<const float *>(device_table(ctx, info->black_row_deltas, a.n_rowd * 4ull, &rc));
        if (rc) return rc;
    }
    if (info->black_col_deltas_len) {
        if (!info->black_col_deltas) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL column deltas");
        a.n_cold = info->black_col_deltas_len;
        a.cold = static_cast<const float *>(device_table(ctx, info->black_col_deltas, a.n_cold * 4ull, &rc));
        if (rc) return rc;
    }
    return FRH_OK;
}

// Launch the op on `st` into `dst`: unpack (normalize false) or normalise. Staged normalisation
// unpacks into pooled scratch first (*d_stage, released by the caller after the stream).
int32_t launch_op(frh_context *ctx, const frh_frame_info *info, bool normalize, const InputDesc &in, InArgs a,
                  void *dst, hipStream_t st, void **d_stage, uint64_t *stage_bytes) {
    hipError_t e;
    if (!normalize) {
        if ((e = launch_unpack(a, in.order, static_cast<uint16_t *>(dst), st)) != hipSuccess)
            return hip_fail(e, "unpack launch");
        return FRH_OK;
    }
    int32_t rc = calib_args(ctx, info, a);
    if (rc) return rc;
    if (in.staged) {
        *stage_bytes = ((uint64_t)a.n * a.rows * 2 + 3) & ~3ull;
        *d_stage = ctx->alloc(*stage_bytes, &e);
        if (!*d_stage) return hip_fail(e, "unpack allocation");
        if ((e = launch_unpack(a, in.order, static_cast<uint16_t *>(*d_stage), st)) != hipSuccess)
            return hip_fail(e, "unpack launch");
        a.kind = FRH_INPUT_U16;
        a.src = *d_stage;
        a.stride = a.n;
        a.x0 = 0;
        a.y0 = 0;
    }
    if ((e = launch_normalize(a, static_cast<float *>(dst), st)) != hipSuccess) return hip_fail(e, "normalize launch");
    return FRH_OK;
}

int32_t run_op(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info, bool normalize, frh_image **out,
               frh_timings *timings) {
    if (!ctx || !frame || !out || (normalize && !info)) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (frame->size < sizeof(frh_frame) || (info && info->size < sizeof(frh_frame_info)))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "descriptor size field too small");
    InputDesc in;
    int32_t rc = check_input(frame, in);
    if (rc) return rc;
    if (!normalize && in.kind == FRH_INPUT_F32) return fail(FRH_ERROR_INVALID_ARGUMENT, "unpack needs integer input");
    (void)hipSetDevice(ctx->device);
    const auto t0 = std::chrono::steady_clock::now();
    InArgs a;
    void *d_in = nullptr, *d_stage = nullptr;
    uint64_t in_bytes = 0, stage_bytes = 0;
    auto cleanup = [&]() {
        if (d_in) ctx->release(d_in, in_bytes);
        if (d_stage) ctx->release(d_stage, stage_bytes);
    };
    hipError_t e;
    rc = input_args(ctx, frame, in, ctx->stream, a, &d_in, &in_bytes);
    if (rc == FRH_OK && (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) rc = hip_fail(e, "upload");
    if (rc) {
        cleanup();
        return rc;
    }
    const auto t1 = std::chrono::steady_clock::now();
    const uint64_t samples = (uint64_t)a.n * a.rows;
    const uint64_t bytes = normalize ? samples * 4 : (samples * 2 + 3) & ~3ull;
    void *d = ctx->alloc(bytes, &e);
    if (!d) {
        cleanup();
        return hip_fail(e, "output allocation");
    }
    frh_image *img = new frh_image{ctx, d, normalize ? samples * 4 : samples * 2, in.w, in.h, in.cpp,
                                   normalize ? FRH_FORMAT_F32 : FRH_FORMAT_U16};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    const bool timed = hipEventCreate(&ev0) == hipSuccess && hipEventCreate(&ev1) == hipSuccess;
    if (timed) (void)hipEventRecord(ev0, ctx->stream);
    rc = launch_op(ctx, info, normalize, in, a, d, ctx->stream, &d_stage, &stage_bytes);
    if (timed) (void)hipEventRecord(ev1, ctx->stream);
    e = hipStreamSynchronize(ctx->stream);
    float kms = 0.0f;
    if (timed && hipEventElapsedTime(&kms, ev0, ev1) != hipSuccess) kms = 0.0f;
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    cleanup();
    if (rc == FRH_OK && e != hipSuccess) rc = hip_fail(e, normalize ? "normalize" : "unpack");
    if (rc) {
        frh_image_free(img);
        return rc;
    }
    if (timings) {
        timings->upload_ms = ms_between(t0, t1);
        timings->compute_ms = ms_between(t1, std::chrono::steady_clock::now());
        timings->kernel_ms = kms;
        timings->has_kernel_ms = timed ? 1 : 0;
        timings->bands = 1;
    }
    *out = img;
    return FRH_OK;
}

int32_t into_op(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info, bool normalize, void *dst,
                uint64_t dst_bytes, void *stream) {
    if (!ctx || !frame || !dst || (normalize && !info)) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (frame->size < sizeof(frh_frame) || (info && info->size < sizeof(frh_frame_info)))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "descriptor size field too small");
    InputDesc in;
    int32_t rc = check_input(frame, in);
    if (rc) return rc;
    if (!normalize && in.kind == FRH_INPUT_F32) return fail(FRH_ERROR_INVALID_ARGUMENT, "unpack needs integer input");
    if (!frame->data_on_device) return fail(FRH_ERROR_INVALID_ARGUMENT, "this entry point needs device input");
    const uint64_t samples = (uint64_t)in.w * in.cpp * in.h;
    if (dst_bytes < samples * (normalize ? 4 : 2)) return fail(FRH_ERROR_INVALID_ARGUMENT, "output buffer too small");
    if (!normalize && (reinterpret_cast<uintptr_t>(dst) & 3u))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "output must be 4-byte aligned");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = (hipStream_t)stream;
    InArgs a;
    void *d_in = nullptr, *d_stage = nullptr;
    uint64_t in_bytes = 0, stage_bytes = 0;
    rc = input_args(ctx, frame, in, st, a, &d_in, &in_bytes);
    if (rc == FRH_OK) rc = launch_op(ctx, info, normalize, in, a, dst, st, &d_stage, &stage_bytes);
    if (d_stage) {
        hipError_t e = hipStreamSynchronize(st);
        ctx->release(d_stage, stage_bytes);
        if (rc == FRH_OK && e != hipSuccess) rc = hip_fail(e, "normalize");
    }
    return rc;
}

}  // namespace

extern "C" int32_t frh_run_unpack(frh_context *ctx, const frh_frame *frame, frh_image **out, frh_timings *timings) {
    try {
        return run_op(ctx, frame, nullptr, false, out, timings);
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_run_normalize(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info,
                                     frh_image **out, frh_timings *timings) {
    try {
        return run_op(ctx, frame, info, true, out, timings);
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_unpack_into(frh_context *ctx, const frh_frame *frame, void *dst, uint64_t dst_bytes,
                                   void *stream) {
    try {
        return into_op(ctx, frame, nullptr, false, dst, dst_bytes, stream);
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_normalize_into(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info, void *dst,
                                      uint64_t dst_bytes, void *stream) {
    try {
        return into_op(ctx, frame, info, true, dst, dst_bytes, stream);
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}
