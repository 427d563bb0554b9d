// C ABI of the develop path (include/flashraw_hip.h): context, validation of the reference's
// develop descriptors, upload, band planning, launch and device-resident images.
// Reference: crates/flashraw-gpu/src/context.rs (context, pool, limits), run.rs (run_banded:
// validation 439-646, bands 650-662, timings 227-239) and image.rs (GpuImage).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "flashraw_hip.h"
#include "frh_device.h"
#include "frh_launch.h"
#include "frh_internal.h"
#include "bil_walk4.h"

using namespace frh;

// ------------------------------------------------------------------------------------------
// Errors: a status per call plus a thread-local message (crates/flashraw-capi/src/lib.rs:1-16)

static thread_local std::string g_last_error;

int32_t frh::fail(int32_t code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int32_t frh::hip_fail(hipError_t e, const char *what) {
    int32_t code = e == hipErrorOutOfMemory ? FRH_ERROR_OUT_OF_MEMORY : FRH_ERROR_GPU_UNAVAILABLE;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

extern "C" const char *frh_last_error_message(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------------------------------
// Cancellation (CancelToken, crates/flashraw-core/src/cancel.rs:26-76; the C API's
// flashraw_cancel_token, crates/flashraw-capi/include/flashraw.h:1247-1275)

struct frh_cancel_token {
    mutable std::atomic<bool> flag{false};
    int32_t (*poll)(void *) = nullptr;
    void *user = nullptr;
};

extern "C" frh_cancel_token *frh_cancel_token_new(void) { return new (std::nothrow) frh_cancel_token(); }

extern "C" frh_cancel_token *frh_cancel_token_with_poll(int32_t (*poll)(void *user), void *user) {
    frh_cancel_token *t = new (std::nothrow) frh_cancel_token();
    if (t) {
        t->poll = poll;
        t->user = user;
    }
    return t;
}

extern "C" void frh_cancel_token_cancel(const frh_cancel_token *t) {
    if (t) t->flag.store(true, std::memory_order_release);
}

bool frh::cancelled(const frh_cancel_token *t) {
    if (!t) return false;
    if (t->flag.load(std::memory_order_acquire)) return true;
    if (t->poll && t->poll(t->user)) {
        t->flag.store(true, std::memory_order_release);
        return true;
    }
    return false;
}

extern "C" int32_t frh_cancel_token_is_cancelled(const frh_cancel_token *t) { return frh::cancelled(t) ? 1 : 0; }

extern "C" void frh_cancel_token_free(frh_cancel_token *t) { delete t; }

int32_t frh::cancel_fail() { return fail(FRH_ERROR_CANCELLED, "cancelled"); }

const frh_cancel_token *frh::spec_cancel(const frh_develop_spec *s) {
    return s && s->size >= sizeof(frh_develop_spec) ? s->cancel : nullptr;
}

// FLASHRAW_GPU=0|off|false|none|cpu (any case) disables the device backend (disabled_by_env,
// crates/flashraw-gpu/src/context.rs:734-747).
static bool disabled_by_env() {
    const char *v = getenv("FLASHRAW_GPU");
    if (!v) return false;
    std::string s(v);
    for (char &c : s) c = (char)tolower((unsigned char)c);
    return s == "0" || s == "off" || s == "false" || s == "none" || s == "cpu";
}

// ------------------------------------------------------------------------------------------
// Tables (crates/flashraw-gpu/src/kernels.rs:348-371; process.rs:1903-1934)

static uint8_t sat_u8(float x) {
    if (!(x >= 1.0f)) return 0;
    if (x >= 255.0f) return 255;
    return (uint8_t)x;
}

// The sRGB 8-bit table: x = i / 4096 in f32, the curve with f32 powf, (s * 255 + 0.5) clamped
// and truncated.
void frh::build_srgb8(uint8_t out[4097]) {
    const float e = 1.0f / 2.4f;
    for (int i = 0; i <= 4096; i++) {
        float x = (float)i / 4096.0f;
        volatile float p = powf(x, e);  // the libm call, as the reference's f32::powf
        float s = (x <= 0.0031308f) ? 12.92f * x : 1.055f * p - 0.055f;
        float v = s * 255.0f + 0.5f;
        v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
        out[i] = sat_u8(v);
    }
}

// The sRGB 16-bit curve at 16385 points (f64, stored as f32) as (value, step) pairs.
static void build_srgb16_pairs(float *pairs) {
    std::vector<float> lut(16385);
    for (int i = 0; i <= 16384; i++) {
        double x = (double)i / 16384.0;
        double s = (x <= 0.0031308) ? 12.92 * x : 1.055 * pow(x, 1.0 / 2.4) - 0.055;
        lut[i] = (float)(s * 65535.0);
    }
    for (int i = 0; i < 16384; i++) {
        pairs[2 * i] = lut[i];
        pairs[2 * i + 1] = lut[i + 1] - lut[i];
    }
}

// The AHD blend weights w = a / (a + b) for homogeneity sums a, b in 0..72 (0.5 when both are 0),
// f32 division as ahd.rs:389 performs it.
static void build_blend_table(float *out) {
    for (int a = 0; a <= 72; a++)
        for (int b = 0; b <= 72; b++) {
            const float fa = (float)a, fb = (float)b;
            out[a * 73 + b] = fa + fb > 0.0f ? fa / (fa + fb) : 0.5f;
        }
}

extern "C" int32_t frh_srgb8_table(uint8_t *out, size_t len) {
    if (!out || len < 4097) return fail(FRH_ERROR_INVALID_ARGUMENT, "sRGB table needs 4097 bytes");
    build_srgb8(out);
    return FRH_OK;
}

extern "C" int32_t frh_check_sqrt(frh_context *ctx, uint64_t *mismatches) {
    if (!ctx || !mismatches) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipSuccess;
    void *d = ctx->alloc(8, &e);
    if (!d) return hip_fail(e, "check buffer");
    unsigned long long n = 0;
    if ((e = hipMemsetAsync(d, 0, 8, ctx->stream)) == hipSuccess &&
        (e = check_sqrt(static_cast<unsigned long long *>(d), ctx->stream)) == hipSuccess &&
        (e = hipMemcpyAsync(&n, d, 8, hipMemcpyDeviceToHost, ctx->stream)) == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    ctx->release(d, 8);
    if (e != hipSuccess) return hip_fail(e, "square root check");
    *mismatches = n;
    return FRH_OK;
}

// ------------------------------------------------------------------------------------------
// Context (context.rs:116-150, 775-842)

extern "C" int32_t frh_available(void) {
    if (disabled_by_env()) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

extern "C" int32_t frh_context_create(int32_t device, frh_context **out) {
    if (!out) return fail(FRH_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (disabled_by_env()) return fail(FRH_ERROR_GPU_UNAVAILABLE, "disabled by `FLASHRAW_GPU`");
    try {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) return fail(FRH_ERROR_GPU_UNAVAILABLE, "no HIP device");
        if (device < 0 || device >= n) return fail(FRH_ERROR_INVALID_ARGUMENT, "device index out of range");
        std::unique_ptr<frh_context> c(new frh_context());
        c->device = device;
        if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
        // this device's copy of the no-allocation batch's scratch array
        if ((e = noalloc_scratch_address(&c->noalloc_scratch)) != hipSuccess)
            return hip_fail(e, "scratch array address");
        hipDeviceProp_t prop;
        if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
        c->num_cus = prop.multiProcessorCount;
        c->name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
        c->total_mem = prop.totalGlobalMem;
        if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
            return hip_fail(e, "hipStreamCreate");
        uint8_t lut[4100] = {0};
        build_srgb8(lut);
        std::vector<float> pairs(2 * 16384);
        build_srgb16_pairs(pairs.data());
        if ((e = hipMalloc(&c->d_lut8, 4100)) != hipSuccess) return hip_fail(e, "hipMalloc");
        if ((e = hipMalloc(&c->d_srgb16, pairs.size() * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc");
        if ((e = hipMemcpy(c->d_lut8, lut, 4100, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "upload");
        if ((e = hipMemcpy(c->d_srgb16, pairs.data(), pairs.size() * sizeof(float), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return hip_fail(e, "upload");
        std::vector<float> wtab(73 * 73);
        build_blend_table(wtab.data());
        if ((e = hipMalloc(&c->d_wtab, wtab.size() * sizeof(float))) != hipSuccess) return hip_fail(e, "hipMalloc");
        if ((e = hipMemcpy(c->d_wtab, wtab.data(), wtab.size() * sizeof(float), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return hip_fail(e, "upload");
        *out = c.release();
        return FRH_OK;
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

frh_context::~frh_context() {
    (void)hipSetDevice(device);
    for (auto &b : pool) (void)hipFree(b.ptr);
    if (d_lut8) (void)hipFree(d_lut8);
    if (d_srgb16) (void)hipFree(d_srgb16);
    if (d_wtab) (void)hipFree(d_wtab);
    for (auto &kv : tables) (void)hipFree(kv.second.d);
    if (h_stage) (void)hipHostFree(h_stage);
    if (stream) (void)hipStreamDestroy(stream);
}

extern "C" void frh_context_destroy(frh_context *ctx) { delete ctx; }

extern "C" int32_t frh_context_info(const frh_context *ctx, char *buf, size_t len) {
    if (!ctx || !buf || len == 0) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    snprintf(buf, len, "%s, %d CUs, %.1f GB", ctx->name.c_str(), ctx->num_cus, ctx->total_mem / 1e9);
    return FRH_OK;
}

extern "C" int32_t frh_set_limits(frh_context *ctx, uint64_t max_binding_bytes) {
    if (!ctx) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL context");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->max_binding = max_binding_bytes;
    return FRH_OK;
}

// A caching device allocator bounded like the reference's pool (POOL_MAX_BYTES, context.rs:153).
void *frh_context::alloc(uint64_t bytes, hipError_t *err) {
    {
        std::lock_guard<std::mutex> g(mu);
        size_t best = pool.size();
        for (size_t i = 0; i < pool.size(); i++)
            if (pool[i].size >= bytes && pool[i].size <= bytes * 2 &&
                (best == pool.size() || pool[i].size < pool[best].size))
                best = i;
        if (best != pool.size()) {
            void *p = pool[best].ptr;
            pooled -= pool[best].size;
            pool.erase(pool.begin() + (long)best);
            *err = hipSuccess;
            return p;
        }
    }
    void *p = nullptr;
    (void)hipSetDevice(device);
    *err = hipMalloc(&p, bytes ? bytes : 1);
    return *err == hipSuccess ? p : nullptr;
}

void frh_context::release(void *p, uint64_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu);
    if (pooled + bytes <= POOL_MAX_BYTES) {
        pool.push_back({p, bytes});
        pooled += bytes;
        return;
    }
    (void)hipSetDevice(device);
    (void)hipFree(p);
}

// ------------------------------------------------------------------------------------------
// Images (image.rs:52-306)

extern "C" uint32_t frh_image_width(const frh_image *i) { return i ? i->w : 0; }
extern "C" uint32_t frh_image_height(const frh_image *i) { return i ? i->h : 0; }
extern "C" uint32_t frh_image_channels(const frh_image *i) { return i ? i->ch : 0; }
extern "C" int32_t frh_image_format(const frh_image *i) { return i ? i->fmt : -1; }
extern "C" const void *frh_image_device_ptr(const frh_image *i) { return i ? i->dptr : nullptr; }
extern "C" uint64_t frh_image_bytes(const frh_image *i) { return i ? i->bytes : 0; }

extern "C" int32_t frh_image_download(const frh_image *img, void *host, uint64_t bytes) {
    if (!img || !host) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (bytes < img->bytes) return fail(FRH_ERROR_INVALID_ARGUMENT, "host buffer too small");
    (void)hipSetDevice(img->ctx->device);
    hipError_t e = hipMemcpy(host, img->dptr, img->bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "download");
    return FRH_OK;
}

extern "C" void frh_image_free(frh_image *img) {
    if (!img) return;
    img->ctx->release(img->dptr, img->bytes);
    delete img;
}

// ------------------------------------------------------------------------------------------
// Develop (run.rs:432-1072)

static uint64_t gcd64(uint64_t a, uint64_t b) { return b == 0 ? a : gcd64(b, a % b); }

// Bytes per sample and samples per pixel of an output encoding.
void frh::out_format(int32_t mode, uint32_t *ch, uint32_t *bps, int32_t *fmt) {
    switch (mode) {
    case OUT_RGBA8: *ch = 4; *bps = 1; *fmt = FRH_FORMAT_RGBA8; break;
    case OUT_LINEAR_F32: *ch = 3; *bps = 4; *fmt = FRH_FORMAT_F32; break;
    default: *ch = 3; *bps = 2; *fmt = FRH_FORMAT_U16; break;
    }
}

bool frh::valid_mode(int32_t m) { return m >= OUT_RGBA8 && m <= OUT_LINEAR_F16; }
bool frh::display_mode(int32_t m) { return m == OUT_RGBA8 || m == OUT_SRGB16; }

static std::tuple<int, uint32_t, uint32_t, int> tune_key(bool ahd, const DevArgs &a) {
    return std::make_tuple(ahd ? 1 : 0, a.dev_w, a.dev_h, (a.out_mode == OUT_RGBA8 && a.out_mode2 == OUT_NONE) ? 1 : 0);
}

// The autotuned tile shape (strip height, waves per block, columns per lane) for this geometry, or
// (fallback, 0, 0): the defaults.
static frh_context::TileShape tuned_shape(const frh_context *ctx, bool ahd, const DevArgs &a, uint32_t fallback) {
    std::lock_guard<std::mutex> g(const_cast<frh_context *>(ctx)->mu);
    auto it = ctx->tuned.find(tune_key(ahd, a));
    return it == ctx->tuned.end() ? frh_context::TileShape{fallback, 0u, 0u} : it->second;
}

// Everything a develop launch needs, validated (run.rs:439-646).
// Calib::new (process.rs:256-286; kernels.rs:216-248) for every component into
// gcal = black[cpp][pw * ph] then scale[cpp][pw * ph]; the pattern is lcm(CFA, black repeat),
// at most 64 positions here.
int32_t frh::build_calib(const frh_frame_info *info, uint32_t cpp, std::vector<float> &gcal, uint32_t &pw_out,
                         uint32_t &ph_out) {
    const bool has_cfa = info->cfa_width != 0 && info->cfa_height != 0;
    const uint64_t rows = info->black_rows ? info->black_rows : 1, cols = info->black_cols ? info->black_cols : 1;
    const uint64_t samples = info->black_samples ? info->black_samples : 1;
    const uint64_t cw0 = has_cfa ? info->cfa_width : 1, ch0 = has_cfa ? info->cfa_height : 1;
    uint64_t pw = cw0 / gcd64(cw0, cols) * cols, ph = ch0 / gcd64(ch0, rows) * rows;
    pw = pw < 64 ? pw : 64;
    ph = ph < 64 ? ph : 64;
    if (pw * ph > 64) return fail(FRH_ERROR_NOT_AVAILABLE, "black-level pattern larger than 64 positions");
    if ((info->black_values_len && !info->black_values) || (info->white_values_len && !info->white_values))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL calibration array");
    gcal.assign(2 * (size_t)cpp * pw * ph, 0.0f);
    for (uint32_t c = 0; c < cpp; c++) {
        for (uint64_t y = 0; y < ph; y++) {
            for (uint64_t x = 0; x < pw; x++) {
                uint64_t idx = ((y % rows) * cols + x % cols) * samples + c % samples;
                float bl = idx < info->black_values_len ? info->black_values[idx] : 0.0f;
                uint32_t colour = has_cfa ? info->cfa[(y % ch0) * cw0 + x % cw0] : c;
                uint32_t nw = info->white_values_len;
                uint32_t wi = nw > 1 ? (colour < nw - 1 ? colour : nw - 1) : 0;
                uint32_t wv = wi < nw ? info->white_values[wi] : (nw ? info->white_values[0] : 0);
                float white = (float)wv;
                gcal[(c * ph + y) * pw + x] = bl;
                gcal[((cpp + c) * ph + y) * pw + x] = white > bl ? 1.0f / (white - bl) : 1.0f;
            }
        }
    }
    pw_out = (uint32_t)pw;
    ph_out = (uint32_t)ph;
    return FRH_OK;
}

// GpuInput validation (run.rs:456-490) and the window's byte range in the raster.
int32_t frh::check_input(const frh_frame *f, InputDesc &in) {
    memset(&in, 0, sizeof(in));
    const uint32_t w = f->window.width, h = f->window.height, cpp = f->cpp ? f->cpp : 1;
    if (w == 0 || h == 0 || (uint64_t)f->window.x + w > f->raster_width ||
        (uint64_t)f->window.y + h > f->raster_height)
        return fail(FRH_ERROR_INVALID_ARGUMENT, "window outside the raster");
    if (!f->data) return fail(FRH_ERROR_INVALID_ARGUMENT, "frame data is NULL");
    in.kind = f->input;
    in.w = w;
    in.h = h;
    in.cpp = cpp;
    in.len = f->data_len;
    const uint64_t raster_row = (uint64_t)f->raster_width * cpp;
    const uint64_t raster_samples = raster_row * f->raster_height;
    if (f->input == FRH_INPUT_PACKED) {
        if (f->bits < 1 || f->bits > 16)
            return fail(FRH_ERROR_INVALID_ARGUMENT, "unsupported packed bit depth " + std::to_string(f->bits));
        if (f->order < FRH_PACK_MSB_BYTES || f->order > FRH_PACK_MSB32)
            return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown pack order");
        in.bits = f->bits;
        in.order = f->order;
        in.row_bytes = (raster_row * f->bits + 7) / 8;
        in.stride = f->row_stride;
        if (in.row_bytes > 0xffffffffull || in.stride > 0xffffffffull)
            return fail(FRH_ERROR_LIMIT_EXCEEDED, "packed row too long");
        // the payload bytes of the window's rows (run.rs:785-796); bytes past the payload read 0
        const uint64_t row_extent = std::max(in.stride, in.row_bytes);
        const uint64_t start = std::min((uint64_t)f->window.y * in.stride, in.len);
        const uint64_t end = std::min(((uint64_t)f->window.y + h - 1) * in.stride + row_extent, in.len);
        in.first_byte = start;
        in.upload_bytes = end > start ? end - start : 0;
    } else if (f->input == FRH_INPUT_U16 || f->input == FRH_INPUT_F32) {
        if (f->data_len < raster_samples)
            return fail(FRH_ERROR_INVALID_ARGUMENT, f->input == FRH_INPUT_U16 ? "u16 raster shorter than its dimensions"
                                                                            : "f32 raster shorter than its dimensions");
        const uint64_t bps = f->input == FRH_INPUT_U16 ? 2 : 4;
        in.stride = raster_row;
        in.first_byte = (uint64_t)f->window.y * raster_row * bps;
        in.upload_bytes = (uint64_t)h * raster_row * bps;
    } else {
        return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown input kind");
    }
    if (f->linearization_len && !f->linearization) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL linearisation table");
    // run.rs:491: the curve applies to integer samples only
    in.lut = f->linearization && f->linearization_len && f->input != FRH_INPUT_F32;
    in.staged = f->input == FRH_INPUT_PACKED || in.lut;
    return FRH_OK;
}

static uint64_t fnv1a(uint64_t h, const void *p, uint64_t n) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (uint64_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// Free least recently used, unpinned entries until `incoming` more bytes fit under the cache's
// bounds. Entries used within the last 16 lookups are kept: a develop looks up at most 8 tables
// between its first lookup and its launch, so a concurrent caller's tables are not freed under
// it. Called with ctx->mu held.
static void evict_tables(frh_context *ctx, uint64_t incoming) {
    std::vector<void *> victims;
    while (ctx->table_bytes + incoming > TABLE_CACHE_BYTES || ctx->tables.size() + 1 > TABLE_CACHE_ENTRIES) {
        auto lru = ctx->tables.end();
        for (auto it = ctx->tables.begin(); it != ctx->tables.end(); ++it)
            if (it->second.pins == 0 && it->second.last_use + 16 < ctx->table_tick &&
                (lru == ctx->tables.end() || it->second.last_use < lru->second.last_use))
                lru = it;
        if (lru == ctx->tables.end()) break;  // everything left is pinned or in use: grow past the bound
        victims.push_back(lru->second.d);
        ctx->table_bytes -= lru->second.dev_bytes;
        ctx->tables.erase(lru);
    }
    if (victims.empty()) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();  // launches queued earlier may still read the evicted tables
    for (void *d : victims) (void)hipFree(d);
}

const void *frh::cached_table(frh_context *ctx, int kind, const void *key, uint64_t key_bytes,
                              const std::function<bool(std::vector<uint8_t> &)> &make, int32_t *rc) {
    *rc = FRH_OK;
    const uint64_t hsh = fnv1a(1469598103934665603ull, key, key_bytes);
    std::lock_guard<std::mutex> g(ctx->mu);
    const uint64_t tick = ++ctx->table_tick;
    auto range = ctx->tables.equal_range(std::make_tuple(kind, hsh, key_bytes));
    for (auto it = range.first; it != range.second; ++it)
        if (memcmp(it->second.key.data(), key, key_bytes) == 0) {
            it->second.last_use = tick;
            return it->second.d;
        }
    std::vector<uint8_t> dev;
    if (!make(dev)) {
        *rc = fail(FRH_ERROR_INTERNAL, "table construction failed");
        return nullptr;
    }
    evict_tables(ctx, dev.size());
    void *d = nullptr;
    hipError_t e;
    (void)hipSetDevice(ctx->device);
    if ((e = hipMalloc(&d, dev.size())) != hipSuccess) {
        *rc = hip_fail(e, "table allocation");
        return nullptr;
    }
    if ((e = hipMemcpy(d, dev.data(), dev.size(), hipMemcpyHostToDevice)) != hipSuccess) {
        (void)hipFree(d);
        *rc = hip_fail(e, "table upload");
        return nullptr;
    }
    TableEntry t{d, dev.size(), std::vector<uint8_t>(static_cast<const uint8_t *>(key),
                                                      static_cast<const uint8_t *>(key) + key_bytes),
                 tick, 0};
    ctx->tables.emplace(std::make_tuple(kind, hsh, key_bytes), std::move(t));
    ctx->table_bytes += dev.size();
    return d;
}

const void *frh::device_table(frh_context *ctx, const void *host, uint64_t bytes, int32_t *rc) {
    *rc = FRH_OK;
    if (!bytes) return nullptr;
    return cached_table(ctx, 0, host, bytes, [&](std::vector<uint8_t> &dev) {
        dev.assign(static_cast<const uint8_t *>(host), static_cast<const uint8_t *>(host) + bytes);
        return true;
    }, rc);
}

void frh::pin_tables(frh_context *ctx, const DevArgs &a, int delta) {
    const void *named[] = {a.rowd, a.cold, a.gcal, a.sites, a.tone_tab, a.curve8, a.etab,
                           a.srgb16 == reinterpret_cast<const float2 *>(ctx->d_srgb16) ? nullptr : a.srgb16};
    std::lock_guard<std::mutex> g(ctx->mu);
    for (const void *p : named) {
        if (!p) continue;
        for (auto &kv : ctx->tables)
            if (kv.second.d == p) {
                kv.second.pins = (uint32_t)((int)kv.second.pins + delta);
                break;
            }
    }
}

extern "C" int32_t frh_context_table_cache(const frh_context *ctx, uint64_t *entries, uint64_t *bytes) {
    if (!ctx) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL context");
    std::lock_guard<std::mutex> g(const_cast<frh_context *>(ctx)->mu);
    if (entries) *entries = ctx->tables.size();
    if (bytes) *bytes = ctx->table_bytes;
    return FRH_OK;
}

// The device tap lists of an edge-aware spec (built once per distinct table contents and kept by
// the context): per canonical position, G0's non-zero taps and, per colour, the taps where some
// direction is non-zero, in tap order (edgeaware/mod.rs:130-160).
static int32_t edge_tables(frh_context *ctx, const frh_edge_spec *e, const EdgeTab **out) {
    const uint32_t np = e->period * e->period;  // canonical positions: 4 (Bayer) or 36 (X-Trans)
    // the key: position count, canonical colours, G0 and directional filters as given
    std::vector<uint8_t> key(4 + np + (size_t)np * 49 * 4 + (size_t)np * 3 * 4 * 49 * 4);
    memcpy(key.data(), &np, 4);
    memcpy(key.data() + 4, e->canon, np);
    memcpy(key.data() + 4 + np, e->g0, (size_t)np * 49 * 4);
    memcpy(key.data() + 4 + np + (size_t)np * 49 * 4, e->dir, (size_t)np * 3 * 4 * 49 * 4);
    int32_t rc = FRH_OK;
    const void *d = cached_table(ctx, 1, key.data(), key.size(), [&](std::vector<uint8_t> &dev) {
        dev.assign(sizeof(EdgeTab), 0);
        EdgeTab &t = *reinterpret_cast<EdgeTab *>(dev.data());
        for (uint32_t ph = 0; ph < np; ph++) {
            t.canon[ph] = std::min<uint32_t>(e->canon[ph], 3);
            for (uint32_t i = 0; i < 49; i++) {
                const float c = e->g0[ph * 49 + i];
                if (c != 0.0f) {
                    t.g0_idx[ph][t.g0_n[ph]] = i;
                    t.g0_c[ph][t.g0_n[ph]] = c;
                    t.g0_n[ph]++;
                }
                if (ph < 4) t.g0d[ph][i] = c;
            }
            for (int c = 0; c < 3; c++)
                for (uint32_t i = 0; i < 49; i++) {
                    float v[4];
                    bool any = false;
                    for (int dd = 0; dd < 4; dd++) {
                        v[dd] = e->dir[((ph * 3 + c) * 4 + dd) * 49 + i];
                        if (ph < 4) t.dird[ph][c][i][dd] = v[dd];
                        any = any || v[dd] != 0.0f;
                    }
                    if (!any) continue;
                    const uint32_t n = t.dir_n[ph][c]++;
                    t.dir_idx[ph][c][n] = i;
                    for (int dd = 0; dd < 4; dd++) t.dir_c[ph][c][n][dd] = v[dd];
                }
        }
        return true;
    }, &rc);
    if (rc) return rc;
    *out = static_cast<const EdgeTab *>(d);
    return FRH_OK;
}

// True when the scalar calibration, the white balance and the matrix bound the values of the AHD
// walker (ahd.rs:143-298) so that
//  - every value it takes a square root of is 0 or at least 2^-96, and finite, and
//  - every product of the RGBA8 finish (matrix coefficient times clipped channel) is 0 or normal.
// The proof: a sample v (an integer) minus a black level b with j fractional bits is a multiple of
// 2^-j, so a non-zero normalised and balanced sample m = min((v - b) * scale * wb, 1) has
// |m| >= 2^-j |scale| |wb| (1 - 2^-20) >= 2^-40, and every float of that size is a multiple of
// q = 2^-63. Sums and differences of multiples of q round to multiples of q, and the walker's
// only other operations before the square roots are clamps and the exact scalings by 2, 0.5 and
// 0.25, at most four halvings deep (the diagonal colour difference), so every square-root
// operand is a multiple of 2^-67: 0 or >= 2^-67. The blend w * a + (1 - w) * b of two such
// candidates, w = 0 or >= 1/72, is 0 or >= 2^-97, so a matrix coefficient of at least 2^-20
// times a non-zero channel is at least 2^-117, a normal float; then scaling the coefficients
// by 4096 scales every product and sum exactly. The bound 2^60 on |m| keeps every value (sums
// of a few m, their square roots, squared differences) far from overflow. With positive scales
// and white balance every zero is +0 (x - x is +0, and nothing before the finish multiplies by a
// negative number), so choosing between equal values, as min, max and clamp do, cannot flip the
// sign of a zero either. The bilinear walker's RGBA8 finish relies on the same bound: its channels
// are means of at most four normalised samples times the white balance, so a non-zero one is at
// least 2^-66 and its products with the pre-scaled matrix are exact too.
static bool ahd_tame(const DevArgs &a) {
    if (!a.bounded || !a.simple_cal) return false;
    double wmin = 1e300, wmax = 0.0;
    for (int c = 0; c < 3; c++) {
        if (!(a.wb[c] > 0.0f)) return false;
        wmin = std::min(wmin, (double)a.wb[c]);
        wmax = std::max(wmax, (double)a.wb[c]);
    }
    for (int i = 0; i < 4; i++) {
        const double b = a.qb[i], k = a.qs[i];
        if (!std::isfinite(b) || !std::isfinite(k) || !(k > 0.0)) return false;
        int j = 0;
        for (double f = b; f != std::floor(f); f *= 2.0)
            if (++j > 40) return false;
        if (std::ldexp(k * wmin * (1.0 - 0x1p-20), -j) < 0x1p-40) return false;
        if ((65535.0 + std::fabs(b)) * k * wmax > 0x1p60) return false;
    }
    for (int i = 0; i < 9; i++)
        if (a.m[i] != 0.0f && std::fabs(a.m[i]) < 0x1p-20f) return false;
    return true;
}

// True when the inputs bound every intermediate value of the edge-aware demosaic: with
// |balanced sample| <= M and |tap| <= C, a 49-tap sum is at most 49 M C <= 1e15, so sqrt(G0) <=
// 3.2e7, the gradient energies E <= 4e16, (E + eps)^2 <= 2e33 stays finite and 1 / that stays
// normal, the colour planes stay below 1e16 and the gate terms (d * inv)^2 below 1e36. Every value
// the kernel multiplies by a zero tap or a zero colour mask is then finite, so skipping those
// products (or multiplying them) leaves every sum unchanged.
static bool edge_tame(const DevArgs &a, const frh_frame_info *info, const frh_edge_spec *e) {
    if (!a.bounded) return false;
    double bmax = 0.0, smax = 0.0, wmax = 1.0, cmax = 0.0;
    for (uint32_t i = 0; i < a.cal_pw * a.cal_ph; i++) {
        bmax = std::max(bmax, (double)std::fabs(a.black[i]));
        smax = std::max(smax, (double)std::fabs(a.scale[i]));
    }
    double rmax = 0.0, cdmax = 0.0;
    for (uint32_t i = 0; i < a.n_rowd; i++) rmax = std::max(rmax, (double)std::fabs(info->black_row_deltas[i]));
    for (uint32_t i = 0; i < a.n_cold; i++) cdmax = std::max(cdmax, (double)std::fabs(info->black_col_deltas[i]));
    for (int i = 0; i < 3; i++) wmax = std::max(wmax, (double)std::fabs(a.wb[i]));
    const uint32_t np = e->period * e->period;
    for (uint32_t i = 0; i < np * 49; i++) cmax = std::max(cmax, (double)std::fabs(e->g0[i]));
    for (uint32_t i = 0; i < np * 3 * 4 * 49; i++) cmax = std::max(cmax, (double)std::fabs(e->dir[i]));
    const double m = (65535.0 + bmax + rmax + cdmax) * smax * wmax * 1.001;
    return 49.0 * m * cmax <= 1e15;
}

// The display outputs' tone and output curve (run.rs:626-675) into a's table pointers.
int32_t frh::output_tables(frh_context *ctx, const frh_tone *tone, const frh_curve *curve, DevArgs &a) {
    int32_t rc = FRH_OK;
    if (tone) {
        const frh_tone *t = tone;
        if (t->size < sizeof(frh_tone)) return fail(FRH_ERROR_INVALID_ARGUMENT, "tone size field too small");
        if (t->table_len) {
            if (!t->table) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL tone table");
            a.tone = 1;
            a.tone_tab = static_cast<const float2 *>(device_table(ctx, t->table, t->table_len * 8ull, &rc));
            if (rc) return rc;
            a.tone_n = t->table_len;
            a.tone_xmax = t->xmax;
            a.tone_scale = t->scale;
            memcpy(a.tone_m, t->to_prophoto, sizeof(a.tone_m));
            memcpy(a.tone_w, t->white, sizeof(a.tone_w));
        }
    }
    if (curve) {
        const frh_curve *c = curve;
        if (c->size < sizeof(frh_curve)) return fail(FRH_ERROR_INVALID_ARGUMENT, "curve size field too small");
        if (c->lut8_len > 1) {
            if (!c->lut8) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL curve table");
            a.curve8 = static_cast<const uint8_t *>(device_table(ctx, c->lut8, c->lut8_len, &rc));
            if (rc) return rc;
            a.curve8_n = c->lut8_len - 1;
        }
        if (c->lut16_len == 16385) {
            if (!c->lut16) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL curve table");
            // (value, step) pairs: the steps are the f32 differences srgb16 computes (process.rs:1953-1958)
            std::vector<float> pairs(2 * 16384);
            for (int i = 0; i < 16384; i++) {
                pairs[2 * i] = c->lut16[i];
                pairs[2 * i + 1] = c->lut16[i + 1] - c->lut16[i];
            }
            a.srgb16 = static_cast<const float2 *>(device_table(ctx, pairs.data(), pairs.size() * 4ull, &rc));
            if (rc) return rc;
        }
    }
    return FRH_OK;
}

int32_t frh::plan_develop(const frh_context *ctx, const frh_frame *f, const frh_frame_info *info,
                            const frh_develop_spec *spec, Plan &p) {
    if (!f || !info || !spec) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL descriptor");
    if (f->size < sizeof(frh_frame) || info->size < sizeof(frh_frame_info) ||
        spec->size < offsetof(frh_develop_spec, cancel))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "descriptor size field too small");
    memset(&p, 0, sizeof(p));
    DevArgs &a = p.a;
    int32_t rc = check_input(f, p.in);
    if (rc) return rc;
    const uint32_t w = p.in.w, h = p.in.h, cpp = p.in.cpp;
    const bool has_cfa = info->cfa_width != 0 && info->cfa_height != 0;
    if (has_cfa && (info->cfa_width > 8 || info->cfa_height > 8))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "CFA pattern larger than 8x8");
    bool bayer = false;
    if (has_cfa && info->cfa_width == 2 && info->cfa_height == 2) {
        int cnt[256] = {0};
        for (int i = 0; i < 4; i++) cnt[info->cfa[i]]++;
        bayer = cnt[0] == 1 && cnt[1] == 2 && cnt[2] == 1;
    }
    if (!valid_mode(spec->output)) return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown output");
    if (spec->output2 != OUT_NONE && !valid_mode(spec->output2))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown second output");
    for (int i = 0; i < 4; i++)
        if (!std::isfinite(spec->wb[i])) return fail(FRH_ERROR_INVALID_ARGUMENT, "non-finite white balance");
    for (int i = 0; i < 9; i++)
        if (!std::isfinite(spec->cam_to_rgb[i])) return fail(FRH_ERROR_INVALID_ARGUMENT, "non-finite matrix");
    if (spec->has_levels && !(std::isfinite(spec->levels[0]) && std::isfinite(spec->levels[1])))
        return fail(FRH_ERROR_INVALID_ARGUMENT, "non-finite levels");

    // Samples: the window of the raster (the upload replaces the pointer for host input; staged
    // input is replaced by its unpacked u16 window).
    const uint64_t row_samples = (uint64_t)f->raster_width * cpp;
    if (p.in.kind == FRH_INPUT_F32)
        a.srcf = static_cast<const float *>(f->data) + (size_t)f->window.y * row_samples + (size_t)f->window.x * cpp;
    else if (p.in.kind == FRH_INPUT_U16)
        a.src = static_cast<const uint16_t *>(f->data) + (size_t)f->window.y * row_samples + (size_t)f->window.x * cpp;
    a.stride = (uint32_t)(p.in.staged ? (uint64_t)w * cpp : row_samples);
    a.cpp = cpp;
    a.width = w;
    a.height = h;
    a.ox = f->origin_x;
    a.oy = f->origin_y;
    // Calib::new (process.rs:256-286) for every component
    std::vector<float> gcal;
    uint32_t pw, ph;
    if ((rc = build_calib(info, cpp, gcal, pw, ph))) return rc;
    a.cal_pw = pw;
    a.cal_ph = ph;
    a.cal_pow2 = ((pw & (pw - 1)) == 0 && (ph & (ph - 1)) == 0) ? 1u : 0u;
    for (uint32_t i = 0; i < pw * ph; i++) {
        a.black[i] = gcal[i];
        a.scale[i] = gcal[(size_t)cpp * pw * ph + i];
    }
    frh_context *mctx = const_cast<frh_context *>(ctx);
    if (info->black_row_deltas_len) {
        if (!info->black_row_deltas) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL row deltas");
        a.n_rowd = info->black_row_deltas_len;
        a.rowd = static_cast<const float *>(device_table(mctx, info->black_row_deltas, a.n_rowd * 4ull, &rc));
        if (rc) return rc;
    }
    if (info->black_col_deltas_len) {
        if (!info->black_col_deltas) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL column deltas");
        a.n_cold = info->black_col_deltas_len;
        a.cold = static_cast<const float *>(device_table(mctx, info->black_col_deltas, a.n_cold * 4ull, &rc));
        if (rc) return rc;
    }
    // The generic kernel: every frame but a single-component Bayer mosaic at full size.
    a.half = spec->half_size ? 1u : 0u;
    a.generic = (cpp == 1 && bayer && !spec->half_size) ? 0u : 1u;
    if (a.generic) {
        a.gmode = (cpp == 1 && has_cfa) ? 1u : (cpp >= 3 ? 2u : 0u);
        a.gcal = static_cast<const float *>(device_table(mctx, gcal.data(), gcal.size() * 4ull, &rc));
        if (rc) return rc;
        if (has_cfa) {
            // cfa_sites (kernels.rs:260-287): colour (clamped to 3) and, per colour R, G, B, the 9-bit
            // mask of 3x3 neighbours of that colour (the site's own colour: the centre only)
            const uint32_t cw = info->cfa_width, chh = info->cfa_height;
            std::vector<uint32_t> sites(cw * chh);
            for (uint32_t py = 0; py < chh; py++)
                for (uint32_t px = 0; px < cw; px++) {
                    const uint32_t own = std::min<uint32_t>(info->cfa[py * cw + px], 3);
                    uint32_t d = own;
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const uint32_t c = std::min<uint32_t>(
                                info->cfa[((py + dy + 4 * chh) % chh) * cw + (px + dx + 4 * cw) % cw], 3);
                            if (c < 3 && (c != own || (dx == 0 && dy == 0)))
                                d |= 1u << (2 + 9 * c + (uint32_t)((dy + 1) * 3 + dx + 1));
                        }
                    sites[py * cw + px] = d;
                }
            a.cfa_w = cw;
            a.cfa_h = chh;
            a.sites = static_cast<const uint32_t *>(device_table(mctx, sites.data(), sites.size() * 4ull, &rc));
            if (rc) return rc;
        }
    }
    if (bayer) {
        for (int i = 0; i < 4; i++) {
            uint32_t colour = info->cfa[i];  // pattern position (i & 1, i >> 1) of sensor parity
            a.chan[i] = colour < 3 ? colour : 3;
            a.chan_bits |= a.chan[i] << (2 * i);
        }
    }
    // The walkers' scalar calibration (pattern of at most 2x2, no deltas) and the bound that lets
    // the RGBA8 fast path clamp with v_med3_f32.
    a.simple_cal = (pw <= 2 && ph <= 2 && !a.n_rowd && !a.n_cold) ? 1u : 0u;
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t X = (i & 1) + a.ox, Y = (i >> 1) + a.oy;
        const uint32_t k = (Y % a.cal_ph) * a.cal_pw + X % a.cal_pw;
        a.qb[i] = a.black[k];
        a.qs[i] = a.scale[k];
    }
    {
        bool ok = true;
        auto fine = [](float v) { return std::isfinite(v) && std::fabs(v) <= 1e6f; };
        for (int i = 0; i < 4; i++) ok = ok && fine(spec->wb[i]);
        for (int i = 0; i < 9; i++) ok = ok && fine(spec->cam_to_rgb[i]);
        if (spec->has_levels) ok = ok && fine(spec->levels[0]) && fine(spec->levels[1]);
        for (uint64_t i = 0; i < pw * ph; i++) ok = ok && fine(a.black[i]);
        for (uint32_t i = 0; i < a.n_rowd; i++) ok = ok && fine(info->black_row_deltas[i]);
        for (uint32_t i = 0; i < a.n_cold; i++) ok = ok && fine(info->black_col_deltas[i]);
        a.bounded = ok ? 1u : 0u;
    }
    // crop.clamp_to(w, h) (crates/flashraw-core/src/types.rs:68-70)
    uint64_t cx = spec->crop.x, cy = spec->crop.y;
    uint64_t cw = spec->crop.width, chh = spec->crop.height;
    if (cx >= w || cy >= h || cw == 0 || chh == 0) return fail(FRH_ERROR_INVALID_ARGUMENT, "develop region too small");
    cw = cw < w - cx ? cw : w - cx;
    chh = chh < h - cy ? chh : h - cy;
    a.crop_x = (uint32_t)cx;
    a.crop_y = (uint32_t)cy;
    a.dev_w = (uint32_t)(spec->half_size ? cw / 2 : cw);
    a.dev_h = (uint32_t)(spec->half_size ? chh / 2 : chh);
    if (a.dev_w == 0 || a.dev_h == 0) return fail(FRH_ERROR_INVALID_ARGUMENT, "develop region too small");
    a.band_y0 = 0;
    a.band_y1 = a.dev_h;
    memcpy(a.wb, spec->wb, sizeof(a.wb));
    memcpy(a.m, spec->cam_to_rgb, sizeof(a.m));
    a.unit_k = 1u;
    for (int c = 0; c < 3; c++) {
        a.kk[c] = a.wb[c] / a.wb[c];
        if (a.kk[c] != 1.0f) a.unit_k = 0u;
    }
    a.ahd_tame = ahd_tame(a) ? 1u : 0u;
    for (int i = 0; i < 9; i++) a.m4k[i] = a.m[i] * 4096.0f;
    a.has_levels = spec->has_levels ? 1 : 0;
    a.lscale = spec->levels[0];
    a.loffset = spec->levels[1];
    a.orientation = (spec->orientation >= 1 && spec->orientation <= 8) ? spec->orientation : 1;
    const bool swap = a.orientation >= 5;
    a.dst_w = swap ? a.dev_h : a.dev_w;
    a.out_mode = spec->output;
    a.out_mode2 = spec->output2;
    a.lut8 = ctx->d_lut8;
    a.wtab = ctx->d_wtab;
    a.srgb16 = reinterpret_cast<const float2 *>(ctx->d_srgb16);
    p.ahd = false;
    {
        const char *impl = getenv("FLASHRAW_BILINEAR_IMPL");
        a.bil_impl = (impl && strcmp(impl, "tiled") == 0) ? 1u : 0u;
        const char *rows = getenv("FLASHRAW_BILINEAR_ROWS"), *wpb = getenv("FLASHRAW_BILINEAR_WPB"),
                   *cpl = getenv("FLASHRAW_BILINEAR_CPL");
        const auto t = tuned_shape(ctx, false, a, ctx->bil_rows);
        a.bil_rows = rows ? (uint32_t)atoi(rows) : t.rows;
        a.bil_wpb = wpb ? (uint32_t)atoi(wpb) : t.wpb;
        a.bil_cpl = cpl ? (uint32_t)atoi(cpl) : t.cpl;
    }
    // The display outputs' tone and output curve (run.rs:626-675)
    if ((rc = output_tables(mctx, spec->tone, spec->curve, a))) return rc;
    // run.rs:545: the edge-aware and AHD demosaics apply to single-component CFA frames at full size
    if (spec->edge && !spec->half_size && cpp == 1 && has_cfa) {
        const frh_edge_spec *e = spec->edge;
        if (e->size < sizeof(frh_edge_spec)) return fail(FRH_ERROR_INVALID_ARGUMENT, "edge spec size too small");
        const uint32_t P = e->period;
        if (e->ahd && P != 2) return fail(FRH_ERROR_NOT_AVAILABLE, "the AHD demosaic takes period-2 specs");
        if (e->ahd ? !bayer : !(P == 2 ? bayer : (P == 6 && info->cfa_width == 6 && info->cfa_height == 6)))
            return fail(FRH_ERROR_NOT_AVAILABLE,
                        "the AHD demosaic takes Bayer mosaics, the edge-aware one Bayer (period 2) or 6x6 "
                        "(period 6) mosaics on this path");
        if (e->canon) {
            if (e->canon_len != P * P) return fail(FRH_ERROR_INVALID_ARGUMENT, "inconsistent edge-aware filter tables");
            // The pattern the spec describes must be the frame's: colour of frame (x, y) is
            // canon[((y + cy) % P) * P + (x + cx) % P].
            for (uint32_t y = 0; y < P; y++)
                for (uint32_t x = 0; x < P; x++) {
                    const uint32_t want =
                        P == 2 ? a.chan[((y + a.oy) & 1) * 2 + ((x + a.ox) & 1)]
                               : std::min<uint32_t>(info->cfa[((y + a.oy) % 6) * 6 + (x + a.ox) % 6], 3);
                    if (std::min<uint32_t>(e->canon[((y + e->cy) % P) * P + (x + e->cx) % P], 3) != want)
                        return fail(FRH_ERROR_INVALID_ARGUMENT, "edge spec pattern does not match the frame's CFA");
                }
        }
        if (e->ahd) {
            // run.rs:588-598: AHD needs period 2 and beta 0
            if (e->beta != 0.0f) return fail(FRH_ERROR_INVALID_ARGUMENT, "inconsistent AHD spec");
            if (e->halo < 5 || e->halo > 64) return fail(FRH_ERROR_INVALID_ARGUMENT, "AHD halo must be in [5, 64]");
            a.ahd_halo = e->halo;
            // Implementation knobs (environment, like the reference's FLASHRAW_* variables): the
            // LDS-tiled AHD kernel instead of the row walker, and the walker's rows per strip.
            const char *impl = getenv("FLASHRAW_AHD_IMPL");
            a.ahd_impl = (impl && strcmp(impl, "tiled") == 0) ? 1u : 0u;
            const char *rows = getenv("FLASHRAW_AHD_ROWS"), *wpb = getenv("FLASHRAW_AHD_WPB");
            const auto t = tuned_shape(ctx, true, a, ctx->ahd_rows);
            a.ahd_rows = rows ? (uint32_t)atoi(rows) : t.rows;
            a.ahd_wpb = wpb ? (uint32_t)atoi(wpb) : t.wpb;
            p.ahd = true;
        } else {
            // the edge-aware filters (edgeaware/mod.rs:57-160): P * P canonical positions
            if (!e->canon || !e->g0 || !e->dir || e->g0_len != P * P * 49 || e->dir_len != P * P * 3 * 4 * 49)
                return fail(FRH_ERROR_INVALID_ARGUMENT,
                            "edge-aware spec needs canon[P*P], g0[P*P*49] and dir[P*P*3*4*49]");
            if (!(e->beta >= 0.0f && e->beta <= 1.0f))
                return fail(FRH_ERROR_INVALID_ARGUMENT, "edge-aware beta must be in [0, 1]");
            if (e->halo < 8 || e->halo > 64)
                return fail(FRH_ERROR_INVALID_ARGUMENT, "edge-aware halo must be in [8, 64]");
            for (uint32_t i = 0; i < P * P * 49; i++)
                if (!std::isfinite(e->g0[i])) return fail(FRH_ERROR_INVALID_ARGUMENT, "non-finite filter tap");
            for (uint32_t i = 0; i < P * P * 3 * 4 * 49; i++)
                if (!std::isfinite(e->dir[i])) return fail(FRH_ERROR_INVALID_ARGUMENT, "non-finite filter tap");
            int32_t rc = edge_tables(const_cast<frh_context *>(ctx), e, &a.etab);
            if (rc) return rc;
            a.edge_aware = 1;
            a.eperiod = P;
            a.ecx = e->cx % P;
            a.ecy = e->cy % P;
            a.ebeta = e->beta;
            // xg = (1 - 2 beta).clamp(0, 1) (edgeaware/mod.rs:547-550), f32 as the reference
            {
                const float xg = 1.0f - 2.0f * e->beta;
                a.exg = xg < 0.0f ? 0.0f : (xg > 1.0f ? 1.0f : xg);
            }
            a.ahd_halo = e->halo;
            a.ecanon_bits = 0;
            if (P == 2)
                for (int i = 0; i < 4; i++) a.ecanon_bits |= (uint32_t)(e->canon[i] & 3u) << (2 * i);
            // 6x6 mosaics: the border pixels take the generic CFA average (sites, gcal stay set)
            a.generic = 0;
            a.edge_tame = (P == 2 && edge_tame(a, info, e)) ? 1u : 0u;
            const char *impl = getenv("FLASHRAW_EDGE_IMPL");  // "general": the tap-list kernel always
            if (impl && strcmp(impl, "general") == 0) a.edge_tame = 0;
        }
    }
    int32_t modes[2] = {spec->output, spec->output2};
    for (int k = 0; k < 2; k++) {
        if (k == 1 && modes[1] == OUT_NONE) break;
        uint32_t bps;
        out_format(modes[k], &p.ch[k], &bps, &p.fmt[k]);
        bool oriented = display_mode(modes[k]);
        p.out_w[k] = oriented && swap ? a.dev_h : a.dev_w;
        p.out_h[k] = oriented && swap ? a.dev_w : a.dev_h;
        p.bytes[k] = (uint64_t)p.out_w[k] * p.out_h[k] * p.ch[k] * bps;
    }
    if (!a.generic && (a.srcf || a.tone || a.curve8)) {
        // the register walkers read u16 rows and encode with the sRGB table: f32 samples, tone
        // and output curves take the LDS-tiled kernels
        a.bil_impl = 1;
        a.ahd_impl = 1;
    }
    p.in_row_bytes = p.in.kind == FRH_INPUT_PACKED ? std::max(p.in.stride, p.in.row_bytes)
                                                    : row_samples * (p.in.kind == FRH_INPUT_F32 ? 4 : 2);
    return FRH_OK;
}

// Developed rows per launch under the context's binding cap (BandGeometry::plan, run.rs:357-389):
// every band's input rows (+ halo) and output rows must fit; at least one 16-row tile.
static uint32_t band_rows(const frh_context *ctx, const Plan &p) {
    if (!ctx->max_binding) return p.a.dev_h;
    uint64_t halo = (p.ahd || p.a.edge_aware) ? p.a.ahd_halo : 1;
    uint64_t out_row = 0;
    for (int k = 0; k < 2; k++) out_row += p.bytes[k] / (p.out_h[k] ? p.out_h[k] : 1);
    uint64_t rows_in = ctx->max_binding / (p.in_row_bytes ? p.in_row_bytes : 1);
    uint64_t n_in = rows_in > 2 * halo ? rows_in - 2 * halo : 0;
    uint64_t n_out = ctx->max_binding / (out_row ? out_row : 1);
    uint64_t n = n_in < n_out ? n_in : n_out;
    if (n >= 16) n -= n % 16;
    if (n == 0) n = 1;
    return n < p.a.dev_h ? (uint32_t)n : p.a.dev_h;
}

// Launch the develop band by band; `cancel` is checked before every band.
static int32_t launch_bands(frh_context *ctx, Plan &p, hipStream_t stream, uint32_t *nbands,
                            const frh_cancel_token *cancel = nullptr) {
    if (p.a.orientation >= 5 && ctx->max_binding)
        return fail(FRH_ERROR_LIMIT_EXCEEDED, "banded runs of rotated outputs are not supported");
    uint32_t n = band_rows(ctx, p);
    uint32_t bands = 0;
    for (uint32_t y0 = 0; y0 < p.a.dev_h; y0 += n) {
        if (cancelled(cancel)) return cancel_fail();
        DevArgs a = p.a;
        a.band_y0 = y0;
        a.band_y1 = y0 + n < p.a.dev_h ? y0 + n : p.a.dev_h;
        hipError_t e = launch_develop(a, p.ahd, stream, ctx->num_cus);
        if (e != hipSuccess) return hip_fail(e, "kernel launch");
        bands++;
    }
    if (nbands) *nbands = bands;
    return FRH_OK;
}

extern "C" int32_t frh_develop_output_size(const frh_frame *f, const frh_develop_spec *spec, int32_t output,
                                           uint32_t *width, uint32_t *height, uint32_t *channels, uint64_t *bytes) {
    if (!f || !spec) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL descriptor");
    if (!valid_mode(output)) return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown output");
    uint64_t w = f->window.width, h = f->window.height;
    uint64_t cx = spec->crop.x, cy = spec->crop.y, cw = spec->crop.width, ch = spec->crop.height;
    if (cx >= w || cy >= h || cw == 0 || ch == 0) return fail(FRH_ERROR_INVALID_ARGUMENT, "develop region too small");
    cw = cw < w - cx ? cw : w - cx;
    ch = ch < h - cy ? ch : h - cy;
    bool swap = display_mode(output) && spec->orientation >= 5 && spec->orientation <= 8;
    uint32_t c, bps;
    int32_t fmt;
    out_format(output, &c, &bps, &fmt);
    if (width) *width = (uint32_t)(swap ? ch : cw);
    if (height) *height = (uint32_t)(swap ? cw : ch);
    if (channels) *channels = c;
    if (bytes) *bytes = cw * ch * c * bps;
    return FRH_OK;
}

// Put the frame's samples where the develop reads them: upload the window's rows of host input
// into a pooled buffer (*d_in), and unpack packed or linearised input into a pooled u16 window
// (*d_stage), on `st`. Points p.a.src / p.a.srcf at the result.
static int32_t stage_frame(frh_context *ctx, const frh_frame *frame, Plan &p, hipStream_t st, void **d_in,
                           uint64_t *in_bytes, void **d_stage, uint64_t *stage_bytes) {
    const InputDesc &in = p.in;
    hipError_t e;
    const uint8_t *base = static_cast<const uint8_t *>(frame->data);
    uint64_t first_row = frame->window.y;  // raster row at `base`
    if (!frame->data_on_device) {
        *in_bytes = in.upload_bytes ? in.upload_bytes : 4;
        *d_in = ctx->alloc(*in_bytes, &e);
        if (!*d_in) return hip_fail(e, "input allocation");
        if (in.upload_bytes &&
            (e = hipMemcpyAsync(*d_in, base + in.first_byte, in.upload_bytes, hipMemcpyHostToDevice, st)) != hipSuccess)
            return hip_fail(e, "upload");
        base = static_cast<const uint8_t *>(*d_in);
        first_row = 0;
    }
    const uint64_t bps = in.kind == FRH_INPUT_F32 ? 4 : 2;
    if (!in.staged) {
        const uint64_t off = (first_row * in.stride + (uint64_t)frame->window.x * in.cpp) * bps;
        if (in.kind == FRH_INPUT_F32) {
            p.a.srcf = reinterpret_cast<const float *>(base + off);
            p.a.src = nullptr;
        } else {
            p.a.src = reinterpret_cast<const uint16_t *>(base + off);
        }
        return FRH_OK;
    }
    InArgs a;
    memset(&a, 0, sizeof(a));
    a.kind = (uint32_t)in.kind;
    a.src = base;
    a.len = frame->data_on_device ? in.len : in.upload_bytes;
    a.stride = in.stride;
    a.row_bytes = (uint32_t)in.row_bytes;
    a.bits = in.bits;
    a.x0 = frame->window.x * in.cpp;
    a.y0 = (uint32_t)first_row;
    a.n = in.w * in.cpp;
    a.rows = in.h;
    if (in.lut) {
        int32_t rc;
        a.lut = static_cast<const uint16_t *>(device_table(ctx, frame->linearization, frame->linearization_len * 2ull, &rc));
        if (rc) return rc;
        a.lut_len = frame->linearization_len;
    }
    *stage_bytes = ((uint64_t)a.n * a.rows * 2 + 3) & ~3ull;
    *d_stage = ctx->alloc(*stage_bytes, &e);
    if (!*d_stage) return hip_fail(e, "unpack allocation");
    if ((e = launch_unpack(a, in.order, static_cast<uint16_t *>(*d_stage), st)) != hipSuccess)
        return hip_fail(e, "unpack launch");
    p.a.src = static_cast<const uint16_t *>(*d_stage);
    p.a.srcf = nullptr;
    return FRH_OK;
}

extern "C" int32_t frh_develop_into(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info,
                                    const frh_develop_spec *spec, void *dst, uint64_t dst_bytes, void *dst2,
                                    uint64_t dst2_bytes, void *stream) {
    try {
        if (!ctx) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL context");
        Plan p;
        int32_t rc = plan_develop(ctx, frame, info, spec, p);
        if (rc) return rc;
        if (!frame->data_on_device) return fail(FRH_ERROR_INVALID_ARGUMENT, "frh_develop_into needs device input");
        if (!dst || dst_bytes < p.bytes[0]) return fail(FRH_ERROR_INVALID_ARGUMENT, "output buffer too small");
        if (spec->output2 != OUT_NONE && (!dst2 || dst2_bytes < p.bytes[1]))
            return fail(FRH_ERROR_INVALID_ARGUMENT, "second output buffer too small");
        p.a.dst = dst;
        p.a.dst2 = dst2;
        const frh_cancel_token *cancel = spec_cancel(spec);
        if (cancelled(cancel)) return cancel_fail();
        (void)hipSetDevice(ctx->device);
        hipStream_t st = (hipStream_t)stream;
        void *d_in = nullptr, *d_stage = nullptr;
        uint64_t in_bytes = 0, stage_bytes = 0;
        rc = stage_frame(ctx, frame, p, st, &d_in, &in_bytes, &d_stage, &stage_bytes);
        if (rc == FRH_OK) rc = launch_bands(ctx, p, st, nullptr, cancel);
        if (d_stage) {
            // the unpacked window is pooled scratch: wait for the develop before returning it
            hipError_t e = hipStreamSynchronize(st);
            ctx->release(d_stage, stage_bytes);
            if (rc == FRH_OK && e != hipSuccess) rc = hip_fail(e, "develop");
        }
        return rc;
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

// ------------------------------------------------------------------------------------------
// Prepared plans

struct frh_plan {
    frh_context *ctx;
    Plan p;
    size_t src_off;  // samples from the raster start to the window origin
};

extern "C" void *frh_context_stream(const frh_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int32_t frh_plan_create(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info,
                                   const frh_develop_spec *spec, frh_plan **out) {
    try {
        if (!ctx || !out) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        std::unique_ptr<frh_plan> pl(new frh_plan());
        pl->ctx = ctx;
        int32_t rc = plan_develop(ctx, frame, info, spec, pl->p);
        if (rc) return rc;
        if (pl->p.in.staged)
            return fail(FRH_ERROR_NOT_AVAILABLE,
                        "prepared plans take u16 or f32 samples: unpack packed or linearised input first (frh_unpack_into)");
        pl->src_off = (size_t)frame->window.y * frame->raster_width * pl->p.in.cpp + (size_t)frame->window.x * pl->p.in.cpp;
        pin_tables(ctx, pl->p.a, +1);  // the plan's tables stay cached while it lives
        *out = pl.release();
        return FRH_OK;
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_plan_launch(const frh_plan *plan, const void *src, void *dst, void *dst2, void *stream) {
    if (!plan || !src || !dst) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (plan->p.a.out_mode2 != OUT_NONE && !dst2) return fail(FRH_ERROR_INVALID_ARGUMENT, "second output is NULL");
    Plan p = plan->p;
    if (p.in.kind == FRH_INPUT_F32)
        p.a.srcf = static_cast<const float *>(src) + plan->src_off;
    else
        p.a.src = static_cast<const uint16_t *>(src) + plan->src_off;
    p.a.dst = dst;
    p.a.dst2 = dst2;
    return launch_bands(plan->ctx, p, (hipStream_t)stream, nullptr);
}

extern "C" int32_t frh_plan_autotune(frh_plan *plan, const void *src, void *dst, void *dst2, void *stream,
                                     int32_t iters, uint32_t *best) {
    try {
        if (!plan || !src || !dst) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        if (plan->p.a.out_mode2 != OUT_NONE && !dst2) return fail(FRH_ERROR_INVALID_ARGUMENT, "second output is NULL");
        if (iters < 1) iters = 1;
        if (plan->p.a.edge_aware) {  // one fixed tile shape: nothing to tune
            if (best) *best = 0;
            return FRH_OK;
        }
        frh_context *ctx = plan->ctx;
        hipStream_t st = (hipStream_t)stream;
        const bool ahd = plan->p.ahd;
        // the tile shape: strip height (rows one wave develops) x waves per block; the LDS-tiled
        // alternates have no shape to tune
        if (ahd ? plan->p.a.ahd_impl != 0 : plan->p.a.bil_impl != 0) {
            if (best) *best = 0;
            return FRH_OK;
        }
        // the four-column bilinear walker is a candidate where the caller's buffers and the
        // frame's geometry allow it (walk4_ok)
        bool four = false;
        if (!ahd) {
            DevArgs t = plan->p.a;
            t.src = static_cast<const uint16_t *>(src) + plan->src_off;
            four = plan->p.in.kind != FRH_INPUT_F32 && walk4_ok(t);
        }
        std::vector<frh_context::TileShape> cands;
        for (uint32_t c = 2; c <= (four ? 4u : 2u); c += 2)
            for (uint32_t w = 1; w <= (ahd ? AHD_MAX_WAVES : BIL_MAX_WAVES); w *= 2) {
                if (ahd)
                    for (uint32_t r = 12; r <= 132; r += 6) cands.push_back({r, w, c});
                else
                    for (uint32_t r = 8; r <= 64; r += 4) cands.push_back({r, w, c});
            }
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            return fail(FRH_ERROR_GPU_UNAVAILABLE, "hipEventCreate failed");
        float best_ms = 1e30f;
        frh_context::TileShape bst = ahd ? frh_context::TileShape{plan->p.a.ahd_rows, plan->p.a.ahd_wpb, 2u}
                                         : frh_context::TileShape{plan->p.a.bil_rows, plan->p.a.bil_wpb,
                                                                  plan->p.a.bil_cpl};
        int32_t rc = FRH_OK;
        for (const auto &c : cands) {
            Plan p = plan->p;
            if (ahd) {
                p.a.ahd_rows = c.rows;
                p.a.ahd_wpb = c.wpb;
            } else {
                p.a.bil_rows = c.rows;
                p.a.bil_wpb = c.wpb;
                p.a.bil_cpl = c.cpl;
            }
            if (p.in.kind == FRH_INPUT_F32)
                p.a.srcf = static_cast<const float *>(src) + plan->src_off;
            else
                p.a.src = static_cast<const uint16_t *>(src) + plan->src_off;
            p.a.dst = dst;
            p.a.dst2 = dst2;
            if ((rc = launch_bands(ctx, p, st, nullptr)) != FRH_OK) break;  // warm
            (void)hipEventRecord(e0, st);
            for (int i = 0; i < iters && rc == FRH_OK; i++) rc = launch_bands(ctx, p, st, nullptr);
            (void)hipEventRecord(e1, st);
            if (rc != FRH_OK) break;
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(FRH_ERROR_GPU_UNAVAILABLE, "autotune launch failed");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best_ms) {
                best_ms = ms;
                bst = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != FRH_OK) return rc;
        if (ahd) {
            plan->p.a.ahd_rows = bst.rows;
            plan->p.a.ahd_wpb = bst.wpb;
        } else {
            plan->p.a.bil_rows = bst.rows;
            plan->p.a.bil_wpb = bst.wpb;
            plan->p.a.bil_cpl = bst.cpl;
        }
        {
            std::lock_guard<std::mutex> g(ctx->mu);
            ctx->tuned[tune_key(ahd, plan->p.a)] = bst;
        }
        if (best) *best = bst.rows;
        return FRH_OK;
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_plan_tile_shape(const frh_plan *plan, uint32_t *rows, uint32_t *waves_per_block,
                                       uint32_t *cols_per_lane) {
    if (!plan) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL plan");
    const DevArgs &a = plan->p.a;
    const bool walker = !a.generic && !a.edge_aware && (plan->p.ahd ? a.ahd_impl == 0 : a.bil_impl == 0);
    if (rows) *rows = walker ? (plan->p.ahd ? a.ahd_rows : a.bil_rows) : 0u;
    if (waves_per_block)
        *waves_per_block = walker ? (plan->p.ahd ? walk_wpb(a.ahd_wpb, AHD_MAX_WAVES) : walk_wpb(a.bil_wpb, BIL_MAX_WAVES)) : 0u;
    // (the four-column walker also needs an aligned source, checked per launch: the plan's own
    // frame pointer stands for it here)
    if (cols_per_lane) *cols_per_lane = walker ? (!plan->p.ahd && a.bil_cpl == 4 && walk4_ok(a) ? 4u : 2u) : 0u;
    return FRH_OK;
}

extern "C" uint64_t frh_plan_output_bytes(const frh_plan *plan, int32_t which) {
    if (!plan || which < 0 || which > 1) return 0;
    return plan->p.bytes[which];
}

extern "C" int32_t frh_plan_kernel(const frh_plan *plan) {
    if (!plan) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL plan");
    const DevArgs &a = plan->p.a;
    if (a.generic) return FRH_KERNEL_GENERIC;
    if (a.edge_aware) return a.edge_tame ? FRH_KERNEL_EDGE_TAME : FRH_KERNEL_EDGE_GENERAL;
    if (plan->p.ahd) return a.ahd_impl ? FRH_KERNEL_AHD_TILED : FRH_KERNEL_AHD_WALK;
    return a.bil_impl ? FRH_KERNEL_BILINEAR_TILED : FRH_KERNEL_BILINEAR_WALK;
}

extern "C" void frh_plan_destroy(frh_plan *plan) {
    if (plan) pin_tables(plan->ctx, plan->p.a, -1);
    delete plan;
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

extern "C" int32_t frh_run_develop(frh_context *ctx, const frh_frame *frame, const frh_frame_info *info,
                                   const frh_develop_spec *spec, frh_image **out, frh_image **out2,
                                   frh_timings *timings) {
    try {
        if (!ctx || !out) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        if (out2) *out2 = nullptr;
        Plan p;
        int32_t rc = plan_develop(ctx, frame, info, spec, p);
        if (rc) return rc;
        if (spec->output2 != OUT_NONE && !out2) return fail(FRH_ERROR_INVALID_ARGUMENT, "second output needs out2");
        const frh_cancel_token *cancel = spec_cancel(spec);
        if (cancelled(cancel)) return cancel_fail();
        (void)hipSetDevice(ctx->device);
        hipError_t e;
        auto t0 = std::chrono::steady_clock::now();
        // Upload: the window's rows of the raster (run.rs:915-923), unless resident; packed and
        // linearised input is unpacked to u16 on the device first.
        void *d_in = nullptr, *d_stage = nullptr;
        uint64_t in_bytes = 0, stage_bytes = 0;
        auto cleanup = [&]() {
            if (d_in) ctx->release(d_in, in_bytes);
            if (d_stage) ctx->release(d_stage, stage_bytes);
        };
        rc = stage_frame(ctx, frame, p, ctx->stream, &d_in, &in_bytes, &d_stage, &stage_bytes);
        if (rc) {
            cleanup();
            return rc;
        }
        if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
            cleanup();
            return hip_fail(e, "upload");
        }
        double upload_ms = ms_since(t0);
        if (cancelled(cancel)) {  // between the upload and the develop (crates/flashraw/src/gpu.rs:341-344)
            cleanup();
            return cancel_fail();
        }
        auto t1 = std::chrono::steady_clock::now();
        frh_image *imgs[2] = {nullptr, nullptr};
        for (int k = 0; k < 2; k++) {
            if (k == 1 && spec->output2 == OUT_NONE) break;
            void *d = ctx->alloc(p.bytes[k], &e);
            if (!d) {
                if (imgs[0]) frh_image_free(imgs[0]);
                cleanup();
                return hip_fail(e, "output allocation");
            }
            imgs[k] = new frh_image{ctx, d, p.bytes[k], p.out_w[k], p.out_h[k], p.ch[k], p.fmt[k]};
        }
        p.a.dst = imgs[0]->dptr;
        p.a.dst2 = imgs[1] ? imgs[1]->dptr : nullptr;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        const bool timed = hipEventCreate(&ev0) == hipSuccess && hipEventCreate(&ev1) == hipSuccess;
        if (timed) (void)hipEventRecord(ev0, ctx->stream);
        uint32_t bands = 0;
        rc = launch_bands(ctx, p, ctx->stream, &bands, cancel);
        if (timed) (void)hipEventRecord(ev1, ctx->stream);
        e = hipStreamSynchronize(ctx->stream);
        float kms = 0.0f;
        if (timed && hipEventElapsedTime(&kms, ev0, ev1) != hipSuccess) kms = 0.0f;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        cleanup();
        if (rc == FRH_OK && e != hipSuccess) rc = hip_fail(e, "develop");
        if (rc == FRH_OK && cancelled(cancel)) rc = cancel_fail();  // after the develop (gpu.rs:475)
        if (rc != FRH_OK) {
            frh_image_free(imgs[0]);
            frh_image_free(imgs[1]);
            return rc;
        }
        if (timings) {
            timings->upload_ms = upload_ms;
            timings->compute_ms = ms_since(t1);
            timings->kernel_ms = kms;
            timings->has_kernel_ms = timed ? 1 : 0;
            timings->bands = bands;
        }
        *out = imgs[0];
        if (out2) *out2 = imgs[1];
        return FRH_OK;
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}
