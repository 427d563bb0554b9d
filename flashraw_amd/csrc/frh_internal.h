// Host-side internals shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "flashraw_hip.h"
#include "frh_device.h"

namespace frh {

// Records `msg` as the calling thread's last error and returns `code`.
int32_t fail(int32_t code, const std::string &msg);
// A HIP error as a status (out of memory, or the device failed).
int32_t hip_fail(hipError_t e, const char *what);
// The 4097-entry sRGB 8-bit table (process.rs:1903-1915).
void build_srgb8(uint8_t out[4097]);

// Whether `t` (NULL: never) has been cancelled; a poll answering non-zero cancels it for good.
bool cancelled(const frh_cancel_token *t);
// FRH_ERROR_CANCELLED with its message.
int32_t cancel_fail();
// The token of a develop spec, NULL when its `size` ends before the field.
const frh_cancel_token *spec_cancel(const frh_develop_spec *s);

// Bytes of freed device buffers a context keeps for reuse. The reference keeps 1 GiB
// (POOL_MAX_BYTES, crates/flashraw-gpu/src/context.rs:153); a 24 MP noise reduction alone holds
// about 2 GB of planes and partial sums at once, and an MI355X has 288 GB, so the pool keeps 8 GiB
// and repeated runs reuse their buffers instead of freeing and allocating them again.
constexpr uint64_t POOL_MAX_BYTES = 8ull << 30;

// The per-frame table cache (calibration, CFA sites, black-level deltas, linearisation and tone
// tables, edge-aware tap lists) keeps at most this many bytes and entries. Beyond either, the
// least recently used entries that no prepared plan holds are freed, after a device
// synchronisation so that no queued launch still reads them.
constexpr uint64_t TABLE_CACHE_BYTES = 64ull << 20;
constexpr uint64_t TABLE_CACHE_ENTRIES = 4096;

// A device copy of a table, found by (kind, hash of its key bytes, key length) and confirmed by
// comparing the key bytes kept on the host. `pins` counts the prepared plans that hold it.
struct TableEntry {
    void *d;
    uint64_t dev_bytes;
    std::vector<uint8_t> key;
    uint64_t last_use;
    uint32_t pins;
};

}  // namespace frh

struct frh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::string name;
    uint64_t total_mem = 0;
    uint8_t *d_lut8 = nullptr;
    float *d_srgb16 = nullptr;
    float *d_wtab = nullptr;  // 73 x 73 AHD blend weights
    uint64_t max_binding = 0;
    uint32_t ahd_rows = 64;  // rows per AHD walker strip
    uint32_t bil_rows = 32;  // rows per bilinear walker strip
    // strip heights found by frh_plan_autotune, by (AHD, developed width, rows, fast path)
    // autotuned walker tile shapes per geometry: strip height, waves per block, columns per lane
    struct TileShape {
        uint32_t rows, wpb, cpl;
    };
    std::map<std::tuple<int, uint32_t, uint32_t, int>, TileShape> tuned;
    // device copies of small per-frame tables (calibration, CFA sites, black-level deltas,
    // linearisation curves, tone and curve tables: kind 0), of edge-aware tap lists (kind 1) and
    // of the Kronecker denoiser's operand tables (kind 2), by (kind, hash of the key bytes, key
    // length); bounded by TABLE_CACHE_BYTES / _ENTRIES
    std::multimap<std::tuple<int, uint64_t, uint64_t>, frh::TableEntry> tables;
    uint64_t table_bytes = 0, table_tick = 0;
    std::mutex mu;
    struct Block {
        void *ptr;
        uint64_t size;
    };
    std::vector<Block> pool;
    uint64_t pooled = 0;
    // pinned host staging for device read-backs (the noise estimate's blocks), grown on demand
    std::mutex stage_mu;
    void *h_stage = nullptr;
    uint64_t h_stage_bytes = 0;
    // device address of the library's fixed scratch array on this context's device (taken once
    // at creation; the array's host-side name is not a device address)
    void *noalloc_scratch = nullptr;

    void *alloc(uint64_t bytes, hipError_t *err);
    void release(void *p, uint64_t bytes);
    ~frh_context();
};

namespace frh {

// A validated develop: kernel arguments (without the buffers a launch names) and the outputs'
// geometry (run.rs:439-760).
// A validated input (GpuInput + window, run.rs:456-490): what to upload and how to read it.
struct InputDesc {
    int32_t kind;          // FRH_INPUT_*
    uint32_t bits;         // packed
    int32_t order;         // packed
    uint64_t stride;       // packed: bytes per raster row; u16 / f32: samples per raster row
    uint64_t row_bytes;    // packed: bytes of sample data per raster row
    uint64_t len;          // packed: payload bytes; u16 / f32: samples
    uint32_t w, h, cpp;    // window
    bool lut;              // a linearisation curve applies (integer input with a non-empty table)
    bool staged;           // the develop reads u16 samples unpacked / linearised first
    // bytes of the raster rows the window needs (uploaded for host input), from `first_byte`
    uint64_t first_byte, upload_bytes;
};

// Validate the frame's input description (status code; the message is the thread's last error).
int32_t check_input(const frh_frame *f, InputDesc &in);

// Calibration tables of every component: black[cpp][pw * ph] then scale[cpp][pw * ph].
int32_t build_calib(const frh_frame_info *info, uint32_t cpp, std::vector<float> &gcal, uint32_t &pw, uint32_t &ph);

// A device copy of `bytes` bytes at `host`, cached by the context (NULL with *rc set on failure).
// The copy stays valid for launches queued before the next call that evicts it (which
// synchronises the device first); a prepared plan pins the tables it names.
const void *device_table(frh_context *ctx, const void *host, uint64_t bytes, int32_t *rc);
// The cached device copy for (kind, key), built by `make` (which fills the device bytes to upload)
// on a miss.
const void *cached_table(frh_context *ctx, int kind, const void *key, uint64_t key_bytes,
                         const std::function<bool(std::vector<uint8_t> &)> &make, int32_t *rc);
// Pin (+1) or unpin (-1) the cached tables a develop's arguments name.
void pin_tables(frh_context *ctx, const DevArgs &a, int delta);

struct Plan {
    DevArgs a;
    bool ahd;
    uint32_t out_w[2], out_h[2], ch[2];
    uint64_t bytes[2];
    int32_t fmt[2];
    uint64_t in_row_bytes;
    InputDesc in;
};

// Output encodings: channels, bytes per sample and image format of `mode`; whether it is one;
// whether it is a display output (oriented, toned).
void out_format(int32_t mode, uint32_t *ch, uint32_t *bps, int32_t *fmt);
bool valid_mode(int32_t m);
bool display_mode(int32_t m);
// The display outputs' tone and output curve into a's table pointers (device copies cached by
// the context).
int32_t output_tables(frh_context *ctx, const frh_tone *tone, const frh_curve *curve, DevArgs &a);
// Validate a develop request and fill `p` (status code; the message is the thread's last error).
int32_t plan_develop(const frh_context *ctx, const frh_frame *f, const frh_frame_info *info,
                     const frh_develop_spec *spec, Plan &p);

// Batches reused by the serving loop: reserve buffers for `capacity` frames once, then assign a
// new set of plans (descriptors uploaded on `st`, asynchronously) and run on `st`.
frh_batch *batch_new(frh_context *ctx, uint32_t capacity);
int32_t batch_assign(frh_batch *b, std::vector<Plan> &&plans, hipStream_t st);
hipError_t batch_run(frh_batch *b, hipStream_t st);
// The device address of the no-allocation batch's scratch array on the current device.
hipError_t noalloc_scratch_address(void **ptr);

}  // namespace frh

struct frh_image {
    frh_context *ctx;
    void *dptr;
    uint64_t bytes;
    uint32_t w, h, ch;
    int32_t fmt;
};
