// Batched develop: many frames of any size, Bayer phase, bit depth, white balance and colour
// matrix in one persistent launch per demosaic kind and Bayer phase (the phase is a template
// parameter of the walkers, as in the single-frame kernels).
//
// Each frame is split exactly as its single-frame launch splits it (the row walkers' strips
// plus the bilinear border pixels, the same kernels' device code), and every strip or group of
// 256 border pixels becomes an item of one work list. A persistent grid, sized to the resident
// waves the kernel's occupancy allows, takes items from an atomic counter until the list is
// empty. An item writes only its own pixels, so a frame's bytes are the same alone, in any
// batch and in any order. The counter is reset by a one-thread kernel ahead of the list kernels
// (no memset node: the launch sequence can be captured in a HIP graph and replayed).
//
// The per-frame strip height is the batch's tile shape; frh_batch_autotune times candidates on
// the batch itself and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ahd_walk.h"
#include "bil_walk.h"
#include "frh_internal.h"

namespace frh {
namespace {

constexpr uint32_t BORDER_ITEM = 256;  // border pixels per item (4 per lane)

struct BatchFrame {
    DevArgs a;
    WalkPlan wp;   // AHD frames
    BWalkPlan bp;  // bilinear frames
    int32_t rp;    // red position of the frame's Bayer phase
    uint32_t walk_items, items;
};

__global__ void k_batch_reset(uint32_t *counters, uint32_t n) {
    if (threadIdx.x < n) counters[threadIdx.x] = 0u;
}

// Work list of one demosaic kind: frames `ids` in order, item_off[k] the first item of ids[k]
// (item_off[n] = total).
struct BatchList {
    uint32_t n, total;
    uint32_t *counter;
    const uint8_t *lut8;
    const float *wtab;
};
// The descriptor and list pointers are kernel parameters of their own, read-only and unaliased:
// a frame's arguments are then fetched with scalar loads, as a single-frame launch's are.

__device__ __forceinline__ uint32_t next_item(uint32_t *counter) {
    uint32_t v = 0;
    if ((threadIdx.x & 63) == 0) v = atomicAdd(counter, 1u);
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint32_t frame_of(const uint32_t *__restrict__ item_off, uint32_t n, uint32_t item) {
    uint32_t lo = 0, hi = n;  // item_off[lo] <= item < item_off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (item_off[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void load_tables(const BatchList &L, Lds &s, float *wt) {
    const uint32_t *l = reinterpret_cast<const uint32_t *>(L.lut8);
    for (int i = threadIdx.x; i < 1025; i += blockDim.x) s.lut[i] = l[i];
    if (wt)
        for (int i = threadIdx.x; i < 73 * 73; i += blockDim.x) wt[i] = L.wtab[i];
    __syncthreads();
}

// Frames of a batch use the scalar calibration (at most a 2x2 black pattern, no deltas), so the
// shared Lds holds only the sRGB table; its black/scale arrays are never read.
template <int RP, bool FAST, bool TAME>
__global__ __launch_bounds__(WALK_WAVES * 64) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_batch_ahd(BatchList L, const BatchFrame *__restrict__ frames,
                                                                    const uint32_t *__restrict__ ids,
                                                                    const uint32_t *__restrict__ item_off) {
    __shared__ Lds s;
    __shared__ float wt[73 * 73];
    load_tables(L, s, wt);
    for (;;) {
        const uint32_t item = next_item(L.counter);
        if (item >= L.total) break;
        const uint32_t k = __builtin_amdgcn_readfirstlane(frame_of(item_off, L.n, item));
        const BatchFrame &F = frames[__builtin_amdgcn_readfirstlane(ids[k])];
        const uint32_t local = item - item_off[k];
        if (local < F.walk_items) {
            walk_item<RP, FAST, TAME>(F.a, F.wp, s, wt, (int)local);
        } else {
            const uint32_t base = (local - F.walk_items) * BORDER_ITEM + (threadIdx.x & 63);
            for (uint32_t j = 0; j < BORDER_ITEM; j += 64)
                if (base + j < F.wp.border.count) border_pixel<false>(F.a, s, F.wp.border, base + j);
        }
    }
}

template <int RP, bool FAST, bool TAME>
__global__ __launch_bounds__(BWALK_WAVES * 64) void k_batch_bilinear(BatchList L, const BatchFrame *__restrict__ frames,
                                                                    const uint32_t *__restrict__ ids,
                                                                    const uint32_t *__restrict__ item_off) {
    __shared__ Lds s;
    load_tables(L, s, nullptr);
    for (;;) {
        const uint32_t item = next_item(L.counter);
        if (item >= L.total) break;
        const uint32_t k = __builtin_amdgcn_readfirstlane(frame_of(item_off, L.n, item));
        const BatchFrame &F = frames[__builtin_amdgcn_readfirstlane(ids[k])];
        const uint32_t local = item - item_off[k];
        if (local < F.walk_items) {
            bwalk_item<RP, FAST, TAME>(F.a, F.bp, s, (int)local);
        } else {
            const uint32_t base = (local - F.walk_items) * BORDER_ITEM + (threadIdx.x & 63);
            for (uint32_t j = 0; j < BORDER_ITEM; j += 64)
                if (base + j < F.bp.border.count) border_pixel<false>(F.a, s, F.bp.border, base + j);
        }
    }
}

}  // namespace
}  // namespace frh

using namespace frh;

constexpr int NLIST = 8;  // (AHD, bilinear) x four Bayer phases

// The work lists of a set of frames: per list its frame count, its item total, the word offset of
// its frame ids in the list words (its item offsets follow them) and whether every frame of it
// takes the RGBA8 fast path.
struct Lists {
    uint32_t n[NLIST] = {}, total[NLIST] = {};
    size_t base[NLIST] = {};
    bool fast[NLIST] = {};
    bool tame[NLIST] = {};  // every AHD frame of the list is AHD-tame (DevArgs::ahd_tame)
    uint64_t items(bool ahd) const {
        uint64_t t = 0;
        for (int l = ahd ? 0 : 4; l < (ahd ? 4 : 8); l++) t += total[l];
        return t;
    }
};

// The device memory the list kernels read: descriptors, list words and one counter per list.
struct BatchDev {
    const BatchFrame *frames;
    const uint32_t *lists;
    uint32_t *counters;
};

// Side streams and events of a batch: its lists run on up to four streams at once, so that one
// persistent kernel's last waves overlap the next kernel's first ones (each kernel alone leaves
// the GPU part idle while its last strips finish). Created on the batch's first launch.
struct ListFork {
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    bool ready = false;
    hipError_t create() {
        hipError_t e = hipSuccess;
        for (int k = 0; k < 3 && e == hipSuccess; k++) e = hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
        for (int k = 0; k < 3 && e == hipSuccess; k++) e = hipEventCreateWithFlags(&join[k], hipEventDisableTiming);
        ready = e == hipSuccess;
        return e;
    }
    ~ListFork() {
        for (int k = 0; k < 3; k++) {
            if (side[k]) (void)hipStreamSynchronize(side[k]);
            if (side[k]) (void)hipStreamDestroy(side[k]);
            if (join[k]) (void)hipEventDestroy(join[k]);
        }
        if (fork) (void)hipEventDestroy(fork);
    }
};

struct frh_batch {
    frh_context *ctx = nullptr;
    std::vector<Plan> plans;
    std::vector<BatchFrame> frames;
    uint32_t ahd_rows = 0, bil_rows = 0;  // 0: each frame's own default
    // device copies
    BatchFrame *d_frames = nullptr;
    uint32_t *d_lists = nullptr;  // per list: ids, then offsets
    uint32_t *d_counters = nullptr;
    size_t n_list_words = 0;
    Lists lists;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    ListFork fork;
    uint32_t capacity = 0;        // frames the device and staging buffers hold
    uint8_t *h_stage = nullptr;   // pinned staging of descriptors and lists
    // recorded after each frh_batch_launch: rebuilding the lists waits for it, so a launch still
    // reading the descriptors is never overwritten (frh_batch_set_strip_rows, frh_batch_autotune)
    hipEvent_t last = nullptr;
    ~frh_batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_def) (void)hipGraphDestroy(graph_def);
        if (last) (void)hipEventSynchronize(last);
        if (last) (void)hipEventDestroy(last);
        if (d_frames) (void)hipFree(d_frames);
        if (d_lists) (void)hipFree(d_lists);
        if (d_counters) (void)hipFree(d_counters);
        if (h_stage) (void)hipHostFree(h_stage);
    }
    uint64_t items(bool ahd) const { return lists.items(ahd); }
    BatchDev dev() const { return BatchDev{d_frames, d_lists, d_counters}; }
};

// Device and pinned staging buffers for `capacity` frames (allocated once; a batch object is
// refilled many times by the serving loop without allocating).
static int32_t batch_reserve(frh_batch *b, uint32_t capacity) {
    if (capacity <= b->capacity) return FRH_OK;
    hipError_t e;
    if (b->d_frames) (void)hipFree(b->d_frames);
    if (b->d_lists) (void)hipFree(b->d_lists);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    b->d_frames = nullptr;
    b->d_lists = nullptr;
    b->h_stage = nullptr;
    b->capacity = 0;
    // per list: its frame ids and its item offsets (one more than its frames)
    const size_t words = 2 * (size_t)capacity + NLIST;
    if ((e = hipMalloc(&b->d_frames, capacity * sizeof(BatchFrame))) != hipSuccess ||
        (e = hipMalloc(&b->d_lists, words * 4)) != hipSuccess ||
        (e = hipHostMalloc(reinterpret_cast<void **>(&b->h_stage), capacity * sizeof(BatchFrame) + words * 4,
                           hipHostMallocDefault)) != hipSuccess)
        return hip_fail(e, "batch buffers");
    if (!b->d_counters && (e = hipMalloc(&b->d_counters, NLIST * sizeof(uint32_t))) != hipSuccess)
        return hip_fail(e, "batch counters");
    b->capacity = capacity;
    b->n_list_words = words;
    return FRH_OK;
}

// The descriptors and work lists of `plans` at strip heights ahd_rows / bil_rows (0: each frame's
// own): `frames` gets one descriptor per plan, `words` every list's frame ids then item offsets.
static void build_lists(const std::vector<Plan> &plans, uint32_t ahd_rows, uint32_t bil_rows,
                        std::vector<BatchFrame> &frames, std::vector<uint32_t> &words, Lists &L) {
    const size_t n = plans.size();
    frames.assign(n, BatchFrame{});
    std::vector<uint32_t> ids[NLIST], off[NLIST];
    for (int l = 0; l < NLIST; l++) {
        off[l].push_back(0);
        L.fast[l] = true;
        L.tame[l] = true;
    }
    for (size_t i = 0; i < n; i++) {
        const Plan &p = plans[i];
        BatchFrame &F = frames[i];
        F.a = p.a;
        F.a.band_y0 = 0;
        F.a.band_y1 = p.a.dev_h;
        F.rp = red_position(F.a);
        const int l = (p.ahd ? 0 : 4) + F.rp;
        if (p.ahd) {
            F.wp = make_walk_plan(F.a, (int)(ahd_rows ? ahd_rows : p.a.ahd_rows));
            F.walk_items = (uint32_t)(F.wp.strips * F.wp.chunks);
            F.items = F.walk_items + (F.wp.border.count + BORDER_ITEM - 1) / BORDER_ITEM;
        } else {
            F.bp = make_bwalk_plan(F.a, (int)(bil_rows ? bil_rows : p.a.bil_rows));
            F.walk_items = (uint32_t)(F.bp.strips * F.bp.chunks);
            F.items = F.walk_items + (F.bp.border.count + BORDER_ITEM - 1) / BORDER_ITEM;
        }
        if (F.items == 0) continue;
        L.fast[l] = L.fast[l] && fast_rgba8(F.a);
        L.tame[l] = L.tame[l] && F.a.ahd_tame != 0;
        ids[l].push_back((uint32_t)i);
        off[l].push_back(off[l].back() + F.items);
    }
    words.clear();
    for (int l = 0; l < NLIST; l++) {
        L.n[l] = (uint32_t)ids[l].size();
        L.total[l] = off[l].back();
        L.base[l] = words.size();
        words.insert(words.end(), ids[l].begin(), ids[l].end());
        words.insert(words.end(), off[l].begin(), off[l].end());
    }
}

// Build the work lists for the current strip heights and upload them on `st` (and wait for the
// upload when `sync`). The upload waits for the batch's last launch, which may still be reading
// the previous lists.
static int32_t batch_build(frh_batch *b, hipStream_t st, bool sync) {
    const size_t n = b->plans.size();
    int32_t rc = batch_reserve(b, (uint32_t)n);
    if (rc) return rc;
    std::vector<uint32_t> words;
    build_lists(b->plans, b->ahd_rows, b->bil_rows, b->frames, words, b->lists);
    if (words.size() > b->n_list_words) return fail(FRH_ERROR_INTERNAL, "batch lists exceed their buffer");
    hipError_t e;
    if (b->last && (e = hipStreamWaitEvent(st, b->last, 0)) != hipSuccess) return hip_fail(e, "batch upload");
    std::memcpy(b->h_stage, b->frames.data(), n * sizeof(BatchFrame));
    std::memcpy(b->h_stage + b->capacity * sizeof(BatchFrame), words.data(), words.size() * 4);
    if ((e = hipMemcpyAsync(b->d_frames, b->h_stage, n * sizeof(BatchFrame), hipMemcpyHostToDevice, st)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(b->d_lists, b->h_stage + b->capacity * sizeof(BatchFrame), words.size() * 4,
                            hipMemcpyHostToDevice, st)) != hipSuccess ||
        (sync && (e = hipStreamSynchronize(st)) != hipSuccess))
        return hip_fail(e, "batch upload");
    if (b->graph) {  // the lists changed: a captured graph is stale
        (void)hipGraphExecDestroy(b->graph);
        (void)hipGraphDestroy(b->graph_def);
        b->graph = nullptr;
        b->graph_def = nullptr;
    }
    return FRH_OK;
}

template <typename K>
static uint32_t resident_blocks(K kernel, int threads, int num_cus) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu * (uint32_t)num_cus;
}

template <int RP, bool FAST, bool TAME>
static void launch_list(const frh_context *ctx, const Lists &Ls, const BatchDev &D, int l, hipStream_t st) {
    BatchList L{};
    L.lut8 = ctx->d_lut8;
    L.wtab = ctx->d_wtab;
    L.n = Ls.n[l];
    L.total = Ls.total[l];
    L.counter = D.counters + l;
    const uint32_t *ids = D.lists + Ls.base[l], *off = ids + Ls.n[l];
    const uint32_t blocks = (Ls.total[l] + 3) / 4;
    if (l < 4) {
        const uint32_t g =
            std::min(resident_blocks(k_batch_ahd<RP, FAST, TAME>, WALK_WAVES * 64, ctx->num_cus), blocks);
        hipLaunchKernelGGL((k_batch_ahd<RP, FAST, TAME>), dim3(g), dim3(WALK_WAVES * 64), 0, st, L, D.frames, ids,
                           off);
    } else {
        const uint32_t g =
            std::min(resident_blocks(k_batch_bilinear<RP, FAST, TAME>, BWALK_WAVES * 64, ctx->num_cus), blocks);
        hipLaunchKernelGGL((k_batch_bilinear<RP, FAST, TAME>), dim3(g), dim3(BWALK_WAVES * 64), 0, st, L, D.frames,
                           ids, off);
    }
}

template <int RP>
static void launch_list_rp(const frh_context *ctx, const Lists &Ls, const BatchDev &D, int l, hipStream_t st) {
    // tame lists: every frame's values are bounded as ahd_tame in api.cpp states (the AHD walker's
    // square roots and clamps, and both walkers' finish with the matrix pre-scaled by 4096)
    const bool tame = Ls.tame[l];
    if (Ls.fast[l] && tame)
        launch_list<RP, true, true>(ctx, Ls, D, l, st);
    else if (Ls.fast[l])
        launch_list<RP, true, false>(ctx, Ls, D, l, st);
    else if (tame)
        launch_list<RP, false, true>(ctx, Ls, D, l, st);
    else
        launch_list<RP, false, false>(ctx, Ls, D, l, st);
}

// The launch sequence: counter reset, then each non-empty list (AHD lists first). With `fk`, the
// lists go round-robin to `st` and the three side streams, forked after the reset and joined
// back into `st` (a capture of `st` records them as parallel graph nodes); without, all on `st`.
static hipError_t enqueue_lists(const frh_context *ctx, const Lists &Ls, const BatchDev &D, hipStream_t st,
                                const ListFork *fk = nullptr) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, D.counters, (uint32_t)NLIST);
    hipError_t e = hipSuccess;
    if (fk && (e = hipEventRecord(fk->fork, st)) != hipSuccess) return e;
    int used = 0;
    for (int l = 0; l < NLIST; l++) {
        if (!Ls.total[l]) continue;
        const int k = fk ? used % 4 : 0;
        hipStream_t s = k ? fk->side[k - 1] : st;
        if (k && used < 4 && (e = hipStreamWaitEvent(s, fk->fork, 0)) != hipSuccess) return e;
        switch (l & 3) {
        case 0: launch_list_rp<0>(ctx, Ls, D, l, s); break;
        case 1: launch_list_rp<1>(ctx, Ls, D, l, s); break;
        case 2: launch_list_rp<2>(ctx, Ls, D, l, s); break;
        default: launch_list_rp<3>(ctx, Ls, D, l, s); break;
        }
        used++;
    }
    if (fk)
        for (int k = 1; k < 4 && k < used; k++) {
            if ((e = hipEventRecord(fk->join[k - 1], fk->side[k - 1])) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(st, fk->join[k - 1], 0)) != hipSuccess) return e;
        }
    return hipGetLastError();
}

static hipError_t batch_enqueue(frh_batch *b, hipStream_t st) {
    if (!b->fork.ready) {
        const hipError_t e = b->fork.create();
        if (e != hipSuccess) return e;
    }
    return enqueue_lists(b->ctx, b->lists, b->dev(), st, &b->fork);
}

extern "C" int32_t frh_batch_create(frh_context *ctx, uint32_t n, const frh_frame *frames,
                                    const frh_frame_info *infos, const frh_develop_spec *specs, void *const *dsts,
                                    frh_batch **out) {
    try {
        if (!ctx || !out || (n && (!frames || !infos || !specs || !dsts)))
            return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        if (n == 0) return fail(FRH_ERROR_INVALID_ARGUMENT, "empty batch");
        std::unique_ptr<frh_batch> b(new frh_batch);
        b->ctx = ctx;
        b->plans.resize(n);
        for (uint32_t i = 0; i < n; i++) {
            Plan &p = b->plans[i];
            int32_t rc = plan_develop(ctx, &frames[i], &infos[i], &specs[i], p);
            if (rc) return rc;
            if (!frames[i].data_on_device) return fail(FRH_ERROR_INVALID_ARGUMENT, "batched frames must be device-resident");
            if (!p.a.simple_cal)
                return fail(FRH_ERROR_NOT_AVAILABLE,
                            "batched frames need a black level of at most a 2x2 pattern without row/column deltas");
            if (p.a.edge_aware) return fail(FRH_ERROR_NOT_AVAILABLE, "batched frames use bilinear or AHD");
            if (p.a.generic || p.in.kind != FRH_INPUT_U16 || p.in.staged)
                return fail(FRH_ERROR_NOT_AVAILABLE, "batched frames are full-size u16 Bayer mosaics");
            if (specs[i].curve || specs[i].tone)
                return fail(FRH_ERROR_NOT_AVAILABLE, "batched frames use the sRGB curve without tone");
            if (specs[i].output2 != OUT_NONE) return fail(FRH_ERROR_NOT_AVAILABLE, "batched frames have one output");
            if (!dsts[i]) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL output");
            p.a.dst = dsts[i];
            p.a.dst2 = nullptr;
        }
        (void)hipSetDevice(ctx->device);
        int32_t rc = batch_build(b.get(), ctx->stream, true);
        if (rc) return rc;
        *out = b.release();
        return FRH_OK;
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_batch_launch(frh_batch *b, void *stream) {
    if (!b) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL batch");
    (void)hipSetDevice(b->ctx->device);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (b->graph) {
        if ((e = hipGraphLaunch(b->graph, st)) != hipSuccess) return hip_fail(e, "graph launch");
    } else if ((e = batch_enqueue(b, st)) != hipSuccess) {
        return hip_fail(e, "batch launch");
    }
    // remember the launch (not while the caller captures `st` into a graph of its own)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
        if (!b->last && (e = hipEventCreateWithFlags(&b->last, hipEventDisableTiming)) != hipSuccess)
            return hip_fail(e, "batch event");
        if ((e = hipEventRecord(b->last, st)) != hipSuccess) return hip_fail(e, "batch event");
    }
    return FRH_OK;
}

extern "C" int32_t frh_batch_capture(frh_batch *b, void *stream) {
    if (!b) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL batch");
    if (!stream) return fail(FRH_ERROR_INVALID_ARGUMENT, "capture needs a stream other than the null stream");
    (void)hipSetDevice(b->ctx->device);
    hipStream_t st = (hipStream_t)stream;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        (void)hipGraphDestroy(b->graph_def);
        b->graph = nullptr;
        b->graph_def = nullptr;
    }
    hipError_t e;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "begin capture");
    const hipError_t le = batch_enqueue(b, st);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (le != hipSuccess || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(le != hipSuccess ? le : e, "capture");
    }
    hipGraphExec_t x = nullptr;
    if ((e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0)) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "graph instantiate");
    }
    b->graph = x;
    b->graph_def = g;
    return FRH_OK;
}

extern "C" int32_t frh_batch_autotune(frh_batch *b, void *stream, int32_t iters, uint32_t *best_ahd,
                                      uint32_t *best_bil) {
    try {
        if (!b) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL batch");
        if (iters < 1) iters = 1;
        (void)hipSetDevice(b->ctx->device);
        hipStream_t st = (hipStream_t)stream;
        const bool had_graph = b->graph != nullptr;
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            return fail(FRH_ERROR_GPU_UNAVAILABLE, "hipEventCreate failed");
        int32_t rc = FRH_OK;
        auto time_it = [&](float *ms) -> int32_t {
            int32_t r = batch_build(b, b->ctx->stream, true);
            if (r) return r;
            hipError_t e = batch_enqueue(b, st);  // warm
            (void)hipEventRecord(e0, st);
            for (int i = 0; i < iters && e == hipSuccess; i++) e = batch_enqueue(b, st);
            (void)hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e != hipSuccess) return hip_fail(e, "autotune launch");
            (void)hipEventElapsedTime(ms, e0, e1);
            return FRH_OK;
        };
        // strip heights are tuned per demosaic kind, AHD first, each with the other at its best
        uint32_t best_a = 64, best_b = 32;
        const uint32_t cand_a[] = {24, 32, 40, 48, 56, 64, 72, 80, 96, 112, 128};
        const uint32_t cand_b[] = {8, 12, 16, 24, 32, 48, 64};
        b->bil_rows = best_b;
        if (b->items(true)) {
            float best = 1e30f;
            for (uint32_t r : cand_a) {
                b->ahd_rows = r;
                float ms = 0.0f;
                if ((rc = time_it(&ms)) != FRH_OK) break;
                if (ms < best) best = ms, best_a = r;
            }
        }
        if (rc == FRH_OK && b->items(false)) {
            b->ahd_rows = best_a;
            float best = 1e30f;
            for (uint32_t r : cand_b) {
                b->bil_rows = r;
                float ms = 0.0f;
                if ((rc = time_it(&ms)) != FRH_OK) break;
                if (ms < best) best = ms, best_b = r;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc != FRH_OK) return rc;
        b->ahd_rows = best_a;
        b->bil_rows = best_b;
        if ((rc = batch_build(b, b->ctx->stream, true)) != FRH_OK) return rc;
        if (best_ahd) *best_ahd = best_a;
        if (best_bil) *best_bil = best_b;
        if (had_graph && stream) return frh_batch_capture(b, stream);
        return FRH_OK;
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_batch_set_strip_rows(frh_batch *b, uint32_t ahd_rows, uint32_t bil_rows) {
    if (!b) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL batch");
    b->ahd_rows = ahd_rows;
    b->bil_rows = bil_rows;
    (void)hipSetDevice(b->ctx->device);
    return batch_build(b, b->ctx->stream, true);
}

extern "C" uint64_t frh_batch_items(const frh_batch *b) { return b ? b->items(true) + b->items(false) : 0; }

extern "C" void frh_batch_destroy(frh_batch *b) {
    if (b) (void)hipSetDevice(b->ctx->device);
    delete b;
}

// ---- internal interface of the serving loop (frh_internal.h) ----

frh_batch *frh::batch_new(frh_context *ctx, uint32_t capacity) {
    std::unique_ptr<frh_batch> b(new frh_batch);
    b->ctx = ctx;
    if (batch_reserve(b.get(), capacity) != FRH_OK) return nullptr;
    return b.release();
}

int32_t frh::batch_assign(frh_batch *b, std::vector<Plan> &&plans, hipStream_t st) {
    b->plans = std::move(plans);
    return batch_build(b, st, false);
}

// (The serving loop runs its batches on its one compute stream: its copy streams already share
// the process's hardware queues.)
hipError_t frh::batch_run(frh_batch *b, hipStream_t st) { return enqueue_lists(b->ctx, b->lists, b->dev(), st); }

// ---- the batched develop that allocates nothing ----
//
// frh_batch_develop_noalloc validates, lists and launches a batch without any device allocation:
// the descriptors, the work lists and the list counters live in `scratch`, a fixed array of this
// library in device memory (one copy per GPU, held until the process ends). The host fills it
// through the arguments of small copy launches on the caller's stream, so the upload is ordered
// with the launches that read it and allocates nothing either.

namespace {
constexpr uint32_t NOALLOC_MAX_FRAMES = 256;
constexpr size_t NA_FRAMES_BYTES = ((size_t)NOALLOC_MAX_FRAMES * sizeof(BatchFrame) + 255) & ~(size_t)255;
constexpr size_t NA_LIST_WORDS = 2 * (size_t)NOALLOC_MAX_FRAMES + NLIST;
constexpr size_t NA_LISTS_BYTES = (NA_LIST_WORDS * 4 + 255) & ~(size_t)255;
constexpr size_t NA_SCRATCH_BYTES = NA_FRAMES_BYTES + NA_LISTS_BYTES + 256;  // + the list counters

// Words one copy launch carries in its arguments (about 2 KB of the kernel argument segment).
constexpr uint32_t PUT_WORDS = 480;
struct PutWords {
    uint32_t *dst;
    uint32_t count;
    uint32_t w[PUT_WORDS];
};

__global__ __launch_bounds__(256) void k_put_words(PutWords p) {
    // read in place from the argument segment (indexing the by-value copy would move it to
    // private memory per lane)
    (void)p;
    const PutWords &q = *(const PutWords *)__builtin_amdgcn_kernarg_segment_ptr();
    for (uint32_t i = threadIdx.x; i < q.count; i += blockDim.x) q.dst[i] = q.w[i];
}

// Copy `count` words from host memory to device `dst` on `st` through launch arguments.
hipError_t put_words(uint32_t *dst, const uint32_t *src, size_t count, hipStream_t st) {
    PutWords p;
    for (size_t i = 0; i < count; i += PUT_WORDS) {
        p.dst = dst + i;
        p.count = (uint32_t)std::min<size_t>(PUT_WORDS, count - i);
        std::memcpy(p.w, src + i, p.count * 4);
        hipLaunchKernelGGL(k_put_words, dim3(1), dim3(256), 0, st, p);
    }
    return hipGetLastError();
}
}  // namespace

static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

// The device address of `scratch` on the current device (its host-side name is not one).
hipError_t frh::noalloc_scratch_address(void **ptr) {
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

extern "C" int32_t frh_batch_develop_noalloc(frh_context *ctx, uint32_t n, const frh_frame *frames,
                                             const frh_frame_info *infos, const frh_develop_spec *specs,
                                             void *const *dsts, uint32_t ahd_rows, uint32_t bilinear_rows,
                                             void *stream) {
    try {
        if (!ctx || (n && (!frames || !infos || !specs || !dsts)))
            return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        if (n == 0) return fail(FRH_ERROR_INVALID_ARGUMENT, "empty batch");
        if (n > NOALLOC_MAX_FRAMES)
            return fail(FRH_ERROR_LIMIT_EXCEEDED, "frh_batch_develop_noalloc takes at most " +
                                                      std::to_string(NOALLOC_MAX_FRAMES) + " frames per call");
        if (!ctx->noalloc_scratch) return fail(FRH_ERROR_GPU_UNAVAILABLE, "the device scratch array is not set up");
        std::vector<Plan> plans(n);
        for (uint32_t i = 0; i < n; i++) {
            const frh_frame &f = frames[i];
            const frh_frame_info &inf = infos[i];
            const frh_develop_spec &s = specs[i];
            if (f.size < sizeof(frh_frame) || inf.size < sizeof(frh_frame_info) ||
                s.size < offsetof(frh_develop_spec, cancel))
                return fail(FRH_ERROR_INVALID_ARGUMENT, "descriptor size field too small");
            // Refused before validation, which would copy their tables to the device: frames
            // other than device-resident u16 Bayer mosaics at full size, black-level deltas,
            // linearisation curves, tone, output curves, edge-aware filters, second outputs.
            bool bayer = inf.cfa_width == 2 && inf.cfa_height == 2;
            if (bayer) {
                int cnt[4] = {0, 0, 0, 0};
                for (int k = 0; k < 4; k++) cnt[std::min<int>(inf.cfa[k], 3)]++;
                bayer = cnt[0] == 1 && cnt[1] == 2 && cnt[2] == 1;
            }
            if (!f.data_on_device || f.input != FRH_INPUT_U16 || f.cpp > 1 || f.linearization_len || !bayer ||
                s.half_size)
                return fail(FRH_ERROR_NOT_AVAILABLE, "frh_batch_develop_noalloc takes device-resident u16 Bayer "
                                                     "mosaics at full size without a linearisation curve");
            if (inf.black_row_deltas_len || inf.black_col_deltas_len || s.tone || s.curve ||
                s.output2 != OUT_NONE || (s.edge && (s.edge->size < sizeof(frh_edge_spec) || !s.edge->ahd)))
                return fail(FRH_ERROR_NOT_AVAILABLE, "frh_batch_develop_noalloc takes bilinear or AHD develops "
                                                     "with one output, no black-level deltas, tone or curve");
            Plan &p = plans[i];
            int32_t rc = plan_develop(ctx, &f, &inf, &s, p);
            if (rc) return rc;
            if (!p.a.simple_cal || p.a.generic || p.a.edge_aware || p.in.staged)
                return fail(FRH_ERROR_NOT_AVAILABLE, "batched frames need a black level of at most a 2x2 pattern");
            if (!dsts[i]) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL output");
            p.a.dst = dsts[i];
            p.a.dst2 = nullptr;
        }
        std::vector<BatchFrame> bf;
        std::vector<uint32_t> words;
        Lists L;
        build_lists(plans, ahd_rows, bilinear_rows, bf, words, L);
        if (words.size() > NA_LIST_WORDS) return fail(FRH_ERROR_INTERNAL, "batch lists exceed the scratch array");
        uint8_t *base = static_cast<uint8_t *>(ctx->noalloc_scratch);
        BatchFrame *d_frames = reinterpret_cast<BatchFrame *>(base);
        uint32_t *d_lists = reinterpret_cast<uint32_t *>(base + NA_FRAMES_BYTES);
        uint32_t *d_counters = reinterpret_cast<uint32_t *>(base + NA_FRAMES_BYTES + NA_LISTS_BYTES);
        static_assert(sizeof(BatchFrame) % 4 == 0, "descriptors are copied as words");
        (void)hipSetDevice(ctx->device);
        hipStream_t st = (hipStream_t)stream;
        hipError_t e;
        if ((e = put_words(reinterpret_cast<uint32_t *>(d_frames), reinterpret_cast<const uint32_t *>(bf.data()),
                           bf.size() * sizeof(BatchFrame) / 4, st)) != hipSuccess ||
            (e = put_words(d_lists, words.data(), words.size(), st)) != hipSuccess)
            return hip_fail(e, "descriptor upload");
        if ((e = enqueue_lists(ctx, L, BatchDev{d_frames, d_lists, d_counters}, st)) != hipSuccess)
            return hip_fail(e, "batch launch");
        return FRH_OK;
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}
