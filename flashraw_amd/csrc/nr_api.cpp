// Luma / chroma noise reduction: the host pipeline of GpuContext::denoise
// (crates/flashraw-gpu/src/denoise.rs:817-1176) over the kernels of nr.hip, and the noise models
// (flashraw::denoise::gpu_models and Model, crates/flashraw/src/denoise/mod.rs:196-222, 705-869)
// computed on the host from the blocks measured on the device (the run's one read-back).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <list>
#include <thread>
#include <vector>

#include "flashraw_hip.h"
#include "frh_internal.h"
#include "nr.h"

using namespace frh;

namespace {

constexpr float CHI2_MEDIAN = 0.4549364f;  // median of chi-square with one degree of freedom
constexpr float SELECTION_BIAS = 1.10f;
constexpr float LEVEL_CAP = 1.5f;
constexpr uint32_t MIN_LEVEL_SIDE = 48;

// f32::total_cmp as an integer key. The map is an involution: applied to a key it gives back the
// float's bits, so order statistics of the keys are those of the floats in total order.
inline int32_t total_key(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}
inline float key_float(int32_t k) {
    const int32_t i = k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
    float f;
    memcpy(&f, &i, 4);
    return f;
}

// Runs fn(0) .. fn(n - 1) on up to 16 host threads (tasks taken from a shared counter).
template <class Fn>
void parallel_for(size_t n, Fn fn) {
    const size_t hw = std::max<size_t>(std::thread::hardware_concurrency(), 1);
    const size_t nt = std::min<size_t>({hw, (size_t)16, n});
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

// The finite blocks of one level in Model::from_stats's order (mod.rs:803-822): ascending luma
// mean in total order, ties in block order (the stable sort), cut into `bins` bins of `per`.
struct Binned {
    const float *stats = nullptr;
    std::vector<uint32_t> idx;
    std::vector<float> act;  // the summed squared AC coefficients of idx[i]
    size_t bins = 0, per = 0;
};

bool sort_blocks(const float *stats, size_t count, Binned &b) {
    b.stats = stats;
    // activities, eight blocks at a time (each block's sum in coefficient order)
    std::vector<float> act(count);
    size_t i0 = 0;
    for (; i0 + 8 <= count; i0 += 8) {
        float a[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const float *s = stats + i0 * 64;
        for (int k = 1; k < 64; k++)
            for (int j = 0; j < 8; j++) a[j] = a[j] + s[j * 64 + k];
        for (int j = 0; j < 8; j++) act[i0 + j] = a[j];
    }
    for (; i0 < count; i0++) {
        const float *s = stats + i0 * 64;
        float a = 0.0f;
        for (int k = 1; k < 64; k++) a = a + s[k];
        act[i0] = a;
    }
    std::vector<uint64_t> keyed;
    keyed.reserve(count);
    for (size_t i = 0; i < count; i++)
        if (std::isfinite(stats[i * 64]) && std::isfinite(act[i]))
            keyed.push_back(((uint64_t)((uint32_t)total_key(stats[i * 64]) ^ 0x80000000u) << 32) | (uint64_t)i);
    if (keyed.size() < 32) return false;
    std::sort(keyed.begin(), keyed.end());  // (mean key, index): the stable order by mean
    const size_t n = keyed.size();
    b.idx.resize(n);
    b.act.resize(n);
    for (size_t i = 0; i < n; i++) {
        const uint32_t j = (uint32_t)keyed[i];
        b.idx[i] = j;
        b.act[i] = act[j];
    }
    b.bins = std::min<size_t>(std::max<size_t>(n / 96, 1), NR_BINS);
    b.per = (n + b.bins - 1) / b.bins;
    return true;
}

// Bin `bin` (mod.rs:823-840): the blocks whose activity is at most the bin's median activity, the
// median of each squared coefficient over them, as a deviation; and the bin's median luma mean.
void bin_model(const Binned &b, size_t bin, float *mean, float sig[64]) {
    const size_t b0 = bin * b.per, len = std::min(b.per, b.idx.size() - b0);
    std::vector<int32_t> act(len);
    for (size_t i = 0; i < len; i++) act[i] = total_key(b.act[b0 + i]);
    std::nth_element(act.begin(), act.begin() + len / 2, act.end());
    const float cut = key_float(act[len / 2]);
    std::vector<uint32_t> flat;
    flat.reserve(len);
    for (size_t i = 0; i < len; i++)
        if (b.act[b0 + i] <= cut) flat.push_back(b.idx[b0 + i]);
    const size_t nf = flat.size();
    std::vector<int32_t> t(63 * nf);  // frequency-major keys of the flat blocks
    for (size_t i = 0; i < nf; i++) {
        const float *s = b.stats + (size_t)flat[i] * 64;
        for (int k = 1; k < 64; k++) t[(size_t)(k - 1) * nf + i] = total_key(s[k]);
    }
    sig[0] = 0.0f;
    for (int k = 1; k < 64; k++) {
        int32_t *v = t.data() + (size_t)(k - 1) * nf;
        std::nth_element(v, v + nf / 2, v + nf);
        sig[k] = std::sqrt(key_float(v[nf / 2]) / CHI2_MEDIAN * SELECTION_BIAS);
    }
    *mean = b.stats[(size_t)b.idx[b0 + len / 2] * 64];
}

// The model's deviations smoothed across neighbouring bins (mod.rs:841-847).
void smooth_bins(NrModel &md) {
    if (md.n <= 2) return;
    float orig[NR_BINS][64];
    memcpy(orig, md.s, sizeof(orig));
    for (uint32_t i = 0; i < md.n; i++) {
        const uint32_t a = i ? i - 1 : 0, b = std::min(i + 1, md.n - 1);
        for (int k = 0; k < 64; k++) md.s[i][k] = 0.25f * orig[a][k] + 0.5f * orig[i][k] + 0.25f * orig[b][k];
    }
}

// Model::from_stats (mod.rs:803-848) of several levels at once (the bins of every level in
// parallel); ok[i] false when level i has fewer than 32 finite blocks.
void models_from_stats(const std::vector<const float *> &stats, const std::vector<size_t> &counts,
                       std::vector<NrModel> &md, std::vector<char> &ok) {
    const size_t nl = stats.size();
    std::vector<Binned> binned(nl);
    md.assign(nl, NrModel{});
    ok.assign(nl, 0);
    parallel_for(nl, [&](size_t i) { ok[i] = sort_blocks(stats[i], counts[i], binned[i]) ? 1 : 0; });
    std::vector<std::pair<size_t, size_t>> tasks;
    for (size_t i = 0; i < nl; i++)
        if (ok[i]) {
            md[i].n = (uint32_t)((binned[i].idx.size() + binned[i].per - 1) / binned[i].per);
            for (size_t bin = 0; bin < md[i].n; bin++) tasks.push_back({i, bin});
        }
    parallel_for(tasks.size(), [&](size_t t) {
        const size_t i = tasks[t].first, bin = tasks[t].second;
        bin_model(binned[i], bin, &md[i].m[bin], md[i].s[bin]);
    });
    for (size_t i = 0; i < nl; i++)
        if (ok[i]) smooth_bins(md[i]);
}

// Model::from_stats (mod.rs:803-848): false when fewer than 32 blocks are finite.
bool from_stats(const float *stats, size_t count, NrModel &md) {
    std::vector<NrModel> m;
    std::vector<char> ok;
    models_from_stats({stats}, {count}, m, ok);
    md = m[0];
    return ok[0] != 0;
}

// Model::from_stats of every region of the device blocks `d_stats` (64 floats each), fitted on the
// device: the regions' models (ok[i] false below 32 finite blocks). False (nothing launched) when a
// region holds more blocks than the fit kernels keep per bin, or on a device error (*err).
bool device_models(frh_context *ctx, hipStream_t st, const float *d_stats, const std::vector<NrRegion> &reg,
                   std::vector<NrModel> &md, std::vector<char> &ok, hipError_t *err) {
    *err = hipSuccess;
    uint64_t total = 0;
    for (const NrRegion &r : reg) {
        if ((uint64_t)r.count > 12ull * NR_FIT_LIST) return false;
        total = std::max<uint64_t>(total, (uint64_t)r.first + r.count);
    }
    md.assign(reg.size(), NrModel{});
    ok.assign(reg.size(), 0);
    if (reg.empty() || total == 0) return true;
    const uint64_t fit_bytes = reg.size() * sizeof(NrFit), reg_bytes = reg.size() * sizeof(NrRegion);
    void *d_reg = ctx->alloc(reg_bytes, err);
    void *d_fit = d_reg ? ctx->alloc(fit_bytes, err) : nullptr;
    void *d_key = d_fit ? ctx->alloc(total * 4, err) : nullptr;
    void *d_act = d_key ? ctx->alloc(total * 4, err) : nullptr;
    void *d_flat = d_act ? ctx->alloc(total * 4, err) : nullptr;
    std::vector<NrFit> fits(reg.size());
    hipError_t e = d_flat ? hipSuccess : (*err != hipSuccess ? *err : hipErrorOutOfMemory);
    if (e == hipSuccess) e = hipMemcpyAsync(d_reg, reg.data(), reg_bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = nr_fit(d_stats, (uint32_t)total, static_cast<const NrRegion *>(d_reg), (uint32_t)reg.size(),
                   static_cast<uint32_t *>(d_key), static_cast<float *>(d_act), static_cast<uint32_t *>(d_flat),
                   static_cast<NrFit *>(d_fit), st);
    if (e == hipSuccess) e = hipMemcpyAsync(fits.data(), d_fit, fit_bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_flat) ctx->release(d_flat, total * 4);
    if (d_act) ctx->release(d_act, total * 4);
    if (d_key) ctx->release(d_key, total * 4);
    if (d_fit) ctx->release(d_fit, fit_bytes);
    if (d_reg) ctx->release(d_reg, reg_bytes);
    if (e != hipSuccess) {
        *err = e;
        return false;
    }
    for (size_t i = 0; i < reg.size(); i++) {
        const NrFit &f = fits[i];
        if (f.n < 32) continue;
        NrModel &m = md[i];
        m.n = f.nbins;
        for (uint32_t b = 0; b < f.nbins; b++) {
            m.m[b] = f.mean[b];
            m.s[b][0] = 0.0f;
            for (int k = 1; k < 64; k++) m.s[b][k] = std::sqrt(key_float(f.sig[b][k]) / CHI2_MEDIAN * SELECTION_BIAS);
        }
        smooth_bins(m);
        ok[i] = 1;
    }
    return true;
}

float var_at(const NrVar &v, float m) {
    uint32_t i = 0;
    while (i < v.n && v.m[i] < m) i++;
    if (i == 0) return v.v[0];
    if (i >= v.n) return v.v[v.n - 1];
    const float a = v.m[i - 1], b = v.m[i];
    const float t = b > a ? (m - a) / (b - a) : 0.0f;
    return v.v[i - 1] + (v.v[i] - v.v[i - 1]) * t;
}

// Model::next_level: a quarter of the mean variance of the 15 lowest AC frequencies per bin.
NrVar next_level(const NrModel &md) {
    NrVar v{};
    v.n = md.n;
    for (uint32_t b = 0; b < md.n; b++) {
        float t = 0.0f;
        for (int l = 0; l < 4; l++)
            for (int k = 0; k < 4; k++) t += md.s[b][l * 8 + k] * md.s[b][l * 8 + k];
        v.m[b] = md.m[b];
        v.v[b] = 0.25f * t / 15.0f;
    }
    return v;
}

// Model::cap: every deviation at most f * sqrt(predicted variance).
void cap(NrModel &md, const NrVar &pred, float f) {
    for (uint32_t b = 0; b < md.n; b++) {
        const float c = f * std::sqrt(std::max(var_at(pred, md.m[b]), 0.0f));
        for (int k = 0; k < 64; k++) md.s[b][k] = std::min(md.s[b][k], c);
    }
}

// Model::pixel_variance: the mean AC variance per bin.
NrVar pixel_variance(const NrModel &md) {
    NrVar v{};
    v.n = md.n;
    for (uint32_t b = 0; b < md.n; b++) {
        float t = 0.0f;
        for (int k = 1; k < 64; k++) t += md.s[b][k] * md.s[b][k];
        v.m[b] = md.m[b];
        v.v[b] = t / 63.0f;
    }
    return v;
}

// estimate_step (mod.rs:790-798, denoise.rs:223-233): 0 when the level is too small to measure.
uint32_t estimate_step(uint32_t w, uint32_t h, uint32_t max_blocks) {
    const uint64_t bx = w / 8, by = h / 8;
    if (bx * by < 32 || max_blocks == 0) return 0;
    const uint32_t s = (uint32_t)std::ceil(std::sqrt((double)(bx * by) / (double)max_blocks));
    return std::max(s, 1u);
}

uint32_t level_count(uint32_t w, uint32_t h, uint32_t max) {
    uint32_t n = 1;
    while (n < max && (w + 1) / 2 >= MIN_LEVEL_SIDE && (h + 1) / 2 >= MIN_LEVEL_SIDE) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
        n++;
    }
    return n;
}

// The orthonormal 8-point DCT-II matrix in f32 (Dct::new, mod.rs:633-641).
void dct_matrix(float m[64]) {
    for (int k = 0; k < 8; k++) {
        const double a = k == 0 ? std::sqrt(1.0 / 8.0) : std::sqrt(2.0 / 8.0);
        for (int i = 0; i < 8; i++) m[k * 8 + i] = (float)(a * std::cos(M_PI * (double)(2 * i + 1) * (double)k / 16.0));
    }
}

// Device planes of one run. A plane put back is reused by the run's later planes (the stream
// orders their uses); every buffer returns to the context's pool only after the stream has
// finished with it.
struct Arena {
    frh_context *ctx;
    hipStream_t st;
    std::vector<std::pair<void *, uint64_t>> held, spare;
    hipError_t err = hipSuccess;
    Arena(frh_context *c, hipStream_t s) : ctx(c), st(s) {}
    float *get(uint64_t floats) {
        const uint64_t bytes = std::max<uint64_t>(floats, 1) * 4;
        for (size_t i = 0; i < spare.size(); i++)
            if (spare[i].second >= bytes && spare[i].second <= 2 * bytes) {
                held.push_back(spare[i]);
                spare.erase(spare.begin() + i);
                return static_cast<float *>(held.back().first);
            }
        hipError_t e = hipSuccess;
        void *p = ctx->alloc(bytes, &e);
        if (!p) {
            err = e == hipSuccess ? hipErrorOutOfMemory : e;
            return nullptr;
        }
        held.push_back({p, bytes});
        return static_cast<float *>(p);
    }
    void put(float *p) {
        for (size_t i = 0; i < held.size(); i++)
            if (held[i].first == p) {
                spare.push_back(held[i]);
                held.erase(held.begin() + i);
                return;
            }
    }
    ~Arena() {
        (void)hipStreamSynchronize(st);
        for (auto &h : held) ctx->release(h.first, h.second);
        for (auto &h : spare) ctx->release(h.first, h.second);
    }
};

struct Level {
    float *p;
    uint32_t w, h;
};

struct Run {
    frh_context *ctx;
    hipStream_t st;
    // host copies of the pass descriptors, alive until the arena's destructor has synchronised
    std::list<std::vector<NrPass>> keep;
    Arena ar;
    const float *dct;
    hipError_t e = hipSuccess;
    Run(frh_context *c, hipStream_t s) : ctx(c), st(s), ar(c, s) {}
    bool ok() const { return e == hipSuccess && ar.err == hipSuccess; }
    void check(hipError_t r) {
        if (e == hipSuccess && r != hipSuccess) e = r;
    }

    // Level l + 1 of the pyramid of 2x2 means.
    Level down(const Level &s) {
        Level d{ar.get((uint64_t)((s.w + 1) / 2) * ((s.h + 1) / 2)), (s.w + 1) / 2, (s.h + 1) / 2};
        if (d.p) check(nr_down(s.p, s.w, s.h, d.p, nullptr, st));
        return d;
    }

    Level box(const float *a, uint32_t w, uint32_t h, uint32_t r) {
        Level o{ar.get((uint64_t)w * h), w, h};
        float *tmp = ar.get((uint64_t)w * h);
        if (o.p && tmp) check(nr_box(a, w, h, r, tmp, o.p, st));
        ar.put(tmp);
        return o;
    }

    // The passes of one launch: descriptors, and tiles / gather chunks so far.
    struct Stage {
        std::vector<NrPass> list;
        uint32_t tiles = 0, cell_blocks = 0, chunks = 0;
    };

    // Add a DCT pass (hard thresholding without `pilot`, Wiener with it) of `src` to stage S; its
    // result plane, or nullptr when a buffer could not be had.
    float *add_pass(Stage &S, const Level &src, const float *guide, const float *pilot, const NrModel *model,
                    float gain, float hard, uint32_t step) {
        NrPass P{};
        P.src = src.p;
        P.guide = guide;
        P.pilot = pilot;
        P.w = src.w;
        P.h = src.h;
        P.step = step;
        P.xlast = src.w + 8;
        P.ylast = src.h + 8;
        P.tiles_x = (P.xlast + NR_TILE) / NR_TILE;
        P.tiles_y = (P.ylast + NR_TILE) / NR_TILE;
        P.gx = P.xlast / step + 1 + (P.xlast % step ? 1 : 0);
        P.gy = P.ylast / step + 1 + (P.ylast % step ? 1 : 0);
        P.tile0 = S.tiles;
        P.cell0 = S.cell_blocks;
        P.chunk0 = S.chunks;
        P.model = model;
        P.dct = dct;
        P.thr = hard * gain;
        P.n2scale = gain * gain;
        P.pacc = ar.get((uint64_t)P.tiles_x * P.tiles_y * NR_EXT * NR_EXT);
        P.wgrid = ar.get((uint64_t)P.gx * P.gy);
        const uint64_t cells = (uint64_t)(P.xlast / step + 1) * (P.ylast / step + 1);
        P.wsum = ar.get(cells);
        P.out = ar.get((uint64_t)src.w * src.h);
        if (!P.pacc || !P.wgrid || !P.wsum || !P.out) return nullptr;
        S.tiles += P.tiles_x * P.tiles_y;
        S.cell_blocks += (uint32_t)((cells + 255) / 256);
        S.chunks += (src.w + 1023) / 1024 * src.h;
        S.list.push_back(P);
        return P.out;
    }

    // Launch every pass of S (descriptors copied to the device; the host copy is kept until the
    // run ends), then hand its partial buffers back.
    void run_stage(Stage &S, bool wiener) {
        if (S.list.empty() || !ok()) return;
        keep.push_back(S.list);
        const uint64_t bytes = S.list.size() * sizeof(NrPass);
        NrPass *d = reinterpret_cast<NrPass *>(ar.get((bytes + 3) / 4));
        if (!d) return;
        check(hipMemcpyAsync(d, keep.back().data(), bytes, hipMemcpyHostToDevice, st));
        check(nr_passes(d, (uint32_t)S.list.size(), S.tiles, S.cell_blocks, S.chunks, wiener, st));
        for (const NrPass &P : S.list) {
            ar.put(P.pacc);
            ar.put(P.wgrid);
            ar.put(P.wsum);
        }
        ar.put(reinterpret_cast<float *>(d));
    }

    // One plane of Run::denoise_plane (denoise.rs:595-633): its pyramid levels, the guide's
    // levels, a device model per level (nullptr: none) and the plane's settings.
    struct Plane {
        const std::vector<Level> *levels, *guides;
        std::vector<const NrModel *> models;
        const frh_nr_plane *ps;
        float *result = nullptr;
    };

    // Run::denoise_plane for several planes at once: every level with a model takes a hard pass
    // and then a Wiener pass (the full-resolution level only the hard pass when hard_only0); a
    // level without a model stays as it is. Each pass reads only its own pyramid level, so each
    // stage is one launch over every plane and level. Then each plane is recomposed from its
    // coarsest level: a finer level's detail over the denoised coarser one.
    bool denoise_planes(std::vector<Plane> &planes, float hard) {
        const size_t np = planes.size();
        std::vector<std::vector<float *>> pil(np), den(np);
        Stage s1, s2;
        for (size_t j = 0; j < np; j++) {
            const Plane &pl = planes[j];
            const size_t nl = pl.levels->size();
            pil[j].assign(nl, nullptr);
            den[j].assign(nl, nullptr);
            for (size_t li = 0; li < nl; li++) {
                if (!pl.models[li]) continue;
                const uint32_t step = li < 2 ? pl.ps->steps[li] : 2;
                pil[j][li] = add_pass(s1, (*pl.levels)[li], (*pl.guides)[li].p, nullptr, pl.models[li], pl.ps->gain,
                                      hard, step);
                if (!pil[j][li]) return false;
            }
        }
        run_stage(s1, false);
        for (size_t j = 0; j < np; j++) {
            const Plane &pl = planes[j];
            for (size_t li = 0; li < pl.levels->size(); li++) {
                if (!pil[j][li] || (li == 0 && pl.ps->hard_only0)) continue;
                const uint32_t step = li < 2 ? pl.ps->steps[li] : 2;
                den[j][li] = add_pass(s2, (*pl.levels)[li], (*pl.guides)[li].p, pil[j][li], pl.models[li], pl.ps->gain,
                                      hard, step);
                if (!den[j][li]) return false;
            }
        }
        run_stage(s2, true);
        for (size_t j = 0; j < np && ok(); j++) {
            const std::vector<Level> &levels = *planes[j].levels;
            float *result = nullptr;
            for (size_t li = levels.size(); li-- > 0 && ok();) {
                const Level &src = levels[li];
                float *d = den[j][li] ? den[j][li] : pil[j][li];
                if (den[j][li]) ar.put(pil[j][li]);
                if (!d) {
                    d = ar.get((uint64_t)src.w * src.h);
                    if (!d) return false;
                    check(hipMemcpyAsync(d, src.p, (size_t)src.w * src.h * 4, hipMemcpyDeviceToDevice, st));
                }
                if (result) {
                    const uint32_t cw = (src.w + 1) / 2, ch = (src.h + 1) / 2;
                    float *diff = ar.get((uint64_t)cw * ch);
                    if (!diff) return false;
                    check(nr_down(d, src.w, src.h, diff, result, st));
                    check(nr_add_up(d, src.w, src.h, diff, st));
                    ar.put(diff);
                    ar.put(result);
                }
                result = d;
            }
            planes[j].result = result;
        }
        return ok();
    }
};

int32_t check_spec(const frh_nr_spec *s) {
    if (!s) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL spec");
    if (s->size < offsetof(frh_nr_spec, cancel)) return fail(FRH_ERROR_INVALID_ARGUMENT, "spec size field too small");
    if (!valid_mode(s->output)) return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown output");
    for (int k = 0; k < 2; k++) {
        const frh_nr_plane &p = k ? s->chroma : s->luma;
        if (!(k ? s->has_chroma : s->has_luma)) continue;
        if (p.levels < 1 || p.levels > 8) return fail(FRH_ERROR_INVALID_ARGUMENT, "pyramid levels must be in 1..8");
        for (int i = 0; i < 2; i++)
            if (p.steps[i] == 0 || 8 % p.steps[i] != 0)
                return fail(FRH_ERROR_INVALID_ARGUMENT, "block steps must be 1, 2, 4 or 8");
        if (!std::isfinite(p.gain) || !std::isfinite(p.keep) || p.gain < 0.0f)
            return fail(FRH_ERROR_INVALID_ARGUMENT, "non-finite plane gain or keep");
    }
    if (!std::isfinite(s->hard) || s->hard < 0.0f) return fail(FRH_ERROR_INVALID_ARGUMENT, "invalid hard threshold");
    return FRH_OK;
}

void out_geometry(const frh_nr_spec *s, uint32_t w, uint32_t h, uint32_t &ow, uint32_t &oh, uint64_t &bytes,
                  uint32_t &ch, int32_t &fmt) {
    uint32_t bps;
    out_format(s->output, &ch, &bps, &fmt);
    const int32_t o = (s->orientation >= 1 && s->orientation <= 8) ? s->orientation : 1;
    const bool swap = display_mode(s->output) && o >= 5;
    ow = swap ? h : w;
    oh = swap ? w : h;
    bytes = (uint64_t)w * h * ch * bps;
}

// GpuContext::denoise (denoise.rs:817-1176) on device buffers.
int32_t denoise(frh_context *ctx, const float *rgb, uint32_t w, uint32_t h, const frh_nr_spec *spec, void *dst,
                hipStream_t st) {
    // the token is checked between the stages: split and pyramids, noise estimate, passes, guided
    // filter, encode (queued work finishes before the arena returns its buffers)
    const frh_cancel_token *cancel = spec->size >= sizeof(frh_nr_spec) ? spec->cancel : nullptr;
    if (cancelled(cancel)) return cancel_fail();
    Run R(ctx, st);
    int32_t rc = FRH_OK;
    R.dct = nullptr;
    {
        float m[64];
        dct_matrix(m);
        R.dct = static_cast<const float *>(device_table(ctx, m, sizeof(m), &rc));
        if (rc) return rc;
    }
    const bool small = w < 8 || h < 8;
    const bool luma = spec->has_luma && !small, chroma = spec->has_chroma && !small;
    const uint64_t n = (uint64_t)w * h;
    Level Y{R.ar.get(n), w, h}, U{R.ar.get(n), w, h}, V{R.ar.get(n), w, h};
    if (!R.ok()) return hip_fail(R.ar.err, "noise reduction buffers");
    R.check(nr_split(rgb, w, h, Y.p, U.p, V.p, st));
    // the luma pyramid (the guides) and the colour pyramids
    const uint32_t nl = std::max({luma ? spec->luma.levels : 1u, chroma ? spec->chroma.levels : 1u, 1u});
    const uint32_t nc = chroma ? spec->chroma.levels : 0u;
    std::vector<Level> ylev{Y}, ulev{U}, vlev{V};
    for (uint32_t l = 1; l < nl && R.ok(); l++) ylev.push_back(R.down(ylev.back()));
    for (uint32_t l = 1; l < nc && R.ok(); l++) {
        ulev.push_back(R.down(ulev.back()));
        vlev.push_back(R.down(vlev.back()));
    }
    if (!R.ok()) return R.e != hipSuccess ? hip_fail(R.e, "noise reduction") : hip_fail(R.ar.err, "noise reduction");
    if (cancelled(cancel)) return cancel_fail();
    // ---- the noise estimate: blocks measured and models fitted on the device ----
    const uint32_t nyl = luma ? spec->luma.levels : 1u;
    const uint32_t want[3] = {(luma || chroma) ? nyl : 0u, nc, nc};
    const std::vector<Level> *planes[3] = {&ylev, &ulev, &vlev};
    struct Region {
        uint32_t pi, l;
        uint64_t first;
        uint32_t cols, rows, step;
    };
    std::vector<Region> regions;
    uint64_t total = 0;
    for (uint32_t pi = 0; pi < 3; pi++)
        for (uint32_t l = 0; l < want[pi]; l++) {
            const Level &p = (*planes[pi])[l];
            const uint32_t step = estimate_step(p.w, p.h, spec->max_blocks);
            if (!step) continue;
            const uint32_t cols = (p.w / 8 + step - 1) / step, rows = (p.h / 8 + step - 1) / step;
            regions.push_back({pi, l, total, cols, rows, step});
            total += (uint64_t)cols * rows;
        }
    // gpu_models: each level's model from its blocks, a coarse level capped by the finer one. The
    // models are fitted on the device; when a level has more blocks than the fit kernels hold per
    // bin, the blocks come back to the host (through the context's pinned staging buffer when it is
    // free, else a pageable buffer of this run's own) and every level is fitted there at once.
    std::vector<NrModel> rmod;
    std::vector<char> rok;
    if (total) {
        float *d_stats = R.ar.get(total * 64);
        if (!d_stats) return hip_fail(R.ar.err, "noise estimate buffer");
        for (const Region &r : regions) {
            const Level &p = (*planes[r.pi])[r.l];
            R.check(nr_stats(p.p, ylev[r.l].p, p.w, r.step, r.cols, r.rows, R.dct, d_stats + r.first * 64, st));
        }
        if (!R.ok()) return hip_fail(R.e, "noise estimate");
        std::vector<NrRegion> nreg;
        for (const Region &r : regions) nreg.push_back(NrRegion{(uint32_t)r.first, r.cols * r.rows});
        hipError_t fe;
        const bool fitted = device_models(ctx, st, d_stats, nreg, rmod, rok, &fe);
        if (fe != hipSuccess) return hip_fail(fe, "noise model fit");
        if (!fitted) {
            std::unique_lock<std::mutex> stage(ctx->stage_mu, std::try_to_lock);
            std::vector<float> pageable;
            float *stats = nullptr;
            if (stage.owns_lock()) {
                if (ctx->h_stage_bytes < total * 256) {
                    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
                    ctx->h_stage = nullptr;
                    ctx->h_stage_bytes = 0;
                    if (hipHostMalloc(&ctx->h_stage, total * 256) == hipSuccess)
                        ctx->h_stage_bytes = total * 256;
                    else
                        ctx->h_stage = nullptr;
                }
                stats = static_cast<float *>(ctx->h_stage);
            }
            if (!stats) {
                pageable.resize(total * 64);
                stats = pageable.data();
            }
            R.check(hipMemcpyAsync(stats, d_stats, total * 256, hipMemcpyDeviceToHost, st));
            R.check(hipStreamSynchronize(st));
            if (!R.ok()) return hip_fail(R.e, "noise estimate");
            std::vector<const float *> rp;
            std::vector<size_t> rn;
            for (const Region &r : regions) {
                rp.push_back(stats + r.first * 64);
                rn.push_back((size_t)r.cols * r.rows);
            }
            models_from_stats(rp, rn, rmod, rok);
        }
        R.ar.put(d_stats);
    }
    std::vector<NrModel> mods[3];
    std::vector<bool> have[3];
    for (uint32_t pi = 0; pi < 3; pi++) {
        mods[pi].resize(want[pi]);
        have[pi].assign(want[pi], false);
        for (uint32_t l = 0; l < want[pi]; l++) {
            size_t ri = regions.size();
            for (size_t i = 0; i < regions.size(); i++)
                if (regions[i].pi == pi && regions[i].l == l) ri = i;
            bool ok = ri < regions.size() && rok[ri];
            if (ok) mods[pi][l] = rmod[ri];
            if (l > 0) {
                if (ok && have[pi][l - 1])
                    cap(mods[pi][l], next_level(mods[pi][l - 1]), LEVEL_CAP);
                else
                    ok = false;
            }
            have[pi][l] = ok;
        }
    }
    // device copies of the models and of the luma noise variance per pixel
    std::vector<NrModel> table;
    std::vector<int> index[3];
    for (uint32_t pi = 0; pi < 3; pi++)
        for (uint32_t l = 0; l < want[pi]; l++) {
            index[pi].push_back(have[pi][l] ? (int)table.size() : -1);
            if (have[pi][l]) table.push_back(mods[pi][l]);
        }
    NrVar nvar{};
    if (want[0] && have[0][0]) nvar = pixel_variance(mods[0][0]);
    NrModel *d_models = nullptr;
    NrVar *d_var = nullptr;
    {
        const uint64_t mb = std::max<size_t>(table.size(), 1) * sizeof(NrModel);
        d_models = reinterpret_cast<NrModel *>(R.ar.get((mb + 3) / 4));
        d_var = reinterpret_cast<NrVar *>(R.ar.get((sizeof(NrVar) + 3) / 4));
        if (!d_models || !d_var) return hip_fail(R.ar.err, "noise model tables");
        if (!table.empty())
            R.check(hipMemcpyAsync(d_models, table.data(), table.size() * sizeof(NrModel), hipMemcpyHostToDevice, st));
        R.check(hipMemcpyAsync(d_var, &nvar, sizeof(NrVar), hipMemcpyHostToDevice, st));
    }
    auto models_of = [&](uint32_t pi, uint32_t count) {
        std::vector<const NrModel *> v;
        for (uint32_t l = 0; l < count; l++)
            v.push_back(l < index[pi].size() && index[pi][l] >= 0 ? d_models + index[pi][l] : nullptr);
        return v;
    };
    if (cancelled(cancel)) return cancel_fail();
    // ---- the passes ----
    // every plane's passes together: luma (its own levels as guides), then the two colour planes
    // (guided by the luma levels)
    const uint32_t nlv = luma ? spec->luma.levels : 0u, ncl = chroma ? spec->chroma.levels : 0u;
    std::vector<Level> lv(ylev.begin(), ylev.begin() + nlv), gl(ylev.begin(), ylev.begin() + ncl);
    std::vector<Level> ul(ulev.begin(), ulev.begin() + ncl), vl(vlev.begin(), vlev.begin() + ncl);
    std::vector<Run::Plane> jobs;
    if (luma) jobs.push_back(Run::Plane{&lv, &lv, models_of(0, nlv), &spec->luma});
    if (chroma) {
        jobs.push_back(Run::Plane{&ul, &gl, models_of(1, ncl), &spec->chroma});
        jobs.push_back(Run::Plane{&vl, &gl, models_of(2, ncl), &spec->chroma});
    }
    if (!jobs.empty() && !R.denoise_planes(jobs, spec->hard))
        return R.e != hipSuccess ? hip_fail(R.e, "noise reduction") : hip_fail(R.ar.err, "noise reduction");
    float *y2 = luma ? jobs[0].result : nullptr;
    if (y2 && spec->luma.keep > 0.0f) R.check(nr_blend(y2, Y.p, spec->luma.keep, n, st));
    const float *y2r = y2 ? y2 : Y.p;
    const float *u2r = U.p, *v2r = V.p;
    if (cancelled(cancel)) return cancel_fail();
    if (chroma && R.ok()) {
        float *ud = jobs[luma ? 1 : 0].result, *vd = jobs[luma ? 2 : 1].result;
        if (!ud || !vd || !R.ok())
            return R.e != hipSuccess ? hip_fail(R.e, "noise reduction") : hip_fail(R.ar.err, "noise reduction");
        // the guided filter (Guide, mod.rs:337-382), coefficients at half resolution
        const uint32_t r = std::max(spec->chroma_radius, 1u);
        const Level full{const_cast<float *>(y2r), w, h};
        Level I = R.down(full);
        const uint64_t hn = (uint64_t)I.w * I.h;
        Level mean = R.box(I.p, I.w, I.h, r);
        float *sq = R.ar.get(hn);
        if (!sq || !mean.p || !I.p) return hip_fail(R.ar.err, "guided filter buffers");
        R.check(nr_mul(I.p, I.p, sq, hn, st));
        Level den = R.box(sq, I.w, I.h, r);
        R.ar.put(sq);
        if (!den.p) return hip_fail(R.ar.err, "guided filter buffers");
        R.check(nr_den(den.p, mean.p, d_var, hn, st));
        float *outs[2] = {nullptr, nullptr};
        float *srcs[2] = {ud, vd};
        const float *origs[2] = {U.p, V.p};
        for (int c = 0; c < 2 && R.ok(); c++) {
            Level ps = R.down(Level{srcs[c], w, h});
            R.ar.put(srcs[c]);
            Level mp = R.box(ps.p, I.w, I.h, r);
            float *prod = R.ar.get(hn);
            if (!ps.p || !mp.p || !prod) return hip_fail(R.ar.err, "guided filter buffers");
            R.check(nr_mul(I.p, ps.p, prod, hn, st));
            R.ar.put(ps.p);
            Level cov = R.box(prod, I.w, I.h, r);
            R.ar.put(prod);
            float *a = R.ar.get(hn), *b = R.ar.get(hn);
            if (!cov.p || !a || !b) return hip_fail(R.ar.err, "guided filter buffers");
            R.check(nr_gf_coef(cov.p, mp.p, mean.p, den.p, a, b, hn, st));
            R.ar.put(cov.p);
            R.ar.put(mp.p);
            Level ma = R.box(a, I.w, I.h, r), mbx = R.box(b, I.w, I.h, r);
            R.ar.put(a);
            R.ar.put(b);
            float *o = R.ar.get(n);
            if (!ma.p || !mbx.p || !o) return hip_fail(R.ar.err, "guided filter buffers");
            R.check(nr_gf_apply(ma.p, mbx.p, y2r, w, h, o, st));
            R.ar.put(ma.p);
            R.ar.put(mbx.p);
            if (spec->chroma.keep > 0.0f) R.check(nr_blend(o, origs[c], spec->chroma.keep, n, st));
            outs[c] = o;
        }
        u2r = outs[0];
        v2r = outs[1];
    }
    if (!R.ok()) return R.e != hipSuccess ? hip_fail(R.e, "noise reduction") : hip_fail(R.ar.err, "noise reduction");
    if (cancelled(cancel)) return cancel_fail();
    // merge and encode (denoise_encode.wgsl: the develop's tone, encoders and orientation)
    DevArgs a;
    memset(&a, 0, sizeof(a));
    a.lut8 = ctx->d_lut8;
    a.srgb16 = reinterpret_cast<const float2 *>(ctx->d_srgb16);
    a.wtab = ctx->d_wtab;
    a.out_mode = spec->output;
    a.out_mode2 = OUT_NONE;
    a.orientation = (spec->orientation >= 1 && spec->orientation <= 8) ? spec->orientation : 1;
    a.dev_w = w;
    a.dev_h = h;
    a.dst_w = (display_mode(spec->output) && a.orientation >= 5) ? h : w;
    a.dst = dst;
    a.band_y0 = 0;
    a.band_y1 = h;
    if ((rc = output_tables(ctx, spec->tone, spec->curve, a))) return rc;
    R.check(nr_encode(a, y2r, u2r, v2r, st));
    // the arena's buffers go back to the pool once the stream has used them
    R.check(hipStreamSynchronize(st));
    if (!R.ok()) return hip_fail(R.e, "noise reduction");
    if (cancelled(cancel)) return cancel_fail();
    return FRH_OK;
}

}  // namespace

extern "C" int32_t frh_nr_spec_for(uint32_t w, uint32_t h, float luma, float chroma, int32_t output,
                                   frh_nr_spec *spec) {
    if (!spec) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL spec");
    auto clamp01 = [](float v) { return v != v ? v : std::min(std::max(v, 0.0f), 1.0f); };
    luma = clamp01(luma);
    chroma = clamp01(chroma);
    memset(spec, 0, sizeof(*spec));
    spec->size = sizeof(frh_nr_spec);
    spec->has_luma = luma > 0.0f;
    spec->has_chroma = chroma > 0.0f;
    // luma: level_count(4), gain 1, step 2, the Wiener pass skipped at full resolution up to
    // strength 0.25, (1 - s)^1.2 of the original noise kept
    spec->luma = frh_nr_plane{level_count(w, h, 4), 1.0f, {2, 2}, luma <= 0.25f ? 1 : 0, std::pow(1.0f - luma, 1.2f)};
    // chroma: level_count(6), deviations raised above strength 0.5, step 4 on the two finest
    // levels, hard pass only at full resolution, 1 - 2s of the original kept
    spec->chroma = frh_nr_plane{level_count(w, h, 6), 1.0f + 2.0f * std::max(chroma - 0.5f, 0.0f), {4, 4}, 1,
                                std::max(1.0f - 2.0f * chroma, 0.0f)};
    const uint32_t rfull = (uint32_t)std::max(std::round(6.0f * (0.5f + chroma)), 1.0f);
    spec->chroma_radius = std::max((rfull + 1) / 2, 1u);
    spec->hard = 2.7f;
    spec->max_blocks = 50000;
    spec->output = output;
    spec->orientation = 1;
    return FRH_OK;
}

extern "C" int32_t frh_nr_model_from_stats(const float *stats, uint32_t count, uint32_t *bins, float *means,
                                           float *sigmas) {
    try {
        if (!stats || !bins || !means || !sigmas) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        NrModel md{};
        if (!from_stats(stats, count, md)) return fail(FRH_ERROR_NOT_AVAILABLE, "fewer than 32 measured blocks");
        *bins = md.n;
        memcpy(means, md.m, md.n * sizeof(float));
        memcpy(sigmas, md.s, (size_t)md.n * 64 * sizeof(float));
        return FRH_OK;
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_nr_model_fit(frh_context *ctx, const float *stats, uint32_t count, uint32_t *bins,
                                    float *means, float *sigmas) {
    try {
        if (!ctx || !stats || !bins || !means || !sigmas) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        (void)hipSetDevice(ctx->device);
        NrModel md{};
        bool ok = false;
        if (count > 0) {
            hipError_t e = hipSuccess;
            const uint64_t bytes = (uint64_t)count * 256;
            void *d = ctx->alloc(bytes, &e);
            if (!d) return hip_fail(e, "noise model blocks");
            e = hipMemcpy(d, stats, bytes, hipMemcpyHostToDevice);
            std::vector<NrModel> m;
            std::vector<char> k;
            hipError_t fe = e;
            const bool fitted =
                e == hipSuccess && device_models(ctx, ctx->stream, static_cast<const float *>(d), {NrRegion{0, count}},
                                                 m, k, &fe);
            ctx->release(d, bytes);
            if (fe != hipSuccess) return hip_fail(fe, "noise model fit");
            if (fitted) {
                ok = k[0] != 0;
                md = m[0];
            } else {
                ok = from_stats(stats, count, md);  // more blocks per bin than the device fit holds
            }
        }
        if (!ok) return fail(FRH_ERROR_NOT_AVAILABLE, "fewer than 32 measured blocks");
        *bins = md.n;
        memcpy(means, md.m, md.n * sizeof(float));
        memcpy(sigmas, md.s, (size_t)md.n * 64 * sizeof(float));
        return FRH_OK;
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_denoise_into(frh_context *ctx, const float *rgb, uint32_t w, uint32_t h,
                                    const frh_nr_spec *spec, void *dst, uint64_t dst_bytes, void *stream) {
    try {
        if (!ctx || !rgb || !dst) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        int32_t rc = check_spec(spec);
        if (rc) return rc;
        if (w == 0 || h == 0) return fail(FRH_ERROR_INVALID_ARGUMENT, "empty image");
        uint32_t ow, oh, ch;
        uint64_t bytes;
        int32_t fmt;
        out_geometry(spec, w, h, ow, oh, bytes, ch, fmt);
        if (dst_bytes < bytes) return fail(FRH_ERROR_INVALID_ARGUMENT, "output buffer too small");
        (void)hipSetDevice(ctx->device);
        return denoise(ctx, rgb, w, h, spec, dst, stream ? (hipStream_t)stream : ctx->stream);
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_run_denoise(frh_context *ctx, const frh_image *input, const frh_nr_spec *spec,
                                   frh_image **out, frh_timings *timings) {
    try {
        if (!ctx || !input || !out) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        if (input->fmt != FRH_FORMAT_F32 || input->ch != 3)
            return fail(FRH_ERROR_INVALID_ARGUMENT, "the noise reduction needs linear f32 RGB input");
        int32_t rc = check_spec(spec);
        if (rc) return rc;
        uint32_t ow, oh, ch;
        uint64_t bytes;
        int32_t fmt;
        out_geometry(spec, input->w, input->h, ow, oh, bytes, ch, fmt);
        (void)hipSetDevice(ctx->device);
        hipError_t e = hipSuccess;
        void *d = ctx->alloc(bytes, &e);
        if (!d) return hip_fail(e, "output allocation");
        frh_image *img = new frh_image{ctx, d, bytes, ow, oh, ch, fmt};
        auto t0 = std::chrono::steady_clock::now();
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        const bool timed = hipEventCreate(&ev0) == hipSuccess && hipEventCreate(&ev1) == hipSuccess;
        if (timed) (void)hipEventRecord(ev0, ctx->stream);
        rc = denoise(ctx, static_cast<const float *>(input->dptr), input->w, input->h, spec, d, ctx->stream);
        if (timed) (void)hipEventRecord(ev1, ctx->stream);
        e = hipStreamSynchronize(ctx->stream);
        float kms = 0.0f;
        if (timed && hipEventElapsedTime(&kms, ev0, ev1) != hipSuccess) kms = 0.0f;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (rc == FRH_OK && e != hipSuccess) rc = hip_fail(e, "noise reduction");
        if (rc != FRH_OK) {
            frh_image_free(img);
            return rc;
        }
        if (timings) {
            timings->upload_ms = 0.0;
            timings->compute_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            timings->kernel_ms = kms;
            timings->has_kernel_ms = timed ? 1 : 0;
            timings->bands = 1;
        }
        *out = img;
        return FRH_OK;
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}
