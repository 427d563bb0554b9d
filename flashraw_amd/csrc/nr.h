// Luma / chroma noise reduction of developed linear RGB: device tables and kernel launchers
// shared by the kernels (nr.hip) and the host pipeline (nr_api.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "frh_device.h"

namespace frh {

constexpr int NR_BINS = 12;  // most bins of a noise model (Model::from_stats, denoise/mod.rs:815)

// The noise model of one pyramid level (denoise.rs:30-39): per bin of luma mean, the deviation
// of each DCT frequency (entry l * 8 + k: horizontal l, vertical k; entry 0 unused).
struct NrModel {
    uint32_t n;
    float m[NR_BINS];
    float s[NR_BINS][64];
};

// Per-pixel noise variance by luma mean (Variance, denoise/mod.rs:255-276); n = 0: no model.
struct NrVar {
    uint32_t n;
    float m[NR_BINS];
    float v[NR_BINS];
};

// One DCT pass over a w x h plane padded symmetrically by 8 (blocks at the positions 0, step,
// 2 step, ... of each padded side, and always its last position `xlast` / `ylast`). Tiles of 64 x
// 64 block positions accumulate their blocks' weighted samples in a 72 x 72 partial each; every
// block's weight goes to a grid of gy x gx block positions (the last position, when it is off the
// step, in the last column / row). Samples whose padded coordinates share px / step and py / step
// are covered by the same blocks, so the weights' sums are formed once per such cell (`wsum`,
// (ylast / step + 1) x (xlast / step + 1) cells). The passes of one stage run in one launch: their
// descriptors sit in device memory, and `tile0` / `cell0` / `chunk0` are the pass's first tile,
// first 256-cell block and first 1024-sample gather chunk in that launch.
struct NrPass {
    const float *src, *guide, *pilot;
    uint32_t w, h;
    uint32_t step;
    uint32_t xlast, ylast;
    uint32_t tiles_x, tiles_y;
    uint32_t gx, gy;
    uint32_t tile0, cell0, chunk0;
    const NrModel *model;
    const float *dct;  // the 8 x 8 DCT-II matrix, dct[k * 8 + i]
    float thr;         // hard threshold times the plane's gain
    float n2scale;     // gain squared
    float *pacc;       // tiles_x * tiles_y partials of NR_EXT x NR_EXT
    float *wgrid;      // gy x gx block weights
    float *wsum;       // the weights covering each cell
    float *out;        // the w x h result
};

// The noise model of one level fitted on the device (Model::from_stats, denoise/mod.rs:803-848):
// the level's measured blocks are stats[first .. first + count) (64 floats each: the luma mean,
// then the squared coefficients). The kernels sort them by mean (rank selections on total-order
// keys, ties in block order), cut them into bins of `per`, find each bin's median activity and
// the median of each squared coefficient over its quieter blocks.
struct NrRegion {
    uint32_t first, count;
};
constexpr uint32_t NR_FIT_LIST = 4608;  // most blocks per bin the fit kernels hold in LDS
struct NrFit {
    uint32_t n;                 // finite blocks
    uint32_t nbins, per;        // bins, blocks per bin (the last may hold fewer)
    uint32_t start[NR_BINS];    // block (region-local) of rank b * per
    uint32_t nflat[NR_BINS];    // blocks at most the bin's median activity
    float mean[NR_BINS];        // the luma mean of the bin's median block
    int32_t sig[NR_BINS][64];   // total-order keys of the median squared coefficients (entry 0 unused)
};

constexpr int NR_TILE = 64;
constexpr int NR_EXT = NR_TILE + 8;

hipError_t nr_split(const float *rgb, uint32_t w, uint32_t h, float *y, float *u, float *v, hipStream_t st);
hipError_t nr_down(const float *src, uint32_t w, uint32_t h, float *dst, const float *minus, hipStream_t st);
hipError_t nr_add_up(float *d, uint32_t w, uint32_t h, const float *coarse, hipStream_t st);
hipError_t nr_blend(float *d, const float *orig, float keep, uint64_t n, hipStream_t st);
hipError_t nr_mul(const float *a, const float *b, float *out, uint64_t n, hipStream_t st);
hipError_t nr_box(const float *a, uint32_t w, uint32_t h, uint32_t r, float *tmp, float *out, hipStream_t st);
hipError_t nr_den(float *var, const float *mean, const NrVar *nv, uint64_t n, hipStream_t st);
hipError_t nr_gf_coef(const float *cov, const float *mp, const float *mean, const float *den, float *a, float *b,
                      uint64_t n, hipStream_t st);
hipError_t nr_gf_apply(const float *ma, const float *mb, const float *g, uint32_t w, uint32_t h, float *out,
                       hipStream_t st);
hipError_t nr_stats(const float *p, const float *g, uint32_t w, uint32_t step, uint32_t cols, uint32_t rows,
                    const float *dct, float *out, hipStream_t st);
// Fit the noise model of every region (fit[i] zeroed first) from `stats`; key, act and flat are
// scratch of `total` entries each.
hipError_t nr_fit(const float *stats, uint32_t total, const NrRegion *reg, uint32_t nreg, uint32_t *key, float *act,
                  uint32_t *flat, NrFit *fit, hipStream_t st);
// The passes d_list[0 .. n - 1] (all Wiener passes when `wiener`, else all hard-threshold passes),
// `tiles` tiles, `cell_blocks` blocks of 256 cells and `chunks` gather chunks in all.
hipError_t nr_passes(const NrPass *d_list, uint32_t n, uint32_t tiles, uint32_t cell_blocks, uint32_t chunks,
                     bool wiener, hipStream_t st);
hipError_t nr_encode(const DevArgs &a, const float *y, const float *u, const float *v, hipStream_t st);

}  // namespace frh
