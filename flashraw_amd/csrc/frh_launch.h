// Host-side launchers of the device kernels (defined next to the kernels they launch).
#pragma once

#include <hip/hip_runtime.h>

#include "frh_device.h"

namespace frh {

// Develop rows [a.band_y0, a.band_y1) of the developed image on `stream` (bilinear, or AHD).
hipError_t launch_develop(const DevArgs &a, bool ahd, hipStream_t stream, int num_cus);

// The edge-aware demosaic (Bayer) over rows [a.band_y0, a.band_y1).
hipError_t launch_develop_edge(const DevArgs &a, hipStream_t stream);

// The AHD row walker over rows [a.band_y0, a.band_y1), `rows_per_strip` rows per wave.
hipError_t launch_develop_ahd_walk(const DevArgs &a, hipStream_t stream, int rows_per_strip);
// Adds to *count the floats (all of [2^-96, 2^128) and +-0) where the kernels' square root differs from
// the correctly rounded one.
hipError_t check_sqrt(unsigned long long *count, hipStream_t stream);

// Unpack / re-window (with the linearisation curve) the window of `in` into u16 samples at `dst`
// (in.n * in.rows samples, 4-byte aligned); PackOrder `order` for packed payloads.
hipError_t launch_unpack(const InArgs &in, int order, uint16_t *dst, hipStream_t stream);

// Normalised samples max((v - black) * scale, 0) of the window of `in` (u16 or f32) into `dst`.
hipError_t launch_normalize(const InArgs &in, float *dst, hipStream_t stream);

// The generic develop (any CFA, no CFA, several components, half size) over rows
// [a.band_y0, a.band_y1).
hipError_t launch_develop_generic(const DevArgs &a, hipStream_t stream);

}  // namespace frh
