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

}  // namespace frh
