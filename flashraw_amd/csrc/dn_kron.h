// The Kronecker-form MFMA CFA-plane denoiser (dn_kron.hip): arguments, operand tables, launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace frh {

struct DnkArgs {
    const uint16_t *src;
    uint16_t *dst;    // may be null
    float *dst_f;     // may be null
    const uint16_t *kf, *ki;  // per-lane operand words of the forward / inverse matrices (dn_kron_tables)
    uint32_t stride;  // samples per mosaic row
    int pw, ph;       // plane size (multiples of 4)
    int plx, ply;     // last block position in x and y (pw - 8, ph - 8)
    int strips, segs, seg_steps, nsteps;  // strips of DNK_STRIP columns; segments of seg_steps 8-row steps
    float black[4];
    float t2a, t2b;   // squared threshold in the forward's scaled units: t2a * max(mean, 0) + t2b
};

// Strip width in plane columns owned by one wave (16 blocks at stride 8: the first a halo, the
// last owning its left half).
constexpr int DNK_STRIP = 116;

void dn_kron_tables(const float d[64], bool bf16, std::vector<uint16_t> &kf, std::vector<uint16_t> &ki);
// Steps per wave for `strips` strips and `nsteps` steps on a device with `num_cus` compute units.
int dn_kron_seg_steps(int strips, int nsteps, int num_cus);
hipError_t launch_dn_kron(const DnkArgs &a, bool bf16, hipStream_t st);

}  // namespace frh
