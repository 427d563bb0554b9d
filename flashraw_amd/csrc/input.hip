// Input kernels (gfx950): packed payloads and u16 rasters to u16 samples through the optional
// linearisation curve (GpuOp::Unpack, crates/flashraw-gpu/src/run.rs:503-512), and normalised
// f32 samples (GpuOp::Normalize, run.rs:513-517). Integer work, bound by HBM: one thread writes
// one 32-bit word (two samples) of the output, which is laid out exactly like a host Vec<u16>.
//
// A packed sample is cut from the three logical bytes at bitpos / 8 of its row (the statement of
// crates/flashraw-gpu/src/kernels.rs:333-357, which the reference proves equal to the codecs'
// bit pump, crates/flashraw-codecs/src/packed.rs:137-180): the word orders read the physical byte
// k ^ 1 (16-bit words) or k ^ 3 (32-bit words), and bytes at or past the row's data or past the
// payload read as 0.
#include <hip/hip_runtime.h>

#include "develop_common.h"
#include "frh_launch.h"

namespace frh {
namespace {

template <int ORD>
__device__ __forceinline__ uint32_t row_byte(const InArgs &a, uint64_t row, uint32_t k) {
    const uint32_t p = ORD == 2 ? (k ^ 1u) : (ORD == 3 ? (k ^ 3u) : k);
    if (p >= a.row_bytes) return 0u;
    const uint64_t at = row + p;
    return at < a.len ? (uint32_t)static_cast<const uint8_t *>(a.src)[at] : 0u;
}

template <int ORD>
__device__ __forceinline__ uint32_t packed_sample(const InArgs &a, uint32_t i, uint32_t y) {
    const uint64_t row = (uint64_t)y * a.stride;
    const uint32_t bitpos = i * a.bits, k = bitpos >> 3, sh = bitpos & 7u;
    const uint32_t b0 = row_byte<ORD>(a, row, k), b1 = row_byte<ORD>(a, row, k + 1u),
                   b2 = row_byte<ORD>(a, row, k + 2u);
    const uint32_t mask = (1u << a.bits) - 1u;
    if (ORD == 1) return ((b0 | (b1 << 8) | (b2 << 16)) >> sh) & mask;
    return (((b0 << 16) | (b1 << 8) | b2) >> (24u - sh - a.bits)) & mask;
}

// Integer sample s (0 .. n * rows) of the window: packed (ORD 0..3) or u16 (ORD 4), linearised.
template <int ORD>
__device__ __forceinline__ uint32_t window_sample(const InArgs &a, uint32_t s) {
    const uint32_t ly = s / a.n, r = s - ly * a.n;
    uint32_t v;
    if (ORD == 4)
        v = static_cast<const uint16_t *>(a.src)[(uint64_t)(a.y0 + ly) * a.stride + a.x0 + r];
    else
        v = packed_sample<ORD>(a, a.x0 + r, a.y0 + ly);
    if (a.lut) v = a.lut[v < a.lut_len ? v : a.lut_len - 1u];
    return v & 0xffffu;
}

template <int ORD>
__global__ __launch_bounds__(256) void k_unpack(InArgs a, uint32_t *dst, uint32_t words) {
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    if (w >= words) return;
    const uint32_t total = a.n * a.rows, s0 = 2u * w;
    uint32_t word = window_sample<ORD>(a, s0);
    if (s0 + 1u < total) {
        word |= window_sample<ORD>(a, s0 + 1u) << 16;
        dst[w] = word;
    } else {
        reinterpret_cast<uint16_t *>(dst)[s0] = (uint16_t)word;  // the last, odd sample
    }
}

// Calib::norm of component c at window (x, y) (process.rs:288-298; normalize.wgsl), then max 0.
__global__ __launch_bounds__(256) void k_normalize(InArgs a, float *dst) {
    const uint64_t s = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (s >= (uint64_t)a.n * a.rows) return;
    const uint32_t ly = (uint32_t)(s / a.n), r = (uint32_t)(s - (uint64_t)ly * a.n);
    const uint32_t x = r / a.cpp, c = r - x * a.cpp;
    const uint64_t at = (uint64_t)(a.y0 + ly) * a.stride + a.x0 + r;
    const float v = a.kind == 1 ? static_cast<const float *>(a.src)[at]
                                : (float)static_cast<const uint16_t *>(a.src)[at];
    const uint32_t X = x + a.ox, Y = ly + a.oy;
    const uint32_t P = a.cal_pw * a.cal_ph;
    const uint32_t i = c * P + (Y % a.cal_ph) * a.cal_pw + X % a.cal_pw;
    float b = a.gcal[i];
    if (a.n_rowd) b += Y < a.n_rowd ? a.rowd[Y] : 0.0f;
    if (a.n_cold) b += X < a.n_cold ? a.cold[X] : 0.0f;
    dst[s] = fmaxf((v - b) * a.gcal[a.cpp * P + i], 0.0f);
}

}  // namespace

hipError_t launch_unpack(const InArgs &in, int order, uint16_t *dst, hipStream_t stream) {
    const uint64_t total = (uint64_t)in.n * in.rows;
    if (total == 0) return hipSuccess;
    const uint32_t words = (uint32_t)((total + 1) / 2);
    const dim3 grid((words + 255) / 256), block(256);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    if (in.kind == 0)
        hipLaunchKernelGGL(k_unpack<4>, grid, block, 0, stream, in, d, words);
    else if (order == 0)
        hipLaunchKernelGGL(k_unpack<0>, grid, block, 0, stream, in, d, words);
    else if (order == 1)
        hipLaunchKernelGGL(k_unpack<1>, grid, block, 0, stream, in, d, words);
    else if (order == 2)
        hipLaunchKernelGGL(k_unpack<2>, grid, block, 0, stream, in, d, words);
    else
        hipLaunchKernelGGL(k_unpack<3>, grid, block, 0, stream, in, d, words);
    return hipGetLastError();
}

hipError_t launch_normalize(const InArgs &in, float *dst, hipStream_t stream) {
    const uint64_t total = (uint64_t)in.n * in.rows;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_normalize, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, in, dst);
    return hipGetLastError();
}

}  // namespace frh
