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

// Eight consecutive samples of a window row per thread. Where the 24 bytes from the row-aligned
// dword holding the group's first bit lie inside the row's data and the payload, the group is cut
// from seven aligned dword loads: each physical little-endian dword becomes its four logical bytes (byte order:
// as is; MSB bytes: swapped; 16-bit words: halves swapped; 32-bit words: as is), and each sample
// is a shift of a 64-bit window - the same bits the three-byte statement reads. Other groups (a
// row's last bytes, a truncated payload) take that statement sample by sample.
template <int ORD>
__device__ __forceinline__ uint32_t logical_word(uint32_t d) {
    if (ORD == 0) return __builtin_bswap32(d);
    if (ORD == 2) return (d >> 16) | (d << 16);
    return d;
}

template <int ORD>
__global__ __launch_bounds__(256) void k_unpack8(InArgs a, uint16_t *dst, uint32_t groups_per_row, uint32_t groups) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    if (gid >= groups) return;
    const uint32_t ly = gid / groups_per_row, g = gid - ly * groups_per_row;
    const uint32_t s0 = 8u * g;  // first sample of the group in the window row
    const uint32_t cnt = min(8u, a.n - s0);
    uint32_t v[8];
    const uint64_t row = (uint64_t)(a.y0 + ly) * a.stride;
    const uint32_t bit0 = (a.x0 + s0) * a.bits, a4 = (bit0 >> 3) & ~3u;
    const uint64_t at = row + a4;  // payload offset of the group's first (row-aligned) dword
    const uint32_t sh = (uint32_t)(at & 3u);
    const bool aligned_src = (reinterpret_cast<uintptr_t>(a.src) & 3u) == 0;
    if (cnt == 8u && aligned_src && a4 + 24u <= a.row_bytes && at - sh + 28u <= a.len) {
        // seven aligned dwords cover the 24 bytes at `at`
        const uint32_t *p = reinterpret_cast<const uint32_t *>(static_cast<const uint8_t *>(a.src) + (at - sh));
        uint32_t P[7];
#pragma unroll
        for (int k = 0; k < 7; k++) P[k] = p[k];
        uint32_t W[6];
#pragma unroll
        for (int k = 0; k < 6; k++)
            W[k] = logical_word<ORD>((uint32_t)((((uint64_t)P[k + 1] << 32) | P[k]) >> (8u * sh)));
        const uint32_t mask = (1u << a.bits) - 1u;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t o = bit0 - a4 * 8u + (uint32_t)j * a.bits, q = o >> 5, r = o & 31u;
            const uint32_t lo = q == 0 ? W[0] : q == 1 ? W[1] : q == 2 ? W[2] : q == 3 ? W[3] : W[4];
            const uint32_t hi = q == 0 ? W[1] : q == 1 ? W[2] : q == 2 ? W[3] : q == 3 ? W[4] : W[5];
            if (ORD == 1) {
                const uint64_t win = (uint64_t)lo | ((uint64_t)hi << 32);
                v[j] = (uint32_t)(win >> r) & mask;
            } else {
                const uint64_t win = ((uint64_t)lo << 32) | hi;
                v[j] = (uint32_t)((win << r) >> (64u - a.bits));
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = (uint32_t)j < cnt ? packed_sample<ORD>(a, a.x0 + s0 + j, a.y0 + ly) : 0u;
    }
    if (a.lut)
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = a.lut[v[j] < a.lut_len ? v[j] : a.lut_len - 1u];
    uint16_t *o = dst + (uint64_t)ly * a.n + s0;
    if (cnt == 8u && ((reinterpret_cast<uintptr_t>(o) & 3u) == 0)) {
        uint32_t *o32 = reinterpret_cast<uint32_t *>(o);
#pragma unroll
        for (int k = 0; k < 4; k++) o32[k] = (v[2 * k] & 0xffffu) | (v[2 * k + 1] << 16);
    } else {
        for (uint32_t j = 0; j < cnt; j++) o[j] = (uint16_t)v[j];
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
    if (in.kind == 0) {
        const uint32_t words = (uint32_t)((total + 1) / 2);
        hipLaunchKernelGGL(k_unpack<4>, dim3((words + 255) / 256), dim3(256), 0, stream, in,
                           reinterpret_cast<uint32_t *>(dst), words);
        return hipGetLastError();
    }
    const uint32_t gpr = (in.n + 7) / 8, groups = gpr * in.rows;
    const dim3 grid((groups + 255) / 256), block(256);
    if (order == 0)
        hipLaunchKernelGGL(k_unpack8<0>, grid, block, 0, stream, in, dst, gpr, groups);
    else if (order == 1)
        hipLaunchKernelGGL(k_unpack8<1>, grid, block, 0, stream, in, dst, gpr, groups);
    else if (order == 2)
        hipLaunchKernelGGL(k_unpack8<2>, grid, block, 0, stream, in, dst, gpr, groups);
    else
        hipLaunchKernelGGL(k_unpack8<3>, grid, block, 0, stream, in, dst, gpr, groups);
    return hipGetLastError();
}

hipError_t launch_normalize(const InArgs &in, float *dst, hipStream_t stream) {
    const uint64_t total = (uint64_t)in.n * in.rows;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_normalize, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, in, dst);
    return hipGetLastError();
}

}  // namespace frh
