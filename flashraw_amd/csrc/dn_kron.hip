// CFA-plane DCT block denoiser, stride 4, as Kronecker-form MFMA contractions (gfx950).
//
// The 2-D orthonormal DCT-II of an 8x8 block is one 64 x 64 matrix (D (x) D, coefficient (u, v)
// against pixel (i, j): d[u][i] d[v][j]) applied to the block's 64 pixels, and the inverse is its
// transpose. One v_mfma_f32_16x16x32_{f16,bf16} multiplies 16 coefficient rows by 32 pixels of
// 16 blocks, so a wave transforms 16 blocks with 8 MFMAs forward and 8 back, with no conversion
// between the separable stages. Operands are f16 / bf16, accumulation f32.
//
// Block classes. At stride 4 every pixel lies in four blocks, one of each class (py, px): blocks
// at (8 a + 4 py, 8 b + 4 px). A wave walks down a strip of the plane in steps of 8 rows and per
// step transforms four tiles of 16 blocks, one per class: class (0, 0) at x = x0 + 8 (n - 1) and
// class (0, 1) at x0 + 8 n - 4 for lane column n = 0..15 (block row y0), classes (1, *) the same
// at block row y0 + 4. Lane column 0 of class (0, 0) is a halo block: it is transformed but owns
// no output. The class-(0, 1) block of column 15 would need column 16's samples, so column 15 owns
// only its left half: a strip owns the 116 plane columns [x0, x0 + 116).
//
// Register layout. In the MFMA's result (lane l holds rows 4 (l / 16) + r of column l % 16 of a
// 16 x 16 tile), the inverse's output row index 16 j + 4 g + r is pixel (4 (j & 1) + g, 4 (j >> 1)
// + r) of block l % 16. A pixel of a class-(0, 0) block and the pixels of the other classes'
// blocks that cover it then sit in the same lane group g and register r: the class-(0, 1) block
// in register j ^ 2 of the same lane column or of column n - 1 (a DPP row shift), the
// class-(1, *) blocks in register j ^ 1 of this step's tiles or of the previous step's (carried
// in registers down the strip). The overlap-add therefore runs in the inverse products'
// accumulators (each output register's chain of MFMAs starts from the carry or from zero), plus
// one DPP-shifted add for the column n - 1 terms; nothing goes through LDS memory.
//
// Throughput. On MI355X the vector ALU and the MFMA unit of a SIMD overlap little, so a step costs
// about its vector instructions plus its MFMAs, and a lone wave per SIMD runs nearly as fast as
// two: the launch is sized to one round of two waves per SIMD (dn_kron_seg_steps), blocks
// outside the plane drop their coefficients through an infinite threshold, the DC broadcast goes
// through ds_bpermute, and stores are issued by every lane (buffer stores, dropped past the end)
// so the step loop has no branches around memory operations.
//
// Inputs. Lane (g, n) loads row 4 h + g of its class-(0, 0) block as two 16-byte loads (both
// planes of the mosaic row parity), converts them to 16-bit operands (sample minus an offset per
// step, taken from one sample of the step's rows, so the operands carry local contrast; the
// saturating 16-bit difference is exact while samples differ by at most 32767, which holds for any
// mosaic of 15 bits or fewer), and builds the class-(0, 1) operand from its own right half and
// column n + 1's left half (DPP). An output's offset term is the mean offset of the blocks that
// cover it.
//
// Threshold. A coefficient c is kept when c^2 >= t^2 with t^2 = thr^2 (sig_a max(mean, 0) +
// sig_b) in the transform's scaled units (the DC coefficient always); the test runs on the
// coefficients rounded to f16, two at a time (one fused multiply-add and a sign mask per pair),
// for bf16 operands too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "dn_kron.h"

namespace frh {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef __bf16 bh2 __attribute__((ext_vector_type(2)));
typedef short s2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mma(u4 a, u4 b, f4 c, _Float16) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c, 0, 0, 0);
}
__device__ __forceinline__ f4 mma(u4 a, u4 b, f4 c, __bf16) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

// Packed 16-bit differences a - b, saturated. (Kept in a function of scalar words: this
// compiler drops the other elements when a bit cast of packed-vector arithmetic is assigned into
// an element of a 4-word vector inside an unrolled loop, so results are assembled from plain
// arrays of words.)
__device__ __forceinline__ uint32_t sub2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(__builtin_bit_cast(s2, a), __builtin_bit_cast(s2, b)));
}
__device__ __forceinline__ u4 words(const uint32_t t[4]) { return u4{t[0], t[1], t[2], t[3]}; }

// Two 16-bit integers (sample minus offset) as a packed pair of operands. For f16 the high half
// is converted in place (SDWA: source word 1 into destination word 1, word 0 preserved), so the
// pair costs two conversions and no packing instruction.
__device__ __forceinline__ uint32_t int_pair(uint32_t w, _Float16) {
    const _Float16 lo = (_Float16)__builtin_bit_cast(s2, w).x;
    uint32_t r = (uint32_t)__builtin_bit_cast(uint16_t, lo);
    asm("v_cvt_f16_i16_sdwa %0, %1 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:WORD_1" : "+v"(r) : "v"(w));
    return r;
}
__device__ __forceinline__ uint32_t int_pair(uint32_t w, __bf16) {
    const s2 v = __builtin_bit_cast(s2, w);
    const bh2 r = {(__bf16)(float)v.x, (__bf16)(float)v.y};
    return __builtin_bit_cast(uint32_t, r);
}

// Row shifts inside 16-lane rows: lane n reads lane n + 1 (shl) / n - 1 (shr); 0 past the row.
__device__ __forceinline__ uint32_t shl1(uint32_t v) { return __builtin_amdgcn_update_dpp(0u, v, 0x101, 0xF, 0xF, true); }
__device__ __forceinline__ float shr1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xF, 0xF, true));
}

// Lanes 0..15's value in every 16-lane row (ds_bpermute from lane n = l % 16: the LDS crossbar,
// no vector ALU slots; `addr` is 4 n).
__device__ __forceinline__ float bcast_row0(float v, int addr) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}

// The thresholded coefficients of one tile (forward result D[4], this lane's block threshold t2
// and the threshold tdc of its coefficient in D[0][0]: t2, or -1 for a kept DC coefficient) as
// the inverse's two K-halves. A block outside the plane has t2 = tdc = +inf: every coefficient is
// dropped by the threshold itself (a select per word after the bit operation would cost 8 more
// instructions).
// Ones in the halves of a pair of coefficients (rounded to f16) where c^2 < t (t as a packed
// pair): one fused multiply-add and a sign mask. Coefficients beyond f16's range round to
// infinity and are kept, as their squares exceed any threshold.
__device__ __forceinline__ uint32_t drop_mask(h2 c, h2 t) {
    const h2 q = __builtin_elementwise_fma(c, c, -t);
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(s2, q) >> (short)15);
}
// A pair of coefficients as packed f16, zeroed where c^2 < t.
__device__ __forceinline__ uint32_t keep_pair(float c0, float c1, h2 t) {
    const h2 c = {(_Float16)c0, (_Float16)c1};
    return __builtin_bit_cast(uint32_t, c) & ~drop_mask(c, t);
}
__device__ __forceinline__ void threshold_pack(const f4 D[4], float t2, float tdc, u4 K[2], _Float16) {
    const h2 Tdc = {(_Float16)tdc, (_Float16)t2};
    const h2 T = {Tdc.y, Tdc.y};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        uint32_t t[4];
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const f4 &src = D[2 * h + (w >> 1)];
            const int r = 2 * (w & 1);
            t[w] = keep_pair(src[r], src[r + 1], (h == 0 && w == 0) ? Tdc : T);
        }
        K[h] = words(t);
    }
}
__device__ __forceinline__ uint32_t bf16_pair(float c0, float c1) {
    const bh2 v = {(__bf16)c0, (__bf16)c1};
    return __builtin_bit_cast(uint32_t, v);
}
// bf16 operands: the same test on the coefficients rounded to f16 (finer than bf16), applied to
// their bf16 pair.
__device__ __forceinline__ void threshold_pack(const f4 D[4], float t2, float tdc, u4 K[2], __bf16) {
    const h2 Tdc = {(_Float16)tdc, (_Float16)t2};
    const h2 T = {Tdc.y, Tdc.y};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        uint32_t t[4];
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const f4 &src = D[2 * h + (w >> 1)];
            const int r = 2 * (w & 1);
            const h2 c = {(_Float16)src[r], (_Float16)src[r + 1]};
            t[w] = bf16_pair(src[r], src[r + 1]) & ~drop_mask(c, (h == 0 && w == 0) ? Tdc : T);
        }
        K[h] = words(t);
    }
}

template <typename T>
struct Wave {
    u4 kf[4][2], ki[4][2];  // forward and inverse constants: M tile j, K-half h
    float t2a2;  // 2 t2a: the DC coefficient's factor in the threshold
    bool dc_lane;
    int bp_addr;  // 4 (l % 16): the DC broadcast's source lane
    T tag;

    // Forward transform of the 16 blocks whose K-halves are q0, q1 and the threshold: the kept
    // coefficients as the inverse's two K-halves. The squared threshold t2a max(mean, 0) + t2b
    // (block mean = 2 DC + the operands' offset above black) is max(2 t2a DC + b, lo) with
    // b = t2a off + t2b and lo = t2b, as t2a >= 0; a block outside the plane passes lo = +inf
    // and dc = +inf (everything dropped), one inside dc = -1 (its DC coefficient kept).
    __device__ __forceinline__ void forward(const u4 &q0, const u4 &q1, float b, float lo, float dc_t, u4 K[2]) const {
        f4 D[4];
        const f4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 4; j++) D[j] = mma(kf[j][1], q1, mma(kf[j][0], q0, z, tag), tag);
        const float dc = bcast_row0(D[0][0], bp_addr);
        // (block mean in DN: the DC coefficient is 8 x mean / 16)
        const float t2 = fmaxf(fmaf(t2a2, dc, b), lo);
        threshold_pack(D, t2, dc_lane ? dc_t : t2, K, tag);
    }
    // Inverse tile j of the coefficients K, accumulated onto acc.
    __device__ __forceinline__ f4 inv(int j, const u4 K[2], f4 acc) const {
        return mma(ki[j][1], K[1], mma(ki[j][0], K[0], acc, tag), tag);
    }
};

// A buffer resource over `bytes` bytes at `base` (wave-uniform): stores at offsets past the end
// are dropped by the hardware's range check.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *base, uint32_t bytes) {
    const uint64_t p = (uint64_t)(uintptr_t)base;
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(p >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)p);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)(((uint64_t)hi << 32) | (uint64_t)lo), (short)0,
                                             __builtin_amdgcn_readfirstlane((int)bytes), 0x00020000);
}
// The offset of a store that must not happen: past any resource (records stay below 2^31).
constexpr uint32_t NO_STORE = 0x80000000u;

// Whether the previous step's stores are pinned between the inverse chains by scheduling barriers:
// measured faster for f16 operands (46.4 -> 44.2 us on the 24 MP frame) and slower for bf16, whose
// threshold runs in f32 (46.9 -> 53.5 us), which keeps the compiler's own order.
template <typename T> constexpr bool kPinStores = false;
template <> constexpr bool kPinStores<_Float16> = true;

template <typename T, bool F32OUT>
__global__ __launch_bounds__(256) void k_dn_kron(DnkArgs A) {
    const int lane = threadIdx.x & 63;
    const int wave = (int)__builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (wave >= 2 * A.strips * A.segs) return;
    const int dy = wave & 1;
    const int rest = wave >> 1;
    const int strip = rest % A.strips, seg = rest / A.strips;
    const int x0 = strip * DNK_STRIP;
    const int y_first = seg * A.seg_steps * 8;
    const int nsteps = min(A.seg_steps, A.nsteps - seg * A.seg_steps);
    const int g = lane >> 4, n = lane & 15;
    const uint16_t *const src = A.src;
    uint16_t *const dst = A.dst;
    float *const dst_f = A.dst_f;
    const int pw = A.pw, ph = A.ph, plx = A.plx, ply = A.ply;
    const uint32_t stride = A.stride;
    const float black[2] = {dy ? A.black[2] : A.black[0], dy ? A.black[3] : A.black[1]};
    // outputs: the host keeps both below 2^31 bytes; a null output is an empty resource
    const uint64_t samples = (uint64_t)2 * ph * stride;
    const __amdgpu_buffer_rsrc_t rdst = rsrc(dst, dst ? (uint32_t)(samples * 2) : 0u);
    const __amdgpu_buffer_rsrc_t rdf = rsrc(dst_f, dst_f ? (uint32_t)(samples * 4) : 0u);

    Wave<T> W;
    {
        const u4 *tf = reinterpret_cast<const u4 *>(A.kf) + lane * 8;
        const u4 *ti = reinterpret_cast<const u4 *>(A.ki) + lane * 8;
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                W.kf[j][h] = tf[2 * j + h];
                W.ki[j][h] = ti[2 * j + h];
            }
    }
    W.t2a2 = 2.0f * A.t2a;
    const float inf = __builtin_inff();
    // a tile kind's threshold floor and DC threshold, inside the plane or not
    auto lo_of = [&](bool live) { return live ? A.t2b : inf; };
    auto dc_of = [&](bool live) { return live ? -1.0f : inf; };
    W.dc_lane = g == 0;
    W.bp_addr = 4 * n;

    // this lane's block columns: class (*, 0) at xb0, class (*, 1) at xb0 + 4
    const int xb0 = x0 + 8 * (n - 1);
    const bool ok0 = xb0 >= 0 && xb0 <= plx, ok1 = xb0 + 4 >= 0 && xb0 + 4 <= plx;
    const int ca = min(max(xb0, 0), pw - 4), cb = min(max(xb0 + 4, 0), pw - 4);

    // The raw samples of plane rows yb + g (rows clamped to the plane): two 16-byte loads per lane.
    struct Raw {
        u4 a, b;
    };
    // (one path for every row, clamped into the plane: no branch splits the step loop)
    auto issue = [&](int yb) {
        Raw r;
        const int rr = min(max(yb + g, 0), ph - 1);
        const uint16_t *p = src + (size_t)(2 * rr + dy) * stride;
        r.a = *reinterpret_cast<const u4 *>(p + 2 * ca);
        r.b = *reinterpret_cast<const u4 *>(p + 2 * cb);
        return r;
    };
    // Both planes' class-(*, 0) samples of a row.
    auto planes = [&](const Raw &r, u4 w[2]) {
        const uint32_t raw[8] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w};
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const uint32_t sel = p ? 0x07060302u : 0x05040100u;
            uint32_t t[4];
#pragma unroll
            for (int k = 0; k < 4; k++) t[k] = __builtin_amdgcn_perm(raw[2 * k + 1], raw[2 * k], sel);
            w[p] = words(t);
        }
    };
    auto convert = [&](const u4 &v, uint32_t off2) {
        uint32_t t[4];
#pragma unroll
        for (int k = 0; k < 4; k++) t[k] = int_pair(sub2(v[k], off2), T());
        return words(t);
    };
    // the class-(*, 1) operand: this block's right half and column n + 1's left half
    auto partner = [&](const u4 &c0) {
        const uint32_t t[4] = {c0[2], c0[3], shl1(c0[0]), shl1(c0[1])};
        return words(t);
    };
    // Per-plane offsets. Step s (block rows y0 = y_first + 8 s and y0 + 4) converts its operands
    // relative to one sample of its own rows, plane row min(y0 + 4, ph - 1) at plane column
    // clamp(x0 + 56, 0, pw - 4) (lane 8's first word of row y0 + 4), so an offset depends only on
    // the block position and a result does not depend on how the strip is cut into segments.
    auto step_offset = [&](const u4 w[2], uint32_t off2[2], float offf[2]) {
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const uint32_t s = (uint32_t)__builtin_amdgcn_readlane(w[p][0], 8) & 0xffffu;
            off2[p] = s | (s << 16);
            offf[p] = (float)s - black[p];  // in DN above black, as the outputs
        }
    };

    // samples of rows y0 + g (both planes), carried down the strip
    u4 wA[2];
    // the class-(1, *) sums of rows 4..7 of the previous block row (registers j = 1, 3), their
    // offsets, and whether those blocks lie inside the plane
    f4 C[2][2];
    float offC[2] = {0.f, 0.f};
    bool c_ok = false;
    const f4 z = {0.f, 0.f, 0.f, 0.f};
    planes(issue(y_first), wA);
    {
        // the halo step: classes (1, *) at y_first - 4 (the previous step's lower blocks), for the
        // carry, with that step's offsets
        const int yh = y_first - 4;
        for (int p = 0; p < 2; p++) C[p][0] = C[p][1] = z;
        if (yh >= 0 && yh <= ply) {
            u4 hw[2];
            uint32_t off2[2];
            planes(issue(yh), hw);
            step_offset(hw, off2, offC);
            for (int p = 0; p < 2; p++) {
                const u4 c0 = convert(hw[p], off2[p]), c1 = convert(wA[p], off2[p]);
                u4 K0[2], K1[2];
                const float b = fmaf(A.t2a, offC[p], A.t2b);
                W.forward(c0, c1, b, lo_of(ok0), dc_of(ok0), K0);
                W.forward(partner(c0), partner(c1), b, lo_of(ok1), dc_of(ok1), K1);
                C[p][0] = W.inv(1, K0, z);
                const f4 s2 = W.inv(3, K1, z);
#pragma unroll
                for (int r = 0; r < 4; r++) C[p][0][r] += shr1(s2[r]);
                C[p][1] = W.inv(1, K1, W.inv(3, K0, z));
            }
            c_ok = true;
        }
    }

    // Stores. Pixel (row, col) of output register j is (y0 + 4 (j & 1) + g, xb0 + 4 (j >> 1) + r),
    // both planes interleaved; where fewer than four blocks cover a pixel (the plane's border rows
    // and columns) its sum is scaled by the cover factor 2 per short direction. The u16 sample is
    // sum * f + offset + black, rounded to nearest and clamped to [0, 65535]: one fma in units of
    // 1/65535 and the hardware's packed unorm16 conversion for two samples (clamp to [0, 1],
    // scale, round); the f32 output (F32OUT) is sum * f + offset. Stores of pixels a lane does
    // not own address past the resource. A store's byte offset is the saturating sum of this
    // lane's part (its row g in the group, its column group) and a part uniform over the wave (the
    // group's first row: rows come in aligned groups of four, as do the plane's height and its
    // border rows); NO_STORE in either part takes the sum past the resource.
    constexpr float kUnorm = 1.0f / 65535.0f;
    float cf[2][4], cfu[2][4];  // column cover factors of this lane's two column groups (and / 65535)
    uint32_t lane16[2], lane32[2];  // this lane's part of the u16 and f32 byte offsets
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int col = xb0 + 4 * h + r;
            cf[h][r] = (col < 4 || col > plx + 3) ? 2.0f : 1.0f;
            cfu[h][r] = cf[h][r] * kUnorm;
        }
        const int col = xb0 + 4 * h;
        // (bitwise, not short-circuit: a select, not a branch)
        const bool owned = (n != 0) & !((n == 15) & (h == 1)) & (col < pw);
        const uint32_t o = (uint32_t)((2 * g + dy) * (int)stride + 2 * col);  // sample index past the group's row
        lane16[h] = owned ? o * 2u : NO_STORE;
        lane32[h] = owned ? o * 4u : NO_STORE;
    }
    auto store_j = [&](int j, const f4 O[2][4], const float ot[2], const float om[2], int y0) {
        {
            const int rowg = y0 + 4 * (j & 1);  // the group's first row (wave-uniform)
            const uint32_t rs = (uint32_t)(2 * rowg) * stride;  // its sample index
            const uint32_t o16 = __builtin_elementwise_add_sat(lane16[j >> 1], rowg < ph ? rs * 2u : NO_STORE);
            const float o0 = (j & 1) ? om[0] : ot[0], o1 = (j & 1) ? om[1] : ot[1];
            const float c0 = (o0 + black[0]) * kUnorm, c1 = (o1 + black[1]) * kUnorm;
            const float fy = (rowg < 4 || rowg >= ph - 4) ? 2.0f : 1.0f;
            uint32_t q[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float f = fy * cfu[j >> 1][r];
                q[r] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pknorm_u16(fmaf(O[0][j][r], f, c0),
                                                                                   fmaf(O[1][j][r], f, c1)));
            }
            __builtin_amdgcn_raw_buffer_store_b128(words(q), rdst, o16, 0, 0);
            if (F32OUT) {
                float v0[4], v1[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float f = fy * cf[j >> 1][r];
                    v0[r] = fmaf(O[0][j][r], f, o0);
                    v1[r] = fmaf(O[1][j][r], f, o1);
                }
                const uint32_t of = __builtin_elementwise_add_sat(lane32[j >> 1], rowg < ph ? rs * 4u : NO_STORE);
                const uint32_t a0[4] = {__builtin_bit_cast(uint32_t, v0[0]), __builtin_bit_cast(uint32_t, v1[0]),
                                        __builtin_bit_cast(uint32_t, v0[1]), __builtin_bit_cast(uint32_t, v1[1])};
                const uint32_t a1[4] = {__builtin_bit_cast(uint32_t, v0[2]), __builtin_bit_cast(uint32_t, v1[2]),
                                        __builtin_bit_cast(uint32_t, v0[3]), __builtin_bit_cast(uint32_t, v1[3])};
                __builtin_amdgcn_raw_buffer_store_b128(words(a0), rdf, of, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b128(words(a1), rdf, __builtin_elementwise_add_sat(of, 16u), 0, 0);
            }
        }
    };
    auto store_step = [&](const f4 O[2][4], const float ot[2], const float om[2], int y0) {
#pragma unroll
        for (int j = 0; j < 4; j++) store_j(j, O, ot, om, y0);
    };

    // The previous step's outputs and offsets: stored while this step's products run (the loop
    // body has no branch, so the compiler interleaves the stores' vector work with the MFMAs).
    f4 P[2][4];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int j = 0; j < 4; j++) P[p][j] = z;
    float pt[2] = {0.f, 0.f}, pm[2] = {0.f, 0.f};
    // rows y0 + 4 + g and y0 + 8 + g of the next step, loaded one step ahead
    Raw nB = issue(y_first + 4), nC = issue(y_first + 8);
    for (int s = 0; s < nsteps; s++) {
        const int y0 = y_first + 8 * s;
        const bool top_ok = y0 <= ply, mid_ok = y0 + 4 <= ply;
        // blocks inside the plane (the others, columns at the strip ends and rows at the bottom, keep nothing)
        const bool kt0 = ok0 && top_ok, kt1 = ok1 && top_ok;
        const bool km0 = ok0 && mid_ok, km1 = ok1 && mid_ok;
        const float lo_t0 = lo_of(kt0), lo_t1 = lo_of(kt1), lo_m0 = lo_of(km0), lo_m1 = lo_of(km1);
        const float dc_t0 = dc_of(kt0), dc_t1 = dc_of(kt1), dc_m0 = dc_of(km0), dc_m1 = dc_of(km1);
        u4 wB[2], wC[2];
        planes(nB, wB);
        planes(nC, wC);
        // (after the last step: this step's rows again, already in the caches, never used)
        const bool more = s + 1 < nsteps;
        nB = issue(more ? y0 + 12 : y0 + 4);
        nC = issue(more ? y0 + 16 : y0 + 8);
        uint32_t off2[2];
        float offf[2];
        step_offset(wB, off2, offf);
        // the offset term of an output: the mean offset of the blocks that cover it
        float off_top[2];
        f4 out[2][4];
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const u4 cA = convert(wA[p], off2[p]), cB = convert(wB[p], off2[p]), cC = convert(wC[p], off2[p]);
            // classes (0, 0), (0, 1) at block row y0 and (1, 0), (1, 1) at y0 + 4
            u4 Kt0[2], Kt1[2], Km0[2], Km1[2];
            const float b = fmaf(A.t2a, offf[p], A.t2b);
            W.forward(cA, cB, b, lo_t0, dc_t0, Kt0);
            W.forward(partner(cA), partner(cB), b, lo_t1, dc_t1, Kt1);
            W.forward(cB, cC, b, lo_m0, dc_m0, Km0);
            W.forward(partner(cB), partner(cC), b, lo_m1, dc_m1, Km1);
            // The overlap-add, accumulated in the inverse products: rows 0..3 (j = 0, 2) of this
            // block row take the previous step's lower blocks (the carry), rows 4..7 (j = 1, 3)
            // this step's; class-(*, 1) blocks add in register j ^ 2, from column n - 1 for j < 2.
            // Between the first plane's product chains go the previous step's stores, one output
            // register each (independent vector work under the MFMAs; the first step has none:
            // its stores address past the resource).
            const int ys = s > 0 ? y0 - 8 : ph;
            const f4 o0 = W.inv(0, Kt0, C[p][0]);
            const f4 s0 = W.inv(2, Kt1, z);
            if (p == 0) {
                store_j(0, P, pt, pm, ys);
                if (kPinStores<T>) __builtin_amdgcn_sched_barrier(0);  // keep the stores with these products
            }
            const f4 o1 = W.inv(0, Km0, W.inv(1, Kt0, z));
            const f4 s1 = W.inv(2, Km1, W.inv(3, Kt1, z));
            if (p == 0) {
                store_j(1, P, pt, pm, ys);
                if (kPinStores<T>) __builtin_amdgcn_sched_barrier(0);  // keep the stores with these products
            }
            out[p][2] = W.inv(0, Kt1, W.inv(2, Kt0, C[p][1]));
            out[p][3] = W.inv(0, Km1, W.inv(2, Km0, W.inv(1, Kt1, W.inv(3, Kt0, z))));
            if (p == 0) {
                store_j(2, P, pt, pm, ys);
                if (kPinStores<T>) __builtin_amdgcn_sched_barrier(0);  // keep the stores with these products
            }
            const f4 c0 = W.inv(1, Km0, z);
            const f4 s2 = W.inv(3, Km1, z);
            C[p][1] = W.inv(1, Km1, W.inv(3, Km0, z));
            if (p == 0) {
                store_j(3, P, pt, pm, ys);
                if (kPinStores<T>) __builtin_amdgcn_sched_barrier(0);  // keep the stores with these products
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                out[p][0][r] = o0[r] + shr1(s0[r]);
                out[p][1][r] = o1[r] + shr1(s1[r]);
                C[p][0][r] = c0[r] + shr1(s2[r]);
            }
            off_top[p] = top_ok ? (c_ok ? (offf[p] + offC[p]) * 0.5f : offf[p]) : offC[p];
            offC[p] = offf[p];
            wA[p] = wC[p];
        }
        c_ok = mid_ok;
#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int j = 0; j < 4; j++) P[p][j] = out[p][j];
            pt[p] = off_top[p];
            pm[p] = offf[p];
        }
    }
    if (nsteps > 0) store_step(P, pt, pm, y_first + 8 * (nsteps - 1));
}

// f32 -> f16 / bf16 bits, round to nearest even (host).
uint16_t to_bits(float x, bool bf16) {
    if (bf16) {
        uint32_t u;
        std::memcpy(&u, &x, 4);
        if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    }
    const _Float16 h = (_Float16)x;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

}  // namespace

// Per lane, the operand words of the forward (coefficient rows x pixels, scaled by 1/16) and
// inverse (pixels x coefficients, scaled by 16 / 4: the overlap-add's uniform weight of the four
// covering blocks) Kronecker matrices in the layout k_dn_kron reads: lane l, M tile j, K-half h,
// element i at ((l * 4 + j) * 2 + h) * 8 + i.
void dn_kron_tables(const float d[64], bool bf16, std::vector<uint16_t> &kf, std::vector<uint16_t> &ki) {
    kf.assign(64 * 64, 0);
    ki.assign(64 * 64, 0);
    for (int l = 0; l < 64; l++) {
        const int g = l >> 4, n = l & 15;
        for (int j = 0; j < 4; j++)
            for (int h = 0; h < 2; h++)
                for (int i = 0; i < 8; i++) {
                    const size_t at = (((size_t)l * 4 + j) * 2 + h) * 8 + i;
                    // forward: coefficient 16 j + n = (u, v); pixel (4 h + g, i)
                    {
                        const int m = 16 * j + n, u = m >> 3, v = m & 7;
                        const double val = (double)d[u * 8 + 4 * h + g] * (double)d[v * 8 + i] / 16.0;
                        kf[at] = to_bits((float)val, bf16);
                    }
                    // inverse: pixel of output row 16 j + n; coefficient of K slot (g, i) of K-half h
                    {
                        const int gp = n >> 2, rp = n & 3;
                        const int prow = 4 * (j & 1) + gp, pcol = 4 * (j >> 1) + rp;
                        const int m = i < 4 ? 16 * (2 * h) + 4 * g + i : 16 * (2 * h + 1) + 4 * g + (i - 4);
                        const int u = m >> 3, v = m & 7;
                        const double val = (double)d[u * 8 + prow] * (double)d[v * 8 + pcol] * 4.0;
                        ki[at] = to_bits((float)val, bf16);
                    }
                }
    }
}

// The kernel holds two waves per SIMD (its registers), and a wave's time is close to fixed per
// step whether it shares its SIMD or not, so the launch is sized to one round of at most two
// waves per SIMD with equal step counts: the fewest steps per wave whose 2 x strips x segments
// waves fit, but at least 4 steps (each segment repeats half a step of halo).
int dn_kron_seg_steps(int strips, int nsteps, int num_cus) {
    const int slots = 2 * 4 * (num_cus > 0 ? num_cus : 256);
    const int per_round = std::max(1, slots / (2 * strips));  // segments per round
    return std::max(4, (nsteps + per_round - 1) / per_round);
}

hipError_t launch_dn_kron(const DnkArgs &a, bool bf16, hipStream_t st) {
    const uint32_t waves = 2u * (uint32_t)a.strips * (uint32_t)a.segs;
    const uint32_t grid = (waves + 3) / 4;
    const bool f32out = a.dst_f != nullptr;
    if (bf16 && f32out)
        hipLaunchKernelGGL((k_dn_kron<__bf16, true>), dim3(grid), dim3(256), 0, st, a);
    else if (bf16)
        hipLaunchKernelGGL((k_dn_kron<__bf16, false>), dim3(grid), dim3(256), 0, st, a);
    else if (f32out)
        hipLaunchKernelGGL((k_dn_kron<_Float16, true>), dim3(grid), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_dn_kron<_Float16, false>), dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace frh
