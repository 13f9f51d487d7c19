// av1pf_cdef.hip - CDEF (AV1 7.15) for gfx950.
//
// Reference semantics (matteopolak/vaco, crates/codec/vaco-codec-av1/src/):
// find_direction cdef.rs:79-124, constrain 126-133, filter_block 142-205,
// adjust_strength 208-220, chroma_direction 222-230, frame walk
// decode.rs:739-861 with cdef_block_skipped 892-896.
//
// One workgroup of 1024 or 512 threads (16 or 8 wave64) per 64x64 luma
// superblock and its co-located chroma, in XCD-contiguous order. The
// superblock plus a 2-sample border is staged in LDS once (8-byte loads
// inside the plane), with a copy shifted by one sample so that every pair of
// samples is one aligned 32-bit LDS read. The direction search keeps each block's 15 line
// sums in registers (one wave per direction, one lane per 8x8 block,
// 16-byte LDS row loads); one wave turns each block's 8 costs into its
// luma and chroma parameters. Then a half-wave filters each 8x8 luma block,
// two samples per lane in packed 16-bit arithmetic, with the block's
// parameters held in registers. Superblocks that touch a plane edge add one
// bit select per tap for unavailable taps.
#include "av1pf_cdef_body.h"

namespace av1pf {

// One workgroup per superblock (cdef_sb), over every frame of the launch.
template <int SX, int SY, bool XCD, int T, bool ARG>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_cdef(
    const CdefFrame *__restrict__ frames, int n_frames, const CdefFrame single, const uint4 *__restrict__ up_src,
    uint4 *__restrict__ up_dst, uint32_t up_n16)
{
    __shared__ CdefLds<SX, SY> S;
    cdef_sb<SX, SY, XCD, T, ARG, false>(S, frames, n_frames, single, up_src, up_dst, up_n16, blockIdx.x, gridDim.x,
                                        ChainSync{});
}

// Kernel-level cases (cdef::find_direction + cdef::filter_block on one
// 12x12 neighbourhood each, av1_cdef.rs:85-103): one wave per case, through
// the frame kernel's own device code: dir_cost for the search, and the
// packed filter_pairs with its tile layout (shifted copy, pre-resolved tap
// offsets), its unavailable-tap select and its clamp skip for filtering.
struct CdefCase {
    uint32_t width, height;
    uint8_t edges, bit_depth, direction, damping;
    uint16_t primary, secondary;
};

// The case tile: 12 rows of kRow samples (column c of the neighbourhood at
// index c + 6, so the block starts at the even index 8), then its copy
// shifted by one sample.
constexpr int kCaseTile = 12 * kRow;

__global__ __launch_bounds__(64) void k_cdef_cases(const uint16_t *__restrict__ inputs,
                                                   const CdefCase *__restrict__ params,
                                                   uint8_t *__restrict__ dirs,
                                                   uint32_t *__restrict__ vars,
                                                   uint16_t *__restrict__ outs, uint32_t n)
{
    __shared__ __align__(16) uint16_t t[2 * kCaseTile + 8];
    __shared__ __align__(16) uint16_t raw[8 * kRow];
    __shared__ uint32_t cost[8];
    const uint32_t cs = blockIdx.x;
    if (cs >= n)
        return;
    const int lane = threadIdx.x;
    const CdefCase P = params[cs];
    const int shift = P.bit_depth - 8;
    const uint16_t *in = inputs + (size_t)cs * 144;
    // Taps outside the edge mask become unavailable, exactly like plane-exterior
    // samples in k_cdef (cdef.rs:161-166, 189-191).
    for (int e = lane; e < 144; e += 64) {
        const int r = e / 12, c = e % 12;
        const int ny = r - 2, nx = c - 2;
        const bool ok = !((nx < 0 && !(P.edges & 1)) || (nx >= (int)P.width && !(P.edges & 2)) ||
                          (ny < 0 && !(P.edges & 4)) || (ny >= (int)P.height && !(P.edges & 8)));
        const uint16_t v = ok ? in[e] : kUnavailable;
        t[r * kRow + c + 6] = v;
        t[kCaseTile + r * kRow + c + 7] = v;
    }
    // The direction search reads the raw 8x8 at (2,2) (av1_cdef.rs:90-95).
    raw[(lane >> 3) * kRow + (lane & 7)] = in[((lane >> 3) + 2) * 12 + (lane & 7) + 2];
    __syncthreads();
    if (lane < 8)
        cost[lane] = dir_cost_any(lane, raw, shift);
    __syncthreads();
    if (lane == 0) {
        int best = 0;
        uint32_t bc = cost[0];
        for (int d = 1; d < 8; d++)
            if (cost[d] > bc) {
                bc = cost[d];
                best = d;
            }
        dirs[cs] = (uint8_t)best;
        vars[cs] = (bc - cost[(best + 4) & 7]) >> 10;
    }
    // Lanes 0-31 each filter one pair of samples (row lane / 4, columns
    // 2 * (lane % 4) and the next). With every edge available no tap is
    // unavailable, so those cases run the interior superblocks' variant.
    if (lane < 32) {
        const Blk B = make_blk<kRow, kCaseTile>(P.direction, P.primary, P.secondary, P.damping, shift);
        const int r = lane >> 2, c = (lane & 3) * 2;
        const int ci = (r + 2) * kRow + c + 8;
        const bool pri_on = P.primary != 0, sec_on = P.secondary != 0;
        s16x2 res;
        if (P.edges == 15)
            filter_pairs<false, 1, 0>(t, ci, B, &res, pri_on, sec_on);
        else
            filter_pairs<true, 1, 0>(t, ci, B, &res, pri_on, sec_on);
        const bool in_block = r < (int)P.height && c < (int)P.width;
        uint16_t *o = outs + (size_t)cs * 64 + r * 8 + c;
        o[0] = in_block ? (uint16_t)res.x : 0;
        o[1] = in_block ? (uint16_t)res.y : 0;  // widths are 4 or 8: both samples or neither
    }
}

template <bool XCD, int T, bool ARG>
static void launch_cdef_mode(const CdefFrame *frames, int n, const CdefFrame &single, uint32_t blocks, int mode,
                             const uint4 *up_src, uint4 *up_dst, uint32_t up_n16, hipStream_t s, hipEvent_t ev0,
                             hipEvent_t ev1)
{
    const dim3 g(blocks), b(T);
    switch (mode) {
    case 0:
        hipExtLaunchKernelGGL((k_cdef<1, 1, XCD, T, ARG>), g, b, 0, s, ev0, ev1, 0, frames, n, single, up_src,
                              up_dst, up_n16);
        break;
    case 1:
        hipExtLaunchKernelGGL((k_cdef<0, 0, XCD, T, ARG>), g, b, 0, s, ev0, ev1, 0, frames, n, single, up_src,
                              up_dst, up_n16);
        break;
    case 2:
        hipExtLaunchKernelGGL((k_cdef<1, 0, XCD, T, ARG>), g, b, 0, s, ev0, ev1, 0, frames, n, single, up_src,
                              up_dst, up_n16);
        break;
    default:
        hipExtLaunchKernelGGL((k_cdef<0, 1, XCD, T, ARG>), g, b, 0, s, ev0, ev1, 0, frames, n, single, up_src,
                              up_dst, up_n16);
        break;
    }
}

template <int T, bool ARG>
static void launch_cdef_t(const CdefFrame *frames, int n, const CdefFrame &single, uint32_t blocks, int mode,
                          bool xcd, const uint4 *us, uint4 *ud, uint32_t n16, hipStream_t s, hipEvent_t ev0,
                          hipEvent_t ev1)
{
    if (xcd)
        launch_cdef_mode<true, T, ARG>(frames, n, single, blocks, mode, us, ud, n16, s, ev0, ev1);
    else
        launch_cdef_mode<false, T, ARG>(frames, n, single, blocks, mode, us, ud, n16, s, ev0, ev1);
}

// mode: 0 = 4:2:0, 1 = 4:4:4 or monochrome, 2 = 4:2:2, 3 = 4:4:0. xcd:
// XCD-contiguous superblock order. waves: 16 or 8 waves (1024 or 512
// threads) per superblock's workgroup. ev0/ev1 (may be null) receive the
// dispatch's start and end.
// single (may be null): the one frame's descriptor, passed by value in the
// kernel arguments (frames and n are then unused); up_src (the slot's host
// memory, at least 16 bytes, read even when up_bytes is 0) and up_dst are
// then 16-byte aligned: see cdef_sb. Without single, up_* are unused.
hipError_t launch_cdef(const CdefFrame *frames, int n, const CdefFrame *single, uint32_t blocks, int mode,
                       bool xcd, int waves, const void *up_src, void *up_dst, size_t up_bytes, hipStream_t s,
                       hipEvent_t ev0, hipEvent_t ev1)
{
    if (blocks == 0)
        return hipSuccess;
    const uint4 *us = static_cast<const uint4 *>(up_src);
    uint4 *ud = static_cast<uint4 *>(up_dst);
    const uint32_t n16 = (uint32_t)((up_bytes + 15) / 16);
    if (single && !up_src)
        return hipErrorInvalidValue;  // the single-frame kernels read (at least) unit 0 of the staged bytes
    const CdefFrame none{};
    const CdefFrame &one = single ? *single : none;
    if (waves == 8) {
        if (single)
            launch_cdef_t<512, true>(frames, n, one, blocks, mode, xcd, us, ud, n16, s, ev0, ev1);
        else
            launch_cdef_t<512, false>(frames, n, one, blocks, mode, xcd, us, ud, n16, s, ev0, ev1);
    } else {
        if (single)
            launch_cdef_t<1024, true>(frames, n, one, blocks, mode, xcd, us, ud, n16, s, ev0, ev1);
        else
            launch_cdef_t<1024, false>(frames, n, one, blocks, mode, xcd, us, ud, n16, s, ev0, ev1);
    }
    return hipGetLastError();
}

hipError_t launch_cdef_cases(const uint16_t *inputs, const void *params, uint8_t *dirs,
                             uint32_t *vars, uint16_t *outs, uint32_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cdef_cases, dim3(n), dim3(64), 0, s, inputs, (const CdefCase *)params,
                       dirs, vars, outs, n);
    return hipGetLastError();
}

}  // namespace av1pf
