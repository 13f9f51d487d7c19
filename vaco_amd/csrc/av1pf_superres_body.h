// av1pf_superres_body.h - the superres work item (AV1 7.16) as a device
// function over caller-provided LDS (k_superres in av1pf_superres.hip).
//
// Reference semantics (matteopolak/vaco, crates/codec/vaco-codec-av1/src/
// superres.rs): subpel_start 145-152, upscale_plane 163-216.
#pragma once

#include "av1pf_common.h"

namespace av1pf {

// UPSCALE_FILTER, superres.rs:25-90 (AV1 spec Upscale_Filter[64][8]).
static __constant__ int16_t kUpscale[64][8] = {
    {0, 0, 0, 128, 0, 0, 0, 0},       {0, 0, -1, 128, 2, -1, 0, 0},
    {0, 1, -3, 127, 4, -2, 1, 0},     {0, 1, -4, 127, 6, -3, 1, 0},
    {0, 2, -6, 126, 8, -3, 1, 0},     {0, 2, -7, 125, 11, -4, 1, 0},
    {-1, 2, -8, 125, 13, -5, 2, 0},   {-1, 3, -9, 124, 15, -6, 2, 0},
    {-1, 3, -10, 123, 18, -6, 2, -1}, {-1, 3, -11, 122, 20, -7, 3, -1},
    {-1, 4, -12, 121, 22, -8, 3, -1}, {-1, 4, -13, 120, 25, -9, 3, -1},
    {-1, 4, -14, 118, 28, -9, 3, -1}, {-1, 4, -15, 117, 30, -10, 4, -1},
    {-1, 5, -16, 116, 32, -11, 4, -1}, {-1, 5, -16, 114, 35, -12, 4, -1},
    {-1, 5, -17, 112, 38, -12, 4, -1}, {-1, 5, -18, 111, 40, -13, 5, -1},
    {-1, 5, -18, 109, 43, -14, 5, -1}, {-1, 6, -19, 107, 45, -14, 5, -1},
    {-1, 6, -19, 105, 48, -15, 5, -1}, {-1, 6, -19, 103, 51, -16, 5, -1},
    {-1, 6, -20, 101, 53, -16, 6, -1}, {-1, 6, -20, 99, 56, -17, 6, -1},
    {-1, 6, -20, 97, 58, -17, 6, -1}, {-1, 6, -20, 95, 61, -18, 6, -1},
    {-2, 7, -20, 93, 64, -18, 6, -2}, {-2, 7, -20, 91, 66, -19, 6, -1},
    {-2, 7, -20, 88, 69, -19, 6, -1}, {-2, 7, -20, 86, 71, -19, 6, -1},
    {-2, 7, -20, 84, 74, -20, 7, -2}, {-2, 7, -20, 81, 76, -20, 7, -1},
    {-2, 7, -20, 79, 79, -20, 7, -2}, {-1, 7, -20, 76, 81, -20, 7, -2},
    {-2, 7, -20, 74, 84, -20, 7, -2}, {-1, 6, -19, 71, 86, -20, 7, -2},
    {-1, 6, -19, 69, 88, -20, 7, -2}, {-1, 6, -19, 66, 91, -20, 7, -2},
    {-2, 6, -18, 64, 93, -20, 7, -2}, {-1, 6, -18, 61, 95, -20, 6, -1},
    {-1, 6, -17, 58, 97, -20, 6, -1}, {-1, 6, -17, 56, 99, -20, 6, -1},
    {-1, 6, -16, 53, 101, -20, 6, -1}, {-1, 5, -16, 51, 103, -19, 6, -1},
    {-1, 5, -15, 48, 105, -19, 6, -1}, {-1, 5, -14, 45, 107, -19, 6, -1},
    {-1, 5, -14, 43, 109, -18, 5, -1}, {-1, 5, -13, 40, 111, -18, 5, -1},
    {-1, 4, -12, 38, 112, -17, 5, -1}, {-1, 4, -12, 35, 114, -16, 5, -1},
    {-1, 4, -11, 32, 116, -16, 5, -1}, {-1, 4, -10, 30, 117, -15, 4, -1},
    {-1, 3, -9, 28, 118, -14, 4, -1}, {-1, 3, -9, 25, 120, -13, 4, -1},
    {-1, 3, -8, 22, 121, -12, 4, -1}, {-1, 3, -7, 20, 122, -11, 3, -1},
    {-1, 2, -6, 18, 123, -10, 3, -1}, {0, 2, -6, 15, 124, -9, 3, -1},
    {0, 2, -5, 13, 125, -8, 2, -1},   {0, 1, -4, 11, 125, -7, 2, 0},
    {0, 1, -3, 8, 126, -6, 2, 0},     {0, 1, -3, 6, 127, -4, 1, 0},
    {0, 1, -2, 4, 127, -3, 1, 0},     {0, 0, -1, 2, 128, -1, 0, 0},
};

constexpr int kSrThreads = 256;
constexpr int kSrPer = kSrSeg / kSrThreads;  // output columns per thread
constexpr int kSrWin = kSrSeg + 16;          // staged source samples per row

// One segment of kSrSeg output columns over up to R rows
// (superres.rs:173-214). Positions, phases and filters depend on the column
// only, so each thread derives them once for its columns and applies them to
// every row. Each row's source window (the union of the segment's taps,
// edge-clamped to the padded width) is staged in LDS; because the step is
// below one source sample the window is exactly the samples the taps read,
// so the sample-range flag is raised exactly when the reference would reject
// a tap.
//
// Rows are staged kSrG at a time into one of two window buffers: while the
// workgroup filters one group, the loads of the next are in flight in
// registers, and they are stored into the other buffer afterwards, so the
// source reads overlap the arithmetic. Each buffer holds win, each row's
// window, and wsh, the same shifted by one sample (window index i at i + 1),
// so the eight taps from any position are four aligned sample pairs of one
// of the two copies.
//
// LDS (sr_lds_bytes<R>() at a 16-byte aligned `lds`): the filter table, then
// the buffers. `local` is the work item's block index within job j. Runs on
// kSrThreads threads.
constexpr int kSrG = 2;  // rows per staged group
constexpr int kSrWinBytes = kSrG * kSrWin * 2;
constexpr int kSrGroupBytes = (kSrWinBytes + kSrG * (kSrWin + 2) * 2 + 15) & ~15;

template <int R>
constexpr int sr_lds_bytes()
{
    return 64 * 8 * 2 + (R > kSrG ? 2 : 1) * kSrGroupBytes;
}

template <int R>
__device__ __forceinline__ void sr_block(const SrJob &j, uint32_t local, uint8_t *lds)
{
    static_assert(R % kSrG == 0, "whole row groups");
    static_assert(kSrWinBytes % 16 == 0 && kSrGroupBytes % 16 == 0, "buffers stay 16-byte aligned");
    int16_t *tab = reinterpret_cast<int16_t *>(lds);
    auto win = [&](int buf) {
        return reinterpret_cast<uint16_t(*)[kSrWin]>(lds + 64 * 8 * 2 + buf * kSrGroupBytes);
    };
    auto wsh = [&](int buf) {
        return reinterpret_cast<uint16_t(*)[kSrWin + 2]>(lds + 64 * 8 * 2 + buf * kSrGroupBytes + kSrWinBytes);
    };
    const uint32_t rb = local / j.segs;
    const uint32_t seg = local - rb * j.segs;
    const int r0 = (int)rb * R;
    const int nr = min(R, (int)j.n_rows - r0);
    const int ngr = (nr + kSrG - 1) / kSrG;  // row groups of this workgroup
    const int tid = (int)threadIdx.x;
    const int x0 = (int)seg * kSrSeg;
    const int xn = min(x0 + kSrSeg, (int)j.out_width) - 1;
    const int lo = ((-(1 << 14) + j.start + x0 * j.step) >> 14) - 3;
    const int hi = ((-(1 << 14) + j.start + xn * j.step) >> 14) + 4;
    const int last = (int)j.in_width - 1;
    // Window origin: lo, or lo rounded down to a multiple of 4 where the
    // window lies inside the row with room for whole 4-sample groups, which
    // are then loaded 8 bytes at a time. Only samples lo .. hi are checked
    // against the bit depth (the taps read no others).
    const int groups = (hi - (lo & ~3)) / 4 + 1;
    const bool vec = j.vec_ok && lo >= 0 && (lo & ~3) + 4 * groups - 1 <= last;
    const int org = vec ? (lo & ~3) : lo;
    const int maxv = j.max_sample;
    bool bad = false;
    // plane row of the workgroup's row r (r < nr)
    auto yrow = [&](int r) -> uint32_t { return j.rows ? glb(j.rows)[r0 + r] : (uint32_t)(r0 + r); };
    gptr<uint16_t> out16 = glb(j.out);
    gptr<uint8_t> out8 = reinterpret_cast<gptr<uint8_t>>(out16);
    const bool u8 = j.out_u8 != 0;

    // Per column of this thread (filled once the table is in LDS): the first
    // tap's window index and the taps as pairs, (0,1) (2,3) (4,5) (6,7), for
    // v_dot2 over sample pairs.
    int cbase[kSrPer];
    s16x2 cf[kSrPer][4];
    auto setup_columns = [&]() {
#pragma unroll
        for (int k = 0; k < kSrPer; k++) {
            const int x = x0 + tid + k * kSrThreads;
            const int src_x = -(1 << 14) + j.start + x * j.step;
            cbase[k] = (src_x >> 14) - 3 - org;
            const uint4 fq = *reinterpret_cast<const uint4 *>(&tab[((src_x & 0x3fff) >> 8) * 8]);
            cf[k][0] = __builtin_bit_cast(s16x2, fq.x);
            cf[k][1] = __builtin_bit_cast(s16x2, fq.y);
            cf[k][2] = __builtin_bit_cast(s16x2, fq.z);
            cf[k][3] = __builtin_bit_cast(s16x2, fq.w);
        }
    };
    // Filter the staged row group g (rows g * kSrG ..) from buffer buf.
    auto filter_group = [&](int g, int buf) {
        uint16_t(*w0)[kSrWin] = win(buf);
        uint16_t(*w1)[kSrWin + 2] = wsh(buf);
#pragma unroll
        for (int rr = 0; rr < kSrG; rr++) {
            const int r = g * kSrG + rr;
            if (r >= nr)
                break;
            const size_t orow = (size_t)yrow(r) * j.out_stride;
#pragma unroll
            for (int k = 0; k < kSrPer; k++) {
                const int x = x0 + tid + k * kSrThreads;
                if (x > xn)
                    break;
                const int base = cbase[k];
                const uint32_t *w =
                    reinterpret_cast<const uint32_t *>((base & 1) ? &w1[rr][base + 1] : &w0[rr][base]);
                int sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[0]), cf[k][0], 0, false);
                sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[1]), cf[k][1], sum, false);
                sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[2]), cf[k][2], sum, false);
                sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[3]), cf[k][3], sum, false);
                const int v = min(max((sum + 64) >> 7, 0), maxv);
                if (u8)
                    out8[orow + x] = (uint8_t)v;
                else
                    out16[orow + x] = (uint16_t)v;
            }
        }
    };

    for (int i = tid; i < 64 * 8; i += kSrThreads)
        tab[i] = kUpscale[i >> 3][i & 7];
    if (vec) {
        // At most two 4-sample groups per thread and row; a row group's loads
        // are all issued before any is used.
        uint64_t q[kSrG][2];
        auto load = [&](int g) {
#pragma unroll
            for (int rr = 0; rr < kSrG; rr++) {
                const int r = g * kSrG + rr;
                gptr<const uint64_t> row = reinterpret_cast<gptr<const uint64_t>>(
                    glb(j.in) + (size_t)yrow(min(r, nr - 1)) * j.in_stride + org);
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int gi = tid + k * kSrThreads;
                    q[rr][k] = r < nr && gi < groups ? row[gi] : 0;
                }
            }
        };
        auto store = [&](int g, int buf) {
            uint16_t(*w0)[kSrWin] = win(buf);
            uint16_t(*w1)[kSrWin + 2] = wsh(buf);
#pragma unroll
            for (int rr = 0; rr < kSrG; rr++) {
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int gi = tid + k * kSrThreads;
                    if (g * kSrG + rr >= nr || gi >= groups)
                        continue;
                    const uint32_t lo32 = (uint32_t)q[rr][k], hi32 = (uint32_t)(q[rr][k] >> 32);
                    const int s0 = org + 4 * gi;
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const int v = (int)(((t < 2 ? lo32 : hi32) >> (16 * (t & 1))) & 0xffffu);
                        bad |= s0 + t >= lo && s0 + t <= hi && v > maxv;
                    }
                    *reinterpret_cast<uint2 *>(&w0[rr][4 * gi]) = make_uint2(lo32, hi32);
                    uint16_t *u = &w1[rr][4 * gi + 1];  // window index i at i + 1
                    u[0] = (uint16_t)lo32;
                    *reinterpret_cast<uint32_t *>(u + 1) = (lo32 >> 16) | (hi32 << 16);
                    u[3] = (uint16_t)(hi32 >> 16);
                }
            }
        };
        load(0);
        store(0, 0);
        __syncthreads();
        setup_columns();
        for (int g = 0; g < ngr; g++) {
            const int buf = R > kSrG ? (g & 1) : 0;
            if (g + 1 < ngr)
                load(g + 1);
            filter_group(g, buf);
            if (g + 1 < ngr) {
                store(g + 1, buf ^ 1);
                __syncthreads();
            }
        }
    } else {
        for (int g = 0; g < ngr; g++) {
            if (g > 0)
                __syncthreads();  // the previous group is filtered
            for (int rr = 0; rr < kSrG && g * kSrG + rr < nr; rr++) {
                gptr<const uint16_t> row = glb(j.in) + (size_t)yrow(g * kSrG + rr) * j.in_stride;
                for (int i = tid; i <= hi - lo; i += kSrThreads) {
                    const int v = row[min(max(lo + i, 0), last)];
                    bad |= v > maxv;
                    win(0)[rr][i] = (uint16_t)v;
                    wsh(0)[rr][i + 1] = (uint16_t)v;
                }
            }
            __syncthreads();
            if (g == 0)
                setup_columns();
            filter_group(g, 0);
        }
    }
    if (__syncthreads_or(bad) && tid == 0)
        raise_status(j.status, kFlagSampleRange);
}

}  // namespace av1pf
