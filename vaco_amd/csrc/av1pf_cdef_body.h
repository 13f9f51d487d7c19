// av1pf_cdef_body.h - CDEF (AV1 7.15) device code for gfx950: the
// direction search, the packed filter, LDS staging and the per-superblock
// body shared by the CDEF kernel and the chained CDEF -> restoration kernel.
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
#pragma once

#include "av1pf_common.h"

namespace av1pf {

// DIRECTIONS, cdef.rs:56-66, as (row, column).
constexpr int8_t kDirDy[8][2] = {{-1, -2}, {0, -1}, {0, 0}, {0, 1},
                                 {1, 2},   {1, 2},  {1, 2}, {1, 2}};
constexpr int8_t kDirDx[8][2] = {{1, 2}, {1, 2}, {1, 2}, {1, 2},
                                 {1, 2}, {0, 1}, {0, 0}, {0, -1}};

// Line-partition bin of pixel (i, j) for direction d (cdef.rs:88-97).
constexpr int bin_of(int d, int i, int j)
{
    return d == 0   ? i + j
           : d == 1 ? i + j / 2
           : d == 2 ? i
           : d == 3 ? 3 + i - j / 2
           : d == 4 ? 7 + i - j
           : d == 5 ? 3 + j - i / 2
           : d == 6 ? j
                    : i / 2 + j;
}

constexpr int bin_count(int d, int b)
{
    int c = 0;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++)
            if (bin_of(d, i, j) == b)
                c++;
    return c;
}

// Samples of bin b in rows i0 .. i0+3.
constexpr int bin_count_rows(int d, int b, int i0)
{
    int c = 0;
    for (int i = i0; i < i0 + 4; i++)
        for (int j = 0; j < 8; j++)
            if (bin_of(d, i, j) == b)
                c++;
    return c;
}

// LDS tiles. A plane's tile holds its rows y0-2 .. with column x0-8 at index
// 0 (luma: 68 rows of kRow samples, columns x0-8 .. x0+71). Each tile is
// followed by its shifted copy (index i at i + 1), SHIFT samples on. The
// luma row is 44 dwords: the 8 rows of a block then start in 8 distinct
// groups of 4 banks, so a half-wave filtering one block along one direction
// reads its taps without bank conflicts. Chroma tiles are 80 samples wide
// (48 when subsampled horizontally) and 68 rows high (36 when subsampled
// vertically), so that two workgroups' tiles fit a CU's LDS in every layout
// and four in 4:2:0.
constexpr int kRow = 88;
constexpr int kTile = 68 * kRow;
template <int SX>
constexpr int chroma_row() { return SX ? 48 : 80; }
template <int SX, int SY>
constexpr int chroma_tile() { return (SY ? 36 : 68) * chroma_row<SX>(); }

// LDS offsets (dy * ROW + dx) of DIRECTIONS[d][k] for directions 4h .. 4h+3,
// 16 bits each, so a lane selects its direction's offset with register
// operations instead of a memory load.
template <int ROW>
constexpr uint64_t dir_offsets(int k, int h)
{
    uint64_t v = 0;
    for (int i = 0; i < 4; i++)
        v |= (uint64_t)(uint16_t)(int16_t)(kDirDy[4 * h + i][k] * ROW + kDirDx[4 * h + i][k]) << (16 * i);
    return v;
}

template <int ROW>
__device__ __forceinline__ int tap_offset(int dir, int k)
{
    constexpr uint64_t t00 = dir_offsets<ROW>(0, 0), t01 = dir_offsets<ROW>(0, 1);
    constexpr uint64_t t10 = dir_offsets<ROW>(1, 0), t11 = dir_offsets<ROW>(1, 1);
    const uint64_t t = k ? (dir < 4 ? t10 : t11) : (dir < 4 ? t00 : t01);
    return (int16_t)(uint16_t)(t >> (16 * (dir & 3)));
}

// In every direction's partition, row i + 4 of a column falls in the bin of
// row i plus a constant (+4, +2, 0 or -2 by direction).
constexpr int row4_offset(int d) { return bin_of(d, 4, 0) - bin_of(d, 0, 0); }

constexpr bool row4_offset_is_uniform(int d)
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++)
            if (bin_of(d, i + 4, j) - bin_of(d, i, j) != row4_offset(d))
                return false;
    return true;
}
static_assert(row4_offset_is_uniform(0) && row4_offset_is_uniform(1) && row4_offset_is_uniform(2) &&
                  row4_offset_is_uniform(3) && row4_offset_is_uniform(4) && row4_offset_is_uniform(5) &&
                  row4_offset_is_uniform(6) && row4_offset_is_uniform(7),
              "row i + 4 is binned at a fixed offset from row i");

// Cost of direction D (cdef.rs:104-113) of the 8x8 block whose first row
// starts at `p` (16-byte aligned LDS rows of kRow samples), with samples
// taken >> shift. Samples (i, j) and (i + 4, j) are paired in one 32-bit
// word (one byte permute per pair) and summed two at a time: the low half of
// acc[b] collects bin b of rows 0-3 and the high half bin b + row4_offset(D)
// of rows 4-7. A half sums at most 8 samples of at most 255, so it cannot
// carry into the other. The -128 of cdef.rs:98 is applied once per bin as
// -128 * count. For in-range samples every cost is below 840*64*128^2 < 2^31.
// SH: shift is nonzero (8-bit frames skip the shifts).
template <int D, bool SH>
__device__ __forceinline__ uint32_t dir_cost(const uint16_t *p, int shift)
{
    constexpr int OFF = row4_offset(D);
    const u16x2 sh = (u16x2)(unsigned short)shift;
    u16x2 acc[15];
#pragma unroll
    for (int b = 0; b < 15; b++)
        acc[b] = (u16x2)0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint4 qa = *reinterpret_cast<const uint4 *>(p + i * kRow);
        const uint4 qc = *reinterpret_cast<const uint4 *>(p + (i + 4) * kRow);
        const uint32_t wa[4] = {qa.x, qa.y, qa.z, qa.w}, wc[4] = {qc.x, qc.y, qc.z, qc.w};
#pragma unroll
        for (int m = 0; m < 4; m++) {
            // (x[i][2m], x[i+4][2m]) and (x[i][2m+1], x[i+4][2m+1])
            const u16x2 ev = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(wc[m], wa[m], 0x05040100u));
            const u16x2 od = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(wc[m], wa[m], 0x07060302u));
            acc[bin_of(D, i, 2 * m)] += SH ? ev >> sh : ev;
            acc[bin_of(D, i, 2 * m + 1)] += SH ? od >> sh : od;
        }
    }
    uint32_t cost = 0;
#pragma unroll
    for (int b = 0; b < 15; b++) {
        const int cnt = bin_count(D, b);
        if (cnt != 0) {
            int part = 0;
            if (bin_count_rows(D, b, 0) != 0)
                part += acc[b].x;
            if (b - OFF >= 0 && b - OFF < 15 && bin_count_rows(D, b, 4) != 0)
                part += acc[b - OFF].y;
            const uint32_t s = (uint32_t)(part - 128 * cnt);  // s * s mod 2^32 is exact for in-range samples
            cost += s * s * (uint32_t)(840 / cnt);
        }
    }
    return cost;
}

template <bool SH>
__device__ __forceinline__ uint32_t dir_cost_sh(int d, const uint16_t *p, int shift)
{
    switch (d) {
    case 0: return dir_cost<0, SH>(p, shift);
    case 1: return dir_cost<1, SH>(p, shift);
    case 2: return dir_cost<2, SH>(p, shift);
    case 3: return dir_cost<3, SH>(p, shift);
    case 4: return dir_cost<4, SH>(p, shift);
    case 5: return dir_cost<5, SH>(p, shift);
    case 6: return dir_cost<6, SH>(p, shift);
    default: return dir_cost<7, SH>(p, shift);
    }
}

__device__ __forceinline__ uint32_t dir_cost_any(int d, const uint16_t *p, int shift)
{
    return shift ? dir_cost_sh<true>(d, p, shift) : dir_cost_sh<false>(d, p, shift);
}

// Filter parameters of one 8x8 block (one plane class), cdef.rs:35-54.
// Tap k's LDS offset (in samples) is that of direction dir, d1 / d2 (k = 0,
// 1), dir+2 (k = 2, 3) or dir+6 (k = 4, 5). The packed path reads sample
// pairs starting at ci + off and ci - off as one aligned 32-bit word: for an
// odd offset that pair starts at an odd index, so it is read from the
// shifted copy of the tile (index i stored at i + 1). offp / offm are those
// two reads' byte offsets from the pair's own address, resolved to the copy
// that makes them aligned.
struct DirOffs {
    int32_t offp[6], offm[6];
};

struct __align__(16) Blk {
    int32_t offp[6], offm[6];
    uint8_t pri, sec;    // strengths after scaling / variance adjustment
    uint8_t shifts;      // damping - ilog2(strength), primary | secondary << 4
    uint8_t flags;       // 1: filter this block, 2: primary taps are (3, 3)
};

// ROW, SHIFT: the row stride and shifted-copy distance of the block's tile.
template <int ROW, int SHIFT>
__device__ __forceinline__ DirOffs dir_offs(int dir)
{
    DirOffs D;
    const int d2 = (dir + 2) & 7, d6 = (dir + 6) & 7;
    int off[6];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        off[k] = tap_offset<ROW>(dir, k);
        off[2 + k] = tap_offset<ROW>(d2, k);
        off[4 + k] = tap_offset<ROW>(d6, k);
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int o = off[k];
        D.offp[k] = 2 * ((o & 1) ? SHIFT + 1 + o : o);
        D.offm[k] = 2 * ((o & 1) ? SHIFT + 1 - o : -o);
    }
    return D;
}

// A block's parameters without its tap offsets: the direction it filters
// along (0 when the primary strength is 0), strengths, shifts and flags. The
// frame kernel keeps these 8 bytes per block in LDS and takes the offsets
// from its per-workgroup table of the 8 directions when filtering.
struct BlkP {
    uint8_t dir, pri, sec, shifts, flags, pad[3];
};

__device__ __forceinline__ BlkP make_blkp(int dir, uint32_t pri, uint32_t sec, int damping, int shift)
{
    BlkP p = {};
    p.dir = (uint8_t)dir;
    p.pri = (uint8_t)min(pri, 255u);
    p.sec = (uint8_t)min(sec, 255u);
    const int ps = pri ? max(0, damping - ilog2u(pri)) : 0;  // constrain, cdef.rs:130
    const int ss = sec ? max(0, damping - ilog2u(sec)) : 0;
    p.shifts = (uint8_t)(min(ps, 15) | (min(ss, 15) << 4));
    p.flags = 1 | (((pri >> shift) & 1) ? 2 : 0);  // cdef.rs:156-160
    return p;
}

// The block's filter parameters: p with the offsets D of its direction.
__device__ __forceinline__ Blk blk_from(const BlkP &p, const DirOffs &D)
{
    Blk B;
    B.pri = p.pri;
    B.sec = p.sec;
    B.shifts = p.shifts;
    B.flags = p.flags;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        B.offp[k] = D.offp[k];
        B.offm[k] = D.offm[k];
    }
    return B;
}

template <int ROW, int SHIFT>
__device__ __forceinline__ Blk make_blk(int dir, uint32_t pri, uint32_t sec, int damping, int shift)
{
    return blk_from(make_blkp(dir, pri, sec, damping, shift), dir_offs<ROW, SHIFT>(dir));
}

// filter_block (cdef.rs:142-205), computed for two horizontally adjacent
// samples of one block at once. Every quantity fits int16: samples <= 4095, differences within
// +-4095, and the tap sum is at most 2*(4+2)*240 + 4*(2+1)*64 = 3648, so
// each operation works on both samples at once (v_pk_*_i16).

// Samples i and i + 1 of a tile as one aligned 32-bit LDS read (i even).
__device__ __forceinline__ s16x2 ld2(const uint16_t *t, int i)
{
    return *reinterpret_cast<const s16x2 *>(t + i);
}

// Samples i and i + 1 at byte address a of a tile (a multiple of 4).
__device__ __forceinline__ s16x2 ldb(const char *a)
{
    return *reinterpret_cast<const s16x2 *>(a);
}

// filter_block (cdef.rs:142-205) for N pairs of horizontally adjacent
// samples of one block, the pairs STRIDE samples apart, starting at tile
// index ci (even). The N computations run interleaved, which fills the
// dependency gaps of each chain.
//
// constrain(d) (cdef.rs:126-133) is sign(d) * min(|d|, T) with T = max(0,
// S - (|d| >> shift)). With samples unsigned, a = p -sat x and b = x -sat p
// (saturating at 0) give |d| = a | b, T = S -sat (|d| >> shift) and
// constrain(d) = min(a, T) - min(b, T): one of a, b is zero.
//
// An unavailable tap (kUnavailable) is skipped by the reference
// (cdef.rs:189-191): no term, no part in the clamp range. Taking the centre
// sample in its place has exactly that effect (constrain(0) = 0, and x is
// already in [mn, mx]), so BORDER superblocks swap it in with one bit
// select per tap.
// The largest and the smallest of three sample pairs, lane by lane, in one
// packed instruction each: samples are at most 4095, and as f16 bit patterns
// those are +0, subnormals and small normals, which f16 orders as the
// integers (the kernels keep f16 subnormals, they do not flush them).
__device__ __forceinline__ u16x2 max3_u16(u16x2 a, u16x2 b, u16x2 c)
{
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 m = __builtin_elementwise_maximum(__builtin_bit_cast(h2, a), __builtin_bit_cast(h2, b));
    return __builtin_bit_cast(u16x2, __builtin_elementwise_maximum(m, __builtin_bit_cast(h2, c)));
}

__device__ __forceinline__ u16x2 min3_u16(u16x2 a, u16x2 b, u16x2 c)
{
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 m = __builtin_elementwise_minimum(__builtin_bit_cast(h2, a), __builtin_bit_cast(h2, b));
    return __builtin_bit_cast(u16x2, __builtin_elementwise_minimum(m, __builtin_bit_cast(h2, c)));
}

// Taps k0 .. k1-1 (both signs) of filter_pairs. CLAMP = false: no clamp
// range is kept (see filter_pairs).
template <bool BORDER, bool CLAMP, int N, int STRIDE, int K0, int K1>
__device__ __forceinline__ void apply_taps(const char *a, const Blk &B, const u16x2 *x, u16x2 S,
                                           u16x2 sh, s16x2 w0, s16x2 w1, s16x2 *sum, u16x2 *mn,
                                           u16x2 *mx)
{
#pragma unroll
    for (int k = K0; k < K1; k++) {
        const s16x2 w = ((k - K0) & 1) ? w1 : w0;
#pragma unroll
        for (int n = 0; n < N; n++) {
            u16x2 ps[2];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const int o = s ? B.offp[k] : B.offm[k];
                u16x2 p = __builtin_bit_cast(u16x2, ldb(a + 2 * STRIDE * n + o));
                if (BORDER) {
                    const uint32_t m =
                        __builtin_bit_cast(uint32_t, __builtin_bit_cast(s16x2, p) >> (s16x2)15);
                    const uint32_t v = (m & __builtin_bit_cast(uint32_t, x[n])) |
                                       (~m & __builtin_bit_cast(uint32_t, p));
                    p = __builtin_bit_cast(u16x2, v);
                }
                const u16x2 up = __builtin_elementwise_sub_sat(p, x[n]);
                const u16x2 dn = __builtin_elementwise_sub_sat(x[n], p);
                const u16x2 T = __builtin_elementwise_sub_sat(S, (up | dn) >> sh);
                const s16x2 c = __builtin_bit_cast(s16x2, __builtin_elementwise_min(up, T)) -
                                __builtin_bit_cast(s16x2, __builtin_elementwise_min(dn, T));
                sum[n] += w * c;
                ps[s] = p;
            }
            if (CLAMP) {  // both signs of the tap at once
                mn[n] = min3_u16(mn[n], ps[0], ps[1]);
                mx[n] = max3_u16(mx[n], ps[0], ps[1]);
            }
        }
    }
}

// pri_on / sec_on: the superblock's preset has a nonzero primary / secondary
// strength (uniform over the workgroup, so the branches are scalar).
//
// With only one of them on, the clamp of cdef.rs:200-202 cannot bind, so it
// and the other direction's taps are skipped: the tap weights of one
// direction sum to 12 (2 * (4 + 2), 2 * (3 + 3) or 4 * (2 + 1)), each term has
// the sign of p - x and at most its size, so with D = the largest p - x over
// the taps (0 if none is above x), 8 + sum <= 8 + 12 D < 16 (D + 1) and y <=
// x + D, the largest tap; likewise y >= the smallest. (An unavailable tap
// contributes nothing either way.)
template <bool BORDER, int N, int STRIDE>
__device__ __forceinline__ void filter_pairs(const uint16_t *t, int ci, const Blk &B, s16x2 *res,
                                             bool pri_on, bool sec_on)
{
    const char *a = reinterpret_cast<const char *>(t + ci);
    u16x2 x[N], mn[N], mx[N];
    s16x2 sum[N];
#pragma unroll
    for (int n = 0; n < N; n++) {
        x[n] = __builtin_bit_cast(u16x2, ldb(a + 2 * STRIDE * n));
        sum[n] = (s16x2)0;
        mn[n] = mx[n] = x[n];
    }
    const u16x2 P = (u16x2)(unsigned short)B.pri, Q = (u16x2)(unsigned short)B.sec;
    const u16x2 ps = (u16x2)(unsigned short)(B.shifts & 15), qs = (u16x2)(unsigned short)(B.shifts >> 4);
    const short t0 = (B.flags & 2) ? 3 : 4;
    const s16x2 w0 = (s16x2)t0, w1 = (s16x2)(short)(6 - t0);
    const bool clamp = pri_on && sec_on;
    if (clamp) {
        apply_taps<BORDER, true, N, STRIDE, 0, 2>(a, B, x, P, ps, w0, w1, sum, mn, mx);
        apply_taps<BORDER, true, N, STRIDE, 2, 6>(a, B, x, Q, qs, (s16x2)2, (s16x2)1, sum, mn, mx);
    } else if (pri_on) {
        apply_taps<BORDER, false, N, STRIDE, 0, 2>(a, B, x, P, ps, w0, w1, sum, mn, mx);
    } else if (sec_on) {
        apply_taps<BORDER, false, N, STRIDE, 2, 6>(a, B, x, Q, qs, (s16x2)2, (s16x2)1, sum, mn, mx);
    }
#pragma unroll
    for (int n = 0; n < N; n++) {
        const s16x2 xs = __builtin_bit_cast(s16x2, x[n]);
        const s16x2 y = xs + ((sum[n] + (s16x2)8 + (sum[n] >> (s16x2)15)) >> (s16x2)4);  // cdef.rs:200-201
        res[n] = clamp ? __builtin_elementwise_min(__builtin_elementwise_max(y, __builtin_bit_cast(s16x2, mn[n])),
                                                   __builtin_bit_cast(s16x2, mx[n]))
                       : y;
    }
}

// Stage rows ys .. ys+ROWS-1, columns xs .. xs+COLS-1 of a plane into an LDS
// tile of row stride ROW at column offset 6, and into the shifted copy at
// dst + SHIFT; positions outside the plane hold kUnavailable. T: threads
// of the workgroup.
template <int ROWS, int COLS, int ROW, int SHIFT, int T>
__device__ __forceinline__ bool stage(uint16_t *dst, gptr<const uint16_t> src, int W, int H, int st,
                                      int xs, int ys, bool border, int maxv)
{
    bool bad = false;
#pragma unroll 5
    for (int e = threadIdx.x; e < ROWS * COLS; e += T) {
        const int r = e / COLS, c = e - r * COLS;
        const int x = xs + c, y = ys + r;
        uint16_t v = kUnavailable;
        if (!border || (x >= 0 && y >= 0 && x < W && y < H)) {
            v = src[(size_t)y * st + x];
            bad |= v > maxv;
        }
        dst[r * ROW + c + 6] = v;
        dst[SHIFT + r * ROW + c + 7] = v;
    }
    return bad;
}

// A window of a plane with 8-byte aligned rows, staged as stage() stages it
// but four samples at a time: the window widened to start at column xs - 2
// (the two extra columns on each side are never read by a tap), loaded in
// two phases so that a thread's loads for every plane are in flight
// together. load() issues the thread's global loads into registers; store()
// writes them, and the shifted copy, to the LDS tile once they arrive.
// Group columns are multiples of 4 and plane widths are multiples of 4, so a
// group lies wholly inside or wholly outside the plane; outside (BORDER only)
// it holds kUnavailable. An outside group's load reads the nearest group
// inside the plane instead, so every load is in bounds and unconditional.
template <int ROWS, int COLS, bool BORDER, int T>
struct VecStage {
    static constexpr int V = (COLS + 4) / 4;  // 4-sample groups per row
    static constexpr int N = ROWS * V;
    static constexpr int K = (N + T - 1) / T;
    static_assert(K <= 32, "one validity bit per group");
    uint2 q[K];
    uint32_t inside;  // bit k: group k lies in the plane

    __device__ __forceinline__ void load(gptr<const uint16_t> src, int W, int H, int st, int xs, int ys)
    {
        inside = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = min((int)threadIdx.x + k * T, N - 1);
            const int r = e / V, v = e - r * V;
            int y = ys + r, x = xs - 2 + 4 * v;
            if (BORDER) {
                if (y >= 0 && y < H && x >= 0 && x < W)
                    inside |= 1u << k;
                y = min(max(y, 0), H - 1);
                x = min(max(x, 0), W - 4);
            } else {
                inside |= 1u << k;
            }
            const uint64_t w = *reinterpret_cast<gptr<const uint64_t>>(src + (size_t)y * st + x);
            q[k] = {(uint32_t)w, (uint32_t)(w >> 32)};
        }
    }

    // ROW, SHIFT: the tile's row stride and shifted-copy distance. Returns
    // whether a sample above maxv was read.
    template <int ROW, int SHIFT>
    __device__ __forceinline__ bool store(uint16_t *dst, int maxv) const
    {
        bool bad = false;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = (int)threadIdx.x + k * T;
            if (N % T != 0 && e >= N)
                break;
            const int r = e / V, v = e - r * V;
            uint2 w = q[k];
            if ((inside >> k) & 1) {
                const uint32_t m = max(max(w.x & 0xffffu, w.x >> 16), max(w.y & 0xffffu, w.y >> 16));
                bad |= (int)m > maxv;
            } else {
                w = {0xffffffffu, 0xffffffffu};
            }
            uint16_t *t = static_cast<uint16_t *>(__builtin_assume_aligned(dst + r * ROW + 4 + 4 * v, 8));
            *reinterpret_cast<uint2 *>(t) = w;
            uint16_t *u = t + SHIFT + 1;  // sample i of the tile at i + 1
            u[0] = (uint16_t)w.x;
            *reinterpret_cast<uint32_t *>(u + 1) = (w.x >> 16) | (w.y << 16);
            u[3] = (uint16_t)(w.y >> 16);
        }
        return bad;
    }
};

// Store a pair of output samples at element index i of a plane of u16
// samples, or of bytes when u8 (8-bit frame output); the second sample only
// when `second` (a crop to an odd width can end between the two).
// COH (chained launch): the output is an intermediate that workgroups on
// other XCDs read within the launch, u16, 4-byte aligned pairs of an even
// width; each pair is one agent-scope store, written through the XCD's L2,
// so that the release which publishes the superblock has no dirty lines of
// it to write back.
template <bool COH = false>
__device__ __forceinline__ void store_pair(gptr<uint16_t> plane, size_t i, s16x2 v, bool u8, bool nt,
                                           bool second)
{
    if (COH) {
        __hip_atomic_store(reinterpret_cast<gptr<uint32_t>>(plane + i),
                           (uint32_t)(uint16_t)v.x | ((uint32_t)(uint16_t)v.y << 16), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    // Pairs at 4-byte (2-byte for bytes) aligned addresses take one store.
    // nt: non-temporal stores, for an output that no launch reads from the
    // L2 right after: lines left dirty in the L2 would be written back by the
    // release at the launch's end, which the next launch waits for.
    if (!nt) {
        if (u8) {
            gptr<uint8_t> o = reinterpret_cast<gptr<uint8_t>>(plane) + i;
            o[0] = (uint8_t)v.x;
            if (second)
                o[1] = (uint8_t)v.y;
        } else {
            gptr<uint16_t> o = plane + i;
            if (second && ((uintptr_t)o & 3) == 0) {
                *reinterpret_cast<gptr<uint32_t>>(o) = (uint32_t)(uint16_t)v.x | ((uint32_t)(uint16_t)v.y << 16);
            } else {
                o[0] = (uint16_t)v.x;
                if (second)
                    o[1] = (uint16_t)v.y;
            }
        }
        return;
    }
    if (u8) {
        gptr<uint8_t> o = reinterpret_cast<gptr<uint8_t>>(plane) + i;
        if (second && ((uintptr_t)o & 1) == 0) {
            __builtin_nontemporal_store((uint16_t)((v.x & 255) | (v.y << 8)), reinterpret_cast<gptr<uint16_t>>(o));
        } else {
            __builtin_nontemporal_store((uint8_t)v.x, o);
            if (second)
                __builtin_nontemporal_store((uint8_t)v.y, o + 1);
        }
    } else {
        gptr<uint16_t> o = plane + i;
        if (second && ((uintptr_t)o & 3) == 0) {
            __builtin_nontemporal_store((uint32_t)(uint16_t)v.x | ((uint32_t)(uint16_t)v.y << 16),
                                        reinterpret_cast<gptr<uint32_t>>(o));
        } else {
            __builtin_nontemporal_store((uint16_t)v.x, o);
            if (second)
                __builtin_nontemporal_store((uint16_t)v.y, o + 1);
        }
    }
}

// The byte fields of a CdefFrame, read as two words (scalar loads; byte
// loads would be vector loads, each waited on before its first use).
struct FrameBytes {
    uint32_t w0, w1;
    __device__ __forceinline__ explicit FrameBytes(const CdefFrame &f)
    {
        static_assert(offsetof(CdefFrame, bit_depth) % 4 == 0 &&
                      offsetof(CdefFrame, ss_x) == offsetof(CdefFrame, bit_depth) + 1 &&
                      offsetof(CdefFrame, ss_y) == offsetof(CdefFrame, bit_depth) + 2 &&
                      offsetof(CdefFrame, nplanes) == offsetof(CdefFrame, bit_depth) + 3 &&
                      offsetof(CdefFrame, damping) == offsetof(CdefFrame, bit_depth) + 4 &&
                      offsetof(CdefFrame, vec_ok) == offsetof(CdefFrame, bit_depth) + 5 &&
                      offsetof(CdefFrame, out_u8) == offsetof(CdefFrame, bit_depth) + 6 &&
                      offsetof(CdefFrame, out_nt) == offsetof(CdefFrame, bit_depth) + 7,
                      "CdefFrame byte fields are read as two words");
        __builtin_memcpy(&w0, &f.bit_depth, 4);
        __builtin_memcpy(&w1, &f.damping, 4);
    }
    __device__ __forceinline__ int bit_depth() const { return (int)(w0 & 255u); }
    __device__ __forceinline__ int nplanes() const { return (int)(w0 >> 24); }
    __device__ __forceinline__ int damping() const { return (int)(w1 & 255u); }
    __device__ __forceinline__ bool vec_ok() const { return ((w1 >> 8) & 255u) != 0; }
    __device__ __forceinline__ bool out_u8() const { return ((w1 >> 16) & 255u) != 0; }
    __device__ __forceinline__ bool out_nt() const { return (w1 >> 24) != 0; }
};

// The descriptor fields the phases after staging use, read at the start of
// the workgroup so that their load latency hides under the staging loads.
struct SbFields {
    gptr<uint16_t> out[3];
    uint32_t out_stride[3];
    int crop_w[3], crop_h[3];
    uint32_t str[8];  // y_pri, y_sec, uv_pri, uv_sec: 8 presets each, 4 per word
    int damping, nplanes;
    bool out_u8, out_nt;

    __device__ __forceinline__ void load(const CdefFrame &f, const FrameBytes &fb)
    {
        static_assert(offsetof(CdefFrame, y_pri) % 4 == 0 && offsetof(CdefFrame, y_sec) ==
                      offsetof(CdefFrame, y_pri) + 8 && offsetof(CdefFrame, uv_sec) ==
                      offsetof(CdefFrame, y_pri) + 24, "strength tables are read as words");
        const uint32_t *w = reinterpret_cast<const uint32_t *>(f.y_pri);
#pragma unroll
        for (int i = 0; i < 8; i++)
            str[i] = w[i];
#pragma unroll
        for (int p = 0; p < 3; p++) {
            out[p] = glb(f.out[p]);
            out_stride[p] = f.out_stride[p];
            crop_w[p] = (int)f.crop_w[p];
            crop_h[p] = (int)f.crop_h[p];
        }
        damping = fb.damping();
        nplanes = fb.nplanes();
        out_u8 = fb.out_u8();
        out_nt = fb.out_nt();
    }

    // Preset ix of table t (0 y_pri, 1 y_sec, 2 uv_pri, 3 uv_sec).
    __device__ __forceinline__ uint32_t strength(int t, int ix) const
    {
        const uint32_t w = (ix & 4) ? str[2 * t + 1] : str[2 * t];
        return (w >> (8 * (ix & 3))) & 255u;
    }
};

// Stage the superblock's windows of every plane (luma at x0-2, y0-2, 68x68;
// chroma likewise subsampled) from 8-byte aligned planes: all loads first,
// then `between()`, then all LDS stores. Returns whether a sample above maxv
// was read.
template <int SX, int SY, bool BORDER, int T, class Between>
__device__ __forceinline__ bool stage_planes(const CdefFrame &f, uint16_t *L, uint16_t *U, uint16_t *V,
                                             int nplanes, int x0, int y0, int maxv, Between &&between)
{
    constexpr int CW = 64 >> SX, CH = 64 >> SY;
    constexpr int CROW = chroma_row<SX>(), CTILE = chroma_tile<SX, SY>();
    const int cx = (x0 >> SX) - 2, cy = (y0 >> SY) - 2;
    VecStage<68, 68, BORDER, T> l;
    VecStage<CH + 4, CW + 4, BORDER, T> u, v;
    l.load(glb(f.in[0]), (int)f.width[0], (int)f.height[0], (int)f.in_stride[0], x0 - 2, y0 - 2);
    if (nplanes > 1) {
        u.load(glb(f.in[1]), (int)f.width[1], (int)f.height[1], (int)f.in_stride[1], cx, cy);
        v.load(glb(f.in[2]), (int)f.width[2], (int)f.height[2], (int)f.in_stride[2], cx, cy);
    }
    between();
    bool bad = l.template store<kRow, kTile>(L, maxv);
    if (nplanes > 1) {
        bad |= u.template store<CROW, CTILE>(U, maxv);
        bad |= v.template store<CROW, CTILE>(V, maxv);
    }
    return bad;
}

// Filter (or pass through) every sample of a staged superblock and store it.
// blist holds the superblock's 64 blocks, the nact blocks to filter first:
// entries are taken in list order, so a wave's blocks are all filtered or
// all passed through except at one boundary, and no wave runs the filter
// with lanes idle for a skipped neighbour.
// Luma: each half-wave takes one 8x8 block (32 lanes = 8 rows x 4 pairs of
// columns), entries hw, hw + T / 32, ... for half-wave hw; all its lanes then
// read their taps at one direction's offsets, which the luma row stride keeps
// in distinct banks. Chroma: 16 lanes per block, pairs of U then V. BORDER: the
// superblock touches a plane edge (unavailable taps; partial superblocks
// store only samples inside the crop; widths are even, so a pair is all in
// or all out, except at an odd crop width where the second sample is
// dropped). L, U, V: the planes' tiles.
template <bool BORDER, int SX, int SY, int T, bool COH = false>
__device__ __forceinline__ void filter_sb(const SbFields &f, const uint16_t *L, const uint16_t *U,
                                          const uint16_t *V, const BlkP (*bp)[64], const DirOffs (*dirs)[8],
                                          const uint8_t *blist,
                                          int nact, int x0, int y0, int ix, bool prio)
{
    constexpr int CROW = chroma_row<SX>();
    // Strengths are per superblock preset (the luma primary is only ever
    // lowered per block, never raised from zero).
    const bool ypri = f.strength(0, ix) != 0, ysec = f.strength(1, ix) != 0;
    const bool cpri = f.strength(2, ix) != 0, csec = f.strength(3, ix) != 0;
    constexpr int BW = 8 >> SX, BH = 8 >> SY, PPP = BW * BH / 2, CPX = BW * BH / 16;
    const int tid = threadIdx.x;
    {
        constexpr int HW = T / 32;  // half-waves: luma blocks per pass
        const int hw = tid >> 5, l = tid & 31, i = l >> 2, j = (l & 3) * 2;
        gptr<uint16_t> dst = f.out[0];
        const int ost = f.out_stride[0], W0 = f.crop_w[0], H0 = f.crop_h[0];
#pragma unroll
        for (int pass = 0; pass < 64 / HW; pass++) {
            const int e = hw + HW * pass, b = blist[e], bi = b >> 3, bj = b & 7;
            const bool act = e < nact;
            const int ci = (2 + bi * 8 + i) * kRow + 8 + bj * 8 + j;
            s16x2 res;
            if (act) {
                const BlkP p = bp[0][b];
                const Blk B = blk_from(p, dirs[0][p.dir]);
                filter_pairs<BORDER, 1, 0>(L, ci, B, &res, ypri, ysec);
            } else {
                res = ld2(L, ci);
            }
            const int x = x0 + bj * 8 + j, y = y0 + bi * 8 + i;
            if (!BORDER || (x < W0 && y < H0))
                store_pair<COH>(dst, (size_t)y * ost + x, res, f.out_u8, f.out_nt, x + 1 < W0);
        }
    }
    constexpr int CB = T / 16;  // chroma blocks per pass
    if (prio)
        __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int cp = 0; cp < 64 / CB; cp++) {
        if (f.nplanes == 1)
            break;
        const int e = (tid >> 4) + CB * cp, b = blist[e], q = tid & 15;
        const int bi = b >> 3, bj = b & 7;
        const bool act = e < nact;
        Blk B = {};
        if (act) {
            const BlkP p = bp[1][b];
            B = blk_from(p, dirs[1][p.dir]);
        }
        const int cx0 = x0 >> SX, cy0 = y0 >> SY, W1 = f.crop_w[1], H1 = f.crop_h[1];
#pragma unroll
        for (int k = 0; k < CPX; k++) {
            const int pi = q + 16 * k;
            const int p = pi < PPP ? 1 : 2, r = pi % PPP;
            const int i = r / (BW / 2), j = (r % (BW / 2)) * 2;
            const int x = cx0 + bj * BW + j, y = cy0 + bi * BH + i;
            const int ci = (2 + bi * BH + i) * CROW + 8 + bj * BW + j;
            const uint16_t *t = p == 1 ? U : V;
            s16x2 res;
            if (act)
                filter_pairs<BORDER, 1, 0>(t, ci, B, &res, cpri, csec);
            else
                res = ld2(t, ci);
            if (!BORDER || (x < W1 && y < H1))
                store_pair<COH>(p == 1 ? f.out[1] : f.out[2],
                           (size_t)y * (p == 1 ? f.out_stride[1] : f.out_stride[2]) + x, res, f.out_u8, f.out_nt,
                           x + 1 < W1);
        }
    }
}


// LDS of one CDEF workgroup: the luma tile and its shifted copy, then U and
// V, each with its copy (V starts 8 samples further on: U and V pairs of one
// half-wave then fall in different banks; a model of the tap addresses gives
// 2.3 accesses per bank per read without the gap, 1.3 with it); the block
// parameters, the directions' tap offsets, direction costs and block list.
template <int SX, int SY>
struct __align__(16) CdefLds {
    uint16_t lt[2 * kTile + 4 * chroma_tile<SX, SY>() + 8];
    DirOffs dirs[2][8];  // the 8 directions' tap offsets, luma then chroma tiles
    // The direction costs are dead once the parameter phase has read them;
    // the block parameters take their place.
    union {
        uint32_t cost[8][64];
        BlkP bp[2][64];  // block parameters, luma then chroma
    };
    unsigned long long bact;  // bit b: block b is filtered
    uint8_t blist[64];        // blocks to filter (ascending), then the rest
    int snact;
    int sidx;
};
// Four 4:2:0 workgroups share a CU's 160 KiB of LDS (4K CDEF 36.2 to 34.1 us,
// 64-frame batch 446 to 424 us against three).
static_assert(sizeof(CdefLds<1, 1>) <= 160 * 1024 / 4, "four 4:2:0 workgroups per CU");

// One 64x64 luma superblock of one frame, all planes (decode.rs:772-856).
// SX, SY: chroma subsampling (monochrome frames use <0, 0> with 1 plane).
// XCD: superblocks numbered XCD-contiguously (xcd_remap).
//
// ARG: one frame, whose descriptor is the kernel argument `single` (kernel
// arguments are read through the scalar cache); otherwise `frames` holds
// n_frames descriptors and up_* are unused.
//
// ARG (up_n16 > 0): the call's staged side information has not been copied
// to the device; the CDEF side information the descriptor points to is read
// in place from coherent host memory, and the launch copies the up_n16
// 16-byte units at up_src (the rest of the call's descriptors, for the
// kernels after it) to up_dst: workgroup b's first wave copies units
// b * 64 + lane, b * 64 + lane + 64 * n_wg, ..., the first issued with
// the staging loads and stored after them.
//
// b, n_wg: the workgroup's index among the n_wg CDEF workgroups of the
// launch. CHAIN (chained launch, ARG only): the superblock's completion is
// published in cs.flags, and the staged-bytes copy in cs.copied.
template <int SX, int SY, bool XCD, int T, bool ARG, bool CHAIN>
__device__ __forceinline__ void cdef_sb(CdefLds<SX, SY> &S, const CdefFrame *__restrict__ frames, int n_frames,
                                        const CdefFrame &single, const uint4 *__restrict__ up_src,
                                        uint4 *__restrict__ up_dst, uint32_t up_n16, uint32_t b, uint32_t n_wg,
                                        const ChainSync &cs)
{
    static_assert(T == 512 || T == 1024, "the direction search takes 512 threads");
    constexpr int CW = 64 >> SX, CH = 64 >> SY;
    constexpr int CROW = chroma_row<SX>(), CTILE = chroma_tile<SX, SY>();
    uint16_t *const L = S.lt, *const U = S.lt + 2 * kTile, *const V = U + 2 * CTILE + 8;
    auto &bp = S.bp;
    auto &cost = S.cost;
    auto &blist = S.blist;
    int &snact = S.snact;
    int &sidx = S.sidx;

    // One frame's launch (a wave or a few of workgroups, all starting
    // together): wave priority falls as the superblock advances (staging 3,
    // direction search 2, luma filtering 1, chroma filtering 0). Of the
    // workgroups a CU holds, the one behind then wins the SIMD arbiter (which
    // otherwise favours the oldest waves), so they finish together instead of
    // the younger one running its last phases alone with half the waves
    // (1080p: 18.0 to 16.9 us). A batch's deep grid, where workgroups arrive
    // as others leave, keeps one priority (its CDEF measured 5% slower with
    // the ladder).
    const bool prio = ARG;
    if (prio)
        __builtin_amdgcn_s_setprio(3);
    const uint32_t wg = XCD ? xcd_remap(b, n_wg) : b;
    const CdefFrame &f = ARG ? single : find_item(frames, n_frames, wg);
    const uint32_t sb = wg - f.block_begin;
    const int sby = (int)(sb / f.sb_cols), sbx = (int)(sb - (uint32_t)sby * f.sb_cols);
    const int tid = threadIdx.x;
    const FrameBytes fb(f);
    const int bd = fb.bit_depth(), shift = bd - 8, maxv = (1 << bd) - 1;
    const int nplanes = fb.nplanes();
    const int x0 = sbx * 64, y0 = sby * 64;
    const int W0 = f.width[0], H0 = f.height[0];
    SbFields sf;
    sf.load(f, fb);
    const bool border = x0 == 0 || y0 == 0 || x0 + 64 >= W0 || y0 + 64 >= H0;

    // The superblock's cdef_idx and the skip flags of its 64 blocks
    // (decode.rs:774-777, 892-896), loaded by the first wave (lane l: block
    // l) between issuing its sample loads and waiting for them, and left
    // untouched until side_info() turns them into the block list after
    // staging: vector loads complete in order, so loads issued before the
    // samples, or used before the samples are stored, would hold up the
    // staging by their latency (they may come from host memory). Every wave
    // issues the same loads with no branch around them, so that the waits for
    // the samples count them exactly; the other waves read the luma plane's
    // first byte instead. Blocks outside the plane read the last block's flag
    // (padded sizes are multiples of 8) and count as skipped. On the ARG
    // path the first wave also loads its share of the staged-bytes copy
    // (lanes past the end re-read unit 0).
    const uint32_t up_i = b * 64u + (uint32_t)tid;
    const bool up_first = ARG && tid < 64 && up_i < up_n16;
    uint4 up_v;
    uint32_t skip_raw;  // nonzero: skipped
    int vidx;
    bool outside;  // the lane's block lies outside the plane
    auto load_side_info = [&]() {
        const bool w0 = tid < 64;
        const int bx = x0 + (tid & 7) * 8, by = y0 + ((tid >> 3) & 7) * 8;
        const int bxc = min(bx, W0 - 8), byc = min(by, H0 - 8);
        outside = bx >= W0 || by >= H0;
        const gptr<const uint8_t> any = reinterpret_cast<gptr<const uint8_t>>(glb(f.in[0]));
        const gptr<const int8_t> ia =
            w0 ? glb(f.cdef_idx) + sby * f.idx_stride + sbx : reinterpret_cast<gptr<const int8_t>>(any);
        const gptr<const uint8_t> sa =
            w0 ? glb(f.skip8) + (size_t)(byc >> 3) * f.skip_stride + (bxc >> 3) : any;
        if (ARG)
            up_v = glb(up_src)[up_first ? up_i : 0u];
        vidx = *ia;
        skip_raw = *sa;
    };
    bool bad = false;
    const int cx = (x0 >> SX) - 2, cy = (y0 >> SY) - 2;
    if (!fb.vec_ok()) {
        load_side_info();
        bad = stage<68, 68, kRow, kTile, T>(L, glb(f.in[0]), W0, H0, f.in_stride[0], x0 - 2, y0 - 2, border,
                                         maxv);
        if (nplanes > 1) {
            bad |= stage<CH + 4, CW + 4, CROW, CTILE, T>(U, glb(f.in[1]), f.width[1], f.height[1],
                                                      f.in_stride[1], cx, cy, border, maxv);
            bad |= stage<CH + 4, CW + 4, CROW, CTILE, T>(V, glb(f.in[2]), f.width[2], f.height[2],
                                                      f.in_stride[2], cx, cy, border, maxv);
        }
    } else {
        // Interior superblocks: rows y0-2 .. y0+65 and columns x0-4 .. x0+67
        // lie inside the plane, because padded sizes are multiples of 8 and
        // the superblock does not touch the last 8 rows or columns (likewise
        // for chroma), so no bounds are checked.
        if (border)
            bad = stage_planes<SX, SY, true, T>(f, L, U, V, nplanes, x0, y0, maxv, load_side_info);
        else
            bad = stage_planes<SX, SY, false, T>(f, L, U, V, nplanes, x0, y0, maxv, load_side_info);
    }
    // Wave 0 stores its share of the staged-bytes copy and turns the loaded
    // skip flags and cdef_idx into the block list.
    auto side_info = [&]() {
        // Wave 1 tabulates the 8 directions' tap offsets, which the
        // filtering reads by each block's direction (their arithmetic then
        // leaves the parameter phase, and a block keeps 8 bytes of LDS).
        if (tid >= 64 && tid < 80) {
            const int d = tid & 7;
            if (tid < 72)
                S.dirs[0][d] = dir_offs<kRow, kTile>(d);
            else
                S.dirs[1][d] = dir_offs<CROW, CTILE>(d);
        }
        if (tid < 64) {
            if (ARG && up_first) {
                gptr<uint4> dst = glb(up_dst);
                dst[up_i] = up_v;
                for (uint32_t i = up_i + 64u * n_wg; i < up_n16; i += 64u * n_wg)
                    dst[i] = glb(up_src)[i];
            }
            // Chained launch: the staged bytes this wave copied are published
            // (the release's wait covers every lane of the wave).
            if (CHAIN && ARG && b * 64u < up_n16 && tid == 0)
                __hip_atomic_fetch_add(cs.copied, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const bool a = vidx >= 0 && skip_raw == 0 && !outside;
            const uint64_t m = __ballot(a);
            if (tid == 0)
                S.bact = m;
            const int below = __popcll(m & ((1ull << tid) - 1)), nact = __popcll(m);
            blist[a ? below : nact + (tid - below)] = (uint8_t)tid;
            if (tid == 0) {
                snact = nact;
                sidx = vidx;
            }
        }
    };
    // Direction search: wave w < 8 computes direction w, lane = block.
    auto search = [&]() {
        if (tid < 512) {
            const int b = tid & 63;
            if ((S.bact >> b) & 1)
                cost[tid >> 6][b] =
                    dir_cost_any(tid >> 6, &L[(2 + (b >> 3) * 8) * kRow + 8 + (b & 7) * 8], shift);
        }
    };
    side_info();
    if (__syncthreads_or(bad) && tid == 0)
        raise_status(f.status, kFlagSampleRange);
    const int idx = sidx;
    if (prio)
        __builtin_amdgcn_s_setprio(2);

    if (idx >= 0) {
        search();
        __syncthreads();
        // Wave 0, lane = block: the block's 8 costs into registers, then
        // its luma and chroma parameters written over the costs' LDS. One
        // wave's LDS accesses complete in order, so its writes cannot
        // overtake any lane's reads of the costs, and no other wave reads
        // them.
        if (tid < 64) {
            const int b = tid;
            BlkP BY = {}, BC = {};
            if ((S.bact >> b) & 1) {
                uint32_t c[8];
#pragma unroll
                for (int d = 0; d < 8; d++)
                    c[d] = cost[d][b];
                int best = 0;
                uint32_t bc = c[0];
#pragma unroll
                for (int d = 1; d < 8; d++) {
                    if (c[d] > bc) {  // first maximum wins, cdef.rs:114-119
                        bc = c[d];
                        best = d;
                    }
                }
                uint32_t opp = c[0];  // cost of the orthogonal direction
#pragma unroll
                for (int d = 1; d < 8; d++)
                    opp = ((best + 4) & 7) == d ? c[d] : opp;
                const uint32_t var = (bc - opp) >> 10;
                const int ix = idx & 7;
                {  // luma, decode.rs:806-818
                    const uint32_t pri = sf.strength(0, ix) << shift;
                    uint32_t adj = 0;
                    if (var != 0) {  // adjust_strength, cdef.rs:208-220
                        const uint32_t sc = var >> 6;
                        const uint32_t lg = sc ? min(ilog2u(sc), 12) : 0;
                        adj = (pri * (4 + lg) + 8) >> 4;
                    }
                    BY = make_blkp(pri ? best : 0, adj, sf.strength(1, ix) << shift, (uint8_t)(sf.damping + shift),
                                   shift);
                }
                {  // chroma
                    const uint32_t pri = sf.strength(2, ix) << shift;
                    // chroma_direction, cdef.rs:222-230: nibble `best` of
                    // {7,0,2,4,5,6,6,6} (4:2:2) or {1,2,2,2,3,4,6,0} (4:4:0).
                    int cdir = best;
                    if (SX && !SY)
                        cdir = (0x66654207u >> (4 * best)) & 7;
                    else if (!SX && SY)
                        cdir = (0x06432221u >> (4 * best)) & 7;
                    BC = make_blkp(pri ? cdir : 0, pri, sf.strength(3, ix) << shift,
                                   (uint8_t)(sf.damping + shift - 1), shift);
                }
            }
            bp[0][b] = BY;
            bp[1][b] = BC;
        }
        __syncthreads();
    }

    const int nact = snact;  // 0 when the superblock has no index
    if (prio)
        __builtin_amdgcn_s_setprio(1);
    if (border)
        filter_sb<true, SX, SY, T, CHAIN>(sf, L, U, V, bp, S.dirs, blist, nact, x0, y0, idx & 7, prio);
    else
        filter_sb<false, SX, SY, T, CHAIN>(sf, L, U, V, bp, S.dirs, blist, nact, x0, y0, idx & 7, prio);
    if (CHAIN) {
        // Every wave's output stores have completed before the barrier; then
        // the superblock is published to the restoration tiles that read it.
        wait_vm_stores();
        __syncthreads();
        if (tid == 0)
            __hip_atomic_store(&cs.flags[sb], cs.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}


}  // namespace av1pf
