// av1pf_kernels.hip - gfx950 kernels of the AV1 post-reconstruction filter
// stage: CDEF (AV1 7.15), superres (7.16), loop restoration (7.17).
//
// Integer work, HBM/VALU bound: no MFMA. Every kernel consumes flat work
// records (av1pf_device.h), so the per-frame and batched paths share code.
//
// Reference semantics followed (matteopolak/vaco,
// crates/codec/vaco-codec-av1/src/):
//   CDEF        cdef.rs:79-230, frame walk decode.rs:739-861, 892-896
//   superres    superres.rs:145-216
//   restoration restoration.rs:178-384
#include <hip/hip_runtime.h>

#include "av1pf_device.h"

namespace av1pf {

// ------------------------------------------------------------------------
// Shared helpers
// ------------------------------------------------------------------------

__device__ __forceinline__ int ilog2u(uint32_t v) { return 31 - __clz(v); }

__device__ __forceinline__ int round2i(int v, int s)
{
    return s == 0 ? v : (v + (1 << (s - 1))) >> s;
}

// Work item owning workgroup `b`: the last item whose block_begin <= b.
template <class T>
__device__ __forceinline__ const T &find_item(const T *items, int n, uint32_t b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block_begin <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return items[lo];
}

// ------------------------------------------------------------------------
// CDEF
// ------------------------------------------------------------------------

// DIRECTIONS, cdef.rs:56-66, as (row, column).
__constant__ int8_t kDirDy[8][2] = {{-1, -2}, {0, -1}, {0, 0}, {0, 1},
                                    {1, 2},   {1, 2},  {1, 2}, {1, 2}};
__constant__ int8_t kDirDx[8][2] = {{1, 2}, {1, 2}, {1, 2}, {1, 2},
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

// Direction cost of one 8x8 block, cdef.rs:104-113. All bins and weights
// are compile-time constants, so the 15 partial sums live in registers.
// Costs are bounded by 840 * 64 * 128^2 < 2^31 for in-range samples.
template <int D>
__device__ __forceinline__ uint32_t dir_cost(const int (&v)[64])
{
    int part[15];
#pragma unroll
    for (int b = 0; b < 15; b++)
        part[b] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++)
            part[bin_of(D, i, j)] += v[i * 8 + j];
    uint32_t cost = 0;
#pragma unroll
    for (int b = 0; b < 15; b++) {
        const int cnt = bin_count(D, b);
        if (cnt != 0) {
            uint32_t s = (uint32_t)(part[b] * part[b]);
            cost += s * (uint32_t)(840 / cnt);
        }
    }
    return cost;
}

// Tile geometry in LDS: rows py0-2 .. py0+ph+1, columns px0-8 .. px0+pw+7,
// so every 8x8 block's first column sits on a 16-byte boundary.
constexpr int kCdefRow = 80;
constexpr int kCdefRows = 68;

struct CdefBlk {
    int16_t off[6];      // LDS offsets: primary d1, d2; dir+2 d1, d2; dir+6 d1, d2
    uint16_t pri, sec;
    uint8_t pri_shift, sec_shift, pri_tap0, active;
};

// constrain, cdef.rs:126-133, with the shift pre-computed per block.
__device__ __forceinline__ int constrain(int diff, int strength, int shift)
{
    int ad = abs(diff);
    int t = max(0, strength - (ad >> shift));
    int m = min(ad, t);
    return diff < 0 ? -m : m;
}

__device__ __forceinline__ void cdef_blk_params(CdefBlk &B, int dir, uint32_t pri,
                                                uint32_t sec, int damping, int shift)
{
    B.pri = (uint16_t)pri;
    B.sec = (uint16_t)sec;
    B.pri_shift = (uint8_t)(pri ? max(0, damping - ilog2u(pri)) : 0);
    B.sec_shift = (uint8_t)(sec ? max(0, damping - ilog2u(sec)) : 0);
    B.pri_tap0 = ((pri >> shift) & 1) ? 3 : 4;
    const int d2 = (dir + 2) & 7, d6 = (dir + 6) & 7;
    for (int k = 0; k < 2; k++) {
        B.off[k] = (int16_t)(kDirDy[dir][k] * kCdefRow + kDirDx[dir][k]);
        B.off[2 + k] = (int16_t)(kDirDy[d2][k] * kCdefRow + kDirDx[d2][k]);
        B.off[4 + k] = (int16_t)(kDirDy[d6][k] * kCdefRow + kDirDx[d6][k]);
    }
}

__device__ __forceinline__ void cdef_tap(const uint16_t *t, int ci, int off, int x, int strength,
                                         int shift, int tap, int &sum, int &mn, int &mx)
{
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int p = t[s ? ci + off : ci - off];
        const bool valid = p != kUnavailable;
        const int c = constrain(p - x, strength, shift);
        sum += valid ? tap * c : 0;
        mn = min(mn, p);
        mx = max(mx, valid ? p : x);
    }
}

// One 64x64 luma superblock of one frame, all planes (decode.rs:772-856).
__global__ __launch_bounds__(256) void k_cdef(const CdefFrame *__restrict__ frames, int n_frames)
{
    __shared__ uint16_t lt[3][kCdefRows * kCdefRow];
    __shared__ CdefBlk blk[3][64];
    __shared__ uint32_t cost[8][64];

    const CdefFrame &f = find_item(frames, n_frames, blockIdx.x);
    struct {
        int sbx, sby;
    } tile;
    {
        const uint32_t sb = blockIdx.x - f.block_begin;
        tile.sby = (int)(sb / f.sb_cols);
        tile.sbx = (int)(sb - (uint32_t)tile.sby * f.sb_cols);
    }
    const int tid = threadIdx.x;
    const int bd = f.bit_depth, shift = bd - 8;
    const int maxv = (1 << bd) - 1;
    const int nplanes = f.nplanes;
    const int idx = f.cdef_idx[tile.sby * f.idx_stride + tile.sbx];

    // Phase A: stage every plane's superblock plus its 2-sample border.
    bool bad = false;
    for (int p = 0; p < nplanes; p++) {
        const int sx = p ? f.ss_x : 0, sy = p ? f.ss_y : 0;
        const int px0 = (tile.sbx * 64) >> sx, py0 = (tile.sby * 64) >> sy;
        const int cols = (64 >> sx) + 4, rows = (64 >> sy) + 4;
        const int W = f.width[p], H = f.height[p];
        const uint16_t *src = f.in[p];
        const int st = f.in_stride[p];
        for (int e = tid; e < rows * cols; e += 256) {
            const int r = e / cols, c = e - r * cols;
            const int x = px0 - 2 + c, y = py0 - 2 + r;
            uint16_t v = kUnavailable;
            if (x >= 0 && y >= 0 && x < W && y < H) {
                v = src[(size_t)y * st + x];
                bad |= v > maxv;
            }
            lt[p][r * kCdefRow + c + 6] = v;
        }
    }
    if (__syncthreads_or(bad) && tid == 0)
        atomicOr(f.status, kFlagSampleRange);

    // Phase B: direction search per active 8x8 luma block. Wave w computes
    // directions w and w+4 of all 64 blocks; lane = block.
    const int b = tid & 63, wave = tid >> 6;
    const int bi = b >> 3, bj = b & 7;
    const int bx = tile.sbx * 64 + bj * 8, by = tile.sby * 64 + bi * 8;
    bool active = idx >= 0 && bx < (int)f.width[0] && by < (int)f.height[0] &&
                  f.skip8[(size_t)(by >> 3) * f.skip_stride + (bx >> 3)] == 0;
    if (idx >= 0) {
        if (active) {
            int v[64];
            const uint16_t *base = &lt[0][(2 + bi * 8) * kCdefRow + 8 + bj * 8];
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++)
                    v[i * 8 + j] = (int)(base[i * kCdefRow + j] >> shift) - 128;
            uint32_t c0, c1;
            switch (wave) {
            case 0: c0 = dir_cost<0>(v); c1 = dir_cost<4>(v); break;
            case 1: c0 = dir_cost<1>(v); c1 = dir_cost<5>(v); break;
            case 2: c0 = dir_cost<2>(v); c1 = dir_cost<6>(v); break;
            default: c0 = dir_cost<3>(v); c1 = dir_cost<7>(v); break;
            }
            cost[wave][b] = c0;
            cost[wave + 4][b] = c1;
        }
        __syncthreads();
        if (tid < 64) {
            if (active) {
                int best = 0;
                uint32_t bc = cost[0][b];
#pragma unroll
                for (int d = 1; d < 8; d++) {
                    const uint32_t c = cost[d][b];
                    if (c > bc) { // first maximum wins, cdef.rs:114-119
                        bc = c;
                        best = d;
                    }
                }
                const uint32_t var = (bc - cost[(best + 4) & 7][b]) >> 10;
                const int ix = idx & 7;
                // luma: decode.rs:806-818
                const uint32_t ypri = (uint32_t)f.y_pri[ix] << shift;
                uint32_t ypa = 0;
                if (var != 0) { // adjust_strength, cdef.rs:208-220
                    const uint32_t sc = var >> 6;
                    const uint32_t lg = sc ? min(ilog2u(sc), 12) : 0;
                    ypa = (ypri * (4 + lg) + 8) >> 4;
                }
                cdef_blk_params(blk[0][b], ypri ? best : 0, ypa, (uint32_t)f.y_sec[ix] << shift,
                                (uint8_t)(f.damping + shift), shift);
                blk[0][b].active = 1;
                if (nplanes > 1) {
                    const uint32_t cpri = (uint32_t)f.uv_pri[ix] << shift;
                    int cdir = best; // chroma_direction, cdef.rs:222-230
                    if (f.ss_x && !f.ss_y) {
                        const int m[8] = {7, 0, 2, 4, 5, 6, 6, 6};
                        cdir = m[best];
                    } else if (!f.ss_x && f.ss_y) {
                        const int m[8] = {1, 2, 2, 2, 3, 4, 6, 0};
                        cdir = m[best];
                    }
                    cdef_blk_params(blk[1][b], cpri ? cdir : 0, cpri,
                                    (uint32_t)f.uv_sec[ix] << shift,
                                    (uint8_t)(f.damping + shift - 1), shift);
                    blk[1][b].active = 1;
                    blk[2][b] = blk[1][b];
                }
            } else {
                blk[0][b].active = 0;
                blk[1][b].active = 0;
                blk[2][b].active = 0;
            }
        }
        __syncthreads();
    }

    // Phase C: filter (or pass through) every sample of the superblock.
    for (int p = 0; p < nplanes; p++) {
        const int sx = p ? f.ss_x : 0, sy = p ? f.ss_y : 0;
        const int px0 = (tile.sbx * 64) >> sx, py0 = (tile.sby * 64) >> sy;
        const int lpw = 6 - sx, pw = 64 >> sx, ph = 64 >> sy;
        const int lbw = 3 - sx, lbh = 3 - sy;
        const int W = f.width[p], H = f.height[p];
        uint16_t *dst = f.out[p];
        const int ost = f.out_stride[p];
        const uint16_t *t = lt[p];
        for (int e = tid; e < pw * ph; e += 256) {
            const int r = e >> lpw, c = e & (pw - 1);
            const int x = px0 + c, y = py0 + r;
            if (x >= W || y >= H)
                continue;
            const int ci = (r + 2) * kCdefRow + c + 8;
            const int cur = t[ci];
            int res = cur;
            if (idx >= 0) {
                const CdefBlk &B = blk[p][(r >> lbh) * 8 + (c >> lbw)];
                if (B.active) {
                    int sum = 0, mn = cur, mx = cur;
                    const int pt0 = B.pri_tap0;
                    cdef_tap(t, ci, B.off[0], cur, B.pri, B.pri_shift, pt0, sum, mn, mx);
                    cdef_tap(t, ci, B.off[1], cur, B.pri, B.pri_shift, 6 - pt0, sum, mn, mx);
                    cdef_tap(t, ci, B.off[2], cur, B.sec, B.sec_shift, 2, sum, mn, mx);
                    cdef_tap(t, ci, B.off[3], cur, B.sec, B.sec_shift, 1, sum, mn, mx);
                    cdef_tap(t, ci, B.off[4], cur, B.sec, B.sec_shift, 2, sum, mn, mx);
                    cdef_tap(t, ci, B.off[5], cur, B.sec, B.sec_shift, 1, sum, mn, mx);
                    res = cur + ((8 + sum - (sum < 0)) >> 4);
                    res = min(max(res, mn), mx);
                }
            }
            dst[(size_t)y * ost + x] = (uint16_t)res;
        }
    }
}

// Kernel-level CDEF cases (cdef::find_direction + cdef::filter_block on a
// 12x12 neighbourhood each): one wave per case, lane = output pixel.
struct CdefCase {
    uint32_t width, height;
    uint8_t edges, bit_depth, direction, damping;
    uint16_t primary, secondary;
};

__global__ __launch_bounds__(64) void k_cdef_cases(const uint16_t *__restrict__ inputs,
                                                   const CdefCase *__restrict__ params,
                                                   uint8_t *__restrict__ dirs,
                                                   uint32_t *__restrict__ vars,
                                                   uint16_t *__restrict__ outs, uint32_t n)
{
    __shared__ uint16_t t[12 * kCdefRow];
    __shared__ uint32_t cost[8];
    const uint32_t cs = blockIdx.x;
    if (cs >= n)
        return;
    const int lane = threadIdx.x;
    const CdefCase P = params[cs];
    const int shift = P.bit_depth - 8;
    // Stage as rows of the CDEF tile layout; columns outside the edge mask
    // become unavailable exactly like plane-exterior samples in k_cdef.
    for (int e = lane; e < 144; e += 64) {
        const int r = e / 12, c = e % 12;
        const int ny = r - 2, nx = c - 2;
        const bool ok = !((nx < 0 && !(P.edges & 1)) || (nx >= (int)P.width && !(P.edges & 2)) ||
                          (ny < 0 && !(P.edges & 4)) || (ny >= (int)P.height && !(P.edges & 8)));
        t[r * kCdefRow + c + 6] = ok ? inputs[(size_t)cs * 144 + e] : kUnavailable;
    }
    __syncthreads();
    // Direction search on the full 8x8 at (2,2) (edges never mask it: the
    // reference reads the raw 8x8, av1_cdef.rs:90-95), 8 lanes per direction.
    if (lane < 8) {
        int v[64];
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 8; j++)
                v[i * 8 + j] = (int)(inputs[(size_t)cs * 144 + (i + 2) * 12 + j + 2] >> shift) - 128;
        uint32_t c;
        switch (lane) {
        case 0: c = dir_cost<0>(v); break;
        case 1: c = dir_cost<1>(v); break;
        case 2: c = dir_cost<2>(v); break;
        case 3: c = dir_cost<3>(v); break;
        case 4: c = dir_cost<4>(v); break;
        case 5: c = dir_cost<5>(v); break;
        case 6: c = dir_cost<6>(v); break;
        default: c = dir_cost<7>(v); break;
        }
        cost[lane] = c;
    }
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
    CdefBlk B;
    cdef_blk_params(B, P.direction, P.primary, P.secondary, P.damping, shift);
    const int r = lane >> 3, c = lane & 7;
    uint16_t res = 0;
    if (r < (int)P.height && c < (int)P.width) {
        const int ci = (r + 2) * kCdefRow + c + 8;
        const int cur = t[ci];
        int sum = 0, mn = cur, mx = cur;
        cdef_tap(t, ci, B.off[0], cur, B.pri, B.pri_shift, B.pri_tap0, sum, mn, mx);
        cdef_tap(t, ci, B.off[1], cur, B.pri, B.pri_shift, 6 - B.pri_tap0, sum, mn, mx);
        cdef_tap(t, ci, B.off[2], cur, B.sec, B.sec_shift, 2, sum, mn, mx);
        cdef_tap(t, ci, B.off[3], cur, B.sec, B.sec_shift, 1, sum, mn, mx);
        cdef_tap(t, ci, B.off[4], cur, B.sec, B.sec_shift, 2, sum, mn, mx);
        cdef_tap(t, ci, B.off[5], cur, B.sec, B.sec_shift, 1, sum, mn, mx);
        int v = cur + ((8 + sum - (sum < 0)) >> 4);
        res = (uint16_t)min(max(v, mn), mx);
    }
    outs[(size_t)cs * 64 + lane] = res;
}

// ------------------------------------------------------------------------
// Superres
// ------------------------------------------------------------------------

// UPSCALE_FILTER, superres.rs:25-90 (AV1 spec Upscale_Filter[64][8]).
__constant__ int16_t kUpscale[64][8] = {
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

// One row segment of kSrSeg output samples (superres.rs:173-214).
__global__ __launch_bounds__(kSrSeg) void k_superres(const SrJob *__restrict__ jobs, int n_jobs)
{
    const SrJob &j = find_item(jobs, n_jobs, blockIdx.x);
    const uint32_t local = blockIdx.x - j.block_begin;
    const uint32_t ri = local / j.segs;
    const uint32_t seg = local - ri * j.segs;
    const uint32_t y = j.rows ? j.rows[ri] : ri;
    const int x = (int)(seg * kSrSeg + threadIdx.x);
    if (x >= (int)j.out_width)
        return;
    const int src_x = -(1 << 14) + j.start + x * j.step;
    const int px = src_x >> 14;
    const int phase = (src_x & 0x3fff) >> 8;
    const uint16_t *row = j.in + (size_t)y * j.in_stride;
    const int last = (int)j.in_width - 1;
    int sum = 0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int sx = min(max(px + k - 3, 0), last);
        const int v = row[sx];
        bad |= v > j.max_sample;
        sum += v * kUpscale[phase][k];
    }
    if (bad)
        atomicOr(j.status, kFlagSampleRange);
    const int r = (sum + 64) >> 7;
    j.out[(size_t)y * j.out_stride + x] = (uint16_t)min(max(r, 0), (int)j.max_sample);
}

// ------------------------------------------------------------------------
// Loop restoration
// ------------------------------------------------------------------------

// x_by_xplus1 closed form of restoration.rs:330-336 for 1 <= z <= 254.
struct XByX1 {
    uint16_t v[256];
    constexpr XByX1() : v()
    {
        for (int z = 0; z < 256; z++)
            v[z] = z >= 255 ? 256 : z == 0 ? 1 : (uint16_t)(((z << 8) + z / 2) / (z + 1));
    }
};
__constant__ XByX1 kXByX1 = XByX1();

// SGR_PARAMS, restoration.rs:286-303: {r0, e0, r1, e1}.
__constant__ int kSgr[16][4] = {
    {2, 12, 1, 4},  {2, 15, 1, 6},  {2, 18, 1, 8},  {2, 21, 1, 9},
    {2, 24, 1, 10}, {2, 29, 1, 11}, {2, 36, 1, 12}, {2, 45, 1, 13},
    {2, 56, 1, 14}, {2, 68, 1, 15}, {0, 0, 1, 5},   {0, 0, 1, 8},
    {0, 0, 1, 11},  {0, 0, 1, 14},  {2, 30, 0, 0},  {2, 75, 0, 0},
};

constexpr int kLrS = 72;   // staged source row stride (samples)
constexpr int kLrAB = 66;  // A/B grid row stride

// box_ab, restoration.rs:314-339, on the staged source (centre at S index).
template <int R>
__device__ __forceinline__ void box_ab(const uint16_t *S, int center, int bd, uint32_t scale,
                                       int &a2_out, int &b2_out)
{
    constexpr int n = (2 * R + 1) * (2 * R + 1);
    constexpr uint32_t recip = ((1u << 12) + n / 2) / n;
    int sum = 0, sumsq = 0;
#pragma unroll
    for (int dy = -R; dy <= R; dy++)
#pragma unroll
        for (int dx = -R; dx <= R; dx++) {
            const int v = S[center + dy * kLrS + dx];
            sum += v;
            sumsq += v * v;
        }
    const int a = round2i(sumsq, 2 * (bd - 8));
    const int d = round2i(sum, bd - 8);
    const int p = max(a * n - d * d, 0);
    const uint64_t ps = (uint64_t)(uint32_t)p * scale;
    const uint64_t z = (ps + (1u << 19)) >> 20;
    const int a2 = z >= 255 ? 256 : kXByX1.v[(int)z];
    const uint64_t bb = (uint64_t)(uint32_t)(256 - a2) * (uint32_t)sum * recip;
    a2_out = a2;
    b2_out = (int)((bb + (1u << 11)) >> 12);
}

__global__ __launch_bounds__(256) void k_lr(const LrPlane *__restrict__ planes, int n_planes)
{
    __shared__ uint16_t S[70 * kLrS];
    __shared__ int M[2 * kLrAB * kLrAB];  // Wiener rows, or SGR A|B grids

    const LrPlane &P = find_item(planes, n_planes, blockIdx.x);
    const uint32_t local = blockIdx.x - P.block_begin;
    const uint32_t rsi = local / P.n_cols;
    const LrRowSeg rs = P.rows[rsi];
    const LrColSeg cs = P.cols[local - rsi * P.n_cols];
    LrUnit tl;
    if (P.units)
        tl = P.units[rs.unit_row * P.unit_cols + cs.unit_col];
    else
        tl.kind = 0;
    const int tid = threadIdx.x;
    const int W = P.width, H = P.height;
    const int bd = P.bit_depth, maxv = (1 << bd) - 1;
    const int tw = cs.w, th = rs.h;
    const int x0 = cs.x0, y0 = rs.y0;

    if (tl.kind == 0) { // RestorationUnit::None: copy the CDEF output
        bool bad = false;
        for (int e = tid; e < tw * th; e += 256) {
            const int r = e / tw, c = e - r * tw;
            const int v = P.after[(size_t)(y0 + r) * P.after_stride + x0 + c];
            bad |= v > maxv;
            P.out[(size_t)(y0 + r) * P.out_stride + x0 + c] = (uint16_t)min(v, maxv);
        }
        if (__syncthreads_or(bad) && tid == 0)
            atomicOr(P.status, kFlagSampleRange);
        return;
    }

    // Stage Source::get (restoration.rs:233-244) over rows y0-3 .. y0+th+2,
    // columns x0-3 .. x0+tw+2.
    const int ss = rs.stripe_start, se = rs.stripe_end;
    bool bad = false;
    const int srows = th + 6, scols = tw + 6;
    for (int e = tid; e < srows * scols; e += 256) {
        const int r = e / scols, c = e - r * scols;
        const int xc = min(max(x0 - 3 + c, 0), W - 1);
        const int yc = min(max(y0 - 3 + r, 0), H - 1);
        int v;
        if (yc < ss)
            v = P.before[(size_t)max(yc, ss - 2) * P.before_stride + xc];
        else if (yc > se)
            v = P.before[(size_t)min(yc, se + 2) * P.before_stride + xc];
        else
            v = P.after[(size_t)yc * P.after_stride + xc];
        bad |= v > maxv;
        S[r * kLrS + c] = (uint16_t)v;
    }
    if (__syncthreads_or(bad) && tid == 0)
        atomicOr(P.status, kFlagSampleRange);

    if (tl.kind == 1) {
        // Wiener, restoration.rs:264-283: horizontal rows into M, then columns.
        const int round_h = bd == 12 ? 5 : 3;
        const int offset = 1 << (bd + 7 - round_h - 1);
        const int limit = (1 << (bd + 1 + 7 - round_h)) - 1;
        const int h0 = tl.horizontal[0], h1 = tl.horizontal[1], h2 = tl.horizontal[2];
        const int h3 = 128 - 2 * (h0 + h1 + h2);
        const int v0 = tl.vertical[0], v1 = tl.vertical[1], v2 = tl.vertical[2];
        const int v3 = 128 - 2 * (v0 + v1 + v2);
        for (int e = tid; e < srows * tw; e += 256) {
            const int r = e / tw, c = e - r * tw;
            const uint16_t *s = &S[r * kLrS + c];
            int h = h0 * (s[0] + s[6]) + h1 * (s[1] + s[5]) + h2 * (s[2] + s[4]) + h3 * s[3];
            h = round2i(h, round_h);
            M[r * 64 + c] = min(max(h, -offset), limit - offset);
        }
        __syncthreads();
        for (int e = tid; e < th * tw; e += 256) {
            const int r = e / tw, c = e - r * tw;
            const int *m = &M[r * 64 + c];
            int s = v0 * (m[0] + m[6 * 64]) + v1 * (m[64] + m[5 * 64]) +
                    v2 * (m[2 * 64] + m[4 * 64]) + v3 * m[3 * 64];
            s = round2i(s, 14 - round_h);
            P.out[(size_t)(y0 + r) * P.out_stride + x0 + c] = (uint16_t)min(max(s, 0), maxv);
        }
        return;
    }

    // Self-guided, restoration.rs:341-384. Each thread owns up to 16 output
    // samples; pass 0 (radius r0, odd rows only) then pass 1 (radius r1)
    // fill the A/B grids over centres (y0-1 .. y0+th, x0-1 .. x0+tw).
    const int set = tl.set & 15;
    const int r0 = kSgr[set][0], e0 = kSgr[set][1], r1 = kSgr[set][2], e1 = kSgr[set][3];
    int *A = M, *B = M + kLrAB * kLrAB;
    int f0[16], f1[16];
    const int crows = th + 2, ccols = tw + 2;

    // pass 0
    if (r0) {
        const uint32_t n2e = 25u * 25u * (uint32_t)e0;
        const uint32_t scale = ((1u << 20) + n2e / 2) / n2e;
        for (int e = tid; e < crows * ccols; e += 256) {
            const int cr = e / ccols, cc = e - cr * ccols;
            if (((y0 - 1 + cr) & 1) == 0)
                continue;
            int a2, b2;
            box_ab<2>(S, (cr + 2) * kLrS + cc + 2, bd, scale, a2, b2);
            A[cr * kLrAB + cc] = a2;
            B[cr * kLrAB + cc] = b2;
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int e = tid + i * 256;
        f0[i] = 0;
        if (e < tw * th) {
            const int r = e / tw, c = e - r * tw;
            const int u = S[(r + 3) * kLrS + c + 3];
            if (!r0) {
                f0[i] = u << 4;
            } else if ((y0 + r) & 1) {
                const int k = (r + 1) * kLrAB + c;
                const int a = 5 * A[k] + 6 * A[k + 1] + 5 * A[k + 2];
                const int bsum = 5 * B[k] + 6 * B[k + 1] + 5 * B[k + 2];
                f0[i] = (a * u + bsum + (1 << 7)) >> 8;
            } else {
                const int k = r * kLrAB + c, k2 = k + 2 * kLrAB;
                const int a = 5 * (A[k] + A[k + 2] + A[k2] + A[k2 + 2]) + 6 * (A[k + 1] + A[k2 + 1]);
                const int bsum = 5 * (B[k] + B[k + 2] + B[k2] + B[k2 + 2]) + 6 * (B[k + 1] + B[k2 + 1]);
                f0[i] = (a * u + bsum + (1 << 8)) >> 9;
            }
        }
    }
    if (r1) {
        __syncthreads(); // A/B of pass 0 fully consumed
        const uint32_t n2e = 9u * 9u * (uint32_t)e1;
        const uint32_t scale = ((1u << 20) + n2e / 2) / n2e;
        for (int e = tid; e < crows * ccols; e += 256) {
            const int cr = e / ccols, cc = e - cr * ccols;
            int a2, b2;
            box_ab<1>(S, (cr + 2) * kLrS + cc + 2, bd, scale, a2, b2);
            A[cr * kLrAB + cc] = a2;
            B[cr * kLrAB + cc] = b2;
        }
        __syncthreads();
    }
    const int w0 = tl.xqd[0], w1 = tl.xqd[1];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int e = tid + i * 256;
        if (e < tw * th) {
            const int r = e / tw, c = e - r * tw;
            const int u = S[(r + 3) * kLrS + c + 3];
            if (!r1) {
                f1[i] = u << 4;
            } else {
                const int k = r * kLrAB + c, k1 = k + kLrAB, k2 = k + 2 * kLrAB;
                const int a = 4 * (A[k + 1] + A[k1] + A[k1 + 1] + A[k1 + 2] + A[k2 + 1]) +
                              3 * (A[k] + A[k + 2] + A[k2] + A[k2 + 2]);
                const int bsum = 4 * (B[k + 1] + B[k1] + B[k1 + 1] + B[k1 + 2] + B[k2 + 1]) +
                                 3 * (B[k] + B[k + 2] + B[k2] + B[k2 + 2]);
                f1[i] = (a * u + bsum + (1 << 8)) >> 9;
            }
            int v = w0 * f0[i] + w1 * (u << 4) + (128 - w0 - w1) * f1[i];
            v = (v + (1 << 10)) >> 11;
            P.out[(size_t)(y0 + r) * P.out_stride + x0 + c] = (uint16_t)min(max(v, 0), maxv);
        }
    }
}

// ------------------------------------------------------------------------
// Host launch wrappers
// ------------------------------------------------------------------------

hipError_t launch_cdef(const CdefFrame *frames, int n, uint32_t blocks, hipStream_t s)
{
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cdef, dim3(blocks), dim3(256), 0, s, frames, n);
    return hipGetLastError();
}

hipError_t launch_superres(const SrJob *jobs, int n, uint32_t blocks, hipStream_t s)
{
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_superres, dim3(blocks), dim3(kSrSeg), 0, s, jobs, n);
    return hipGetLastError();
}

hipError_t launch_lr(const LrPlane *planes, int n, uint32_t blocks, hipStream_t s)
{
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_lr, dim3(blocks), dim3(256), 0, s, planes, n);
    return hipGetLastError();
}

hipError_t launch_cdef_cases(const uint16_t *inputs, const void *params, uint8_t *dirs,
                             uint32_t *vars, uint16_t *outs, uint32_t n, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_cdef_cases, dim3(n), dim3(64), 0, s, inputs,
                       (const CdefCase *)params, dirs, vars, outs, n);
    return hipGetLastError();
}

}  // namespace av1pf
