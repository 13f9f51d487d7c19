// av1pf_lr_body.h - loop restoration (AV1 7.17) device code for gfx950,
// shared by the restoration kernel and the chained CDEF -> restoration
// kernel.
//
// Reference semantics (matteopolak/vaco, crates/codec/vaco-codec-av1/src/
// restoration.rs): restore_plane 178-207, Source 209-245, wiener 259-283,
// box_ab 314-339, guided_pass 341-371, self_guided 373-384; copy_visible_plane
// decode.rs:673-681.
//
// A tile is at most 64x64 output samples inside one restoration stripe and
// one restoration unit, so its filter type is uniform. The host sorts tiles
// by type and one launch of k_lr (8 threads per row of the tuned tile
// height: 512, 256 or 128) covers all of them; each
// workgroup branches once on its tile's unit type. The host orders the list
// (tuning): either self-guided tiles, then Wiener, then copies, so the
// longest tiles start first, or plane order numbered XCD-contiguously, so
// neighbouring tiles share halo lines in one L2. Each kind is a branch-free
// path of its own:
//   copy        RestorationUnit::None and planes without restoration
//   Wiener      ~19 KB LDS at 64 rows: staged source, 16-bit row pass, column pass
//   self-guided ~39 KB LDS at 64 rows: staged source, column sums of x and x^2, box
//               sums per centre, A/B grids (built in registers and written
//               over the column sums), 3x3 weighted combination
// Each staged window follows Source::get: CDEF output inside the stripe,
// pre-CDEF rows across its borders, edge clamping.
#pragma once

#include "av1pf_common.h"
#include "av1pf_superres_body.h"

namespace av1pf {

constexpr int kS = 72;   // staged row stride (samples), 70 used
constexpr int kG = 68;   // A/B grid row stride (66 used; rows 16-byte aligned)

// SGR_PARAMS, restoration.rs:286-303: {r0, e0, r1, e1}.
constexpr int kSgr[16][4] = {
    {2, 12, 1, 4},  {2, 15, 1, 6},  {2, 18, 1, 8},  {2, 21, 1, 9},
    {2, 24, 1, 10}, {2, 29, 1, 11}, {2, 36, 1, 12}, {2, 45, 1, 13},
    {2, 56, 1, 14}, {2, 68, 1, 15}, {0, 0, 1, 5},   {0, 0, 1, 8},
    {0, 0, 1, 11},  {0, 0, 1, 14},  {2, 30, 0, 0},  {2, 75, 0, 0},
};

// One guided pass of a parameter set: radius (0: pass off), the scale of
// restoration.rs:316-317 for n = (2r+1)^2 and its eps, and the clamp
// threshold of ab_of (the smallest p whose z reaches 255).
struct SgrPass {
    uint32_t r, scale, thr, pad;
};

constexpr SgrPass sgr_pass(int r, int eps)
{
    if (r == 0)
        return SgrPass{0, 0, 0, 0};
    const uint32_t n = (uint32_t)((2 * r + 1) * (2 * r + 1));
    const uint32_t n2e = n * n * (uint32_t)eps;
    const uint32_t scale = ((1u << 20) + n2e / 2) / n2e;
    const uint32_t thr = ((255u << 20) - (1u << 19) + scale - 1) / scale;
    return SgrPass{(uint32_t)r, scale, thr, 0};
}

struct SgrPasses {
    SgrPass p[16][2];
};

constexpr SgrPasses sgr_passes()
{
    SgrPasses t{};
    for (int s = 0; s < 16; s++) {
        t.p[s][0] = sgr_pass(kSgr[s][0], kSgr[s][1]);
        t.p[s][1] = sgr_pass(kSgr[s][2], kSgr[s][3]);
    }
    return t;
}

// What ab_of relies on, for every pass in use: z = (p * scale + 2^19) >> 20
// reaches 255 first at p = thr (so clamping p to thr keeps every z that
// selects a table entry, and a2 = 256 beyond it), and thr * scale + 2^19
// fits 32 bits.
constexpr bool sgr_passes_exact()
{
    const SgrPasses t = sgr_passes();
    for (int s = 0; s < 16; s++)
        for (int i = 0; i < 2; i++) {
            const SgrPass g = t.p[s][i];
            if (g.r == 0)
                continue;
            const uint64_t top = (uint64_t)g.thr * g.scale + (1u << 19), below = top - g.scale;
            if (top >= (1ull << 32) || (top >> 20) < 255 || (below >> 20) >= 255)
                return false;
        }
    return true;
}
static_assert(sgr_passes_exact(), "self-guided clamp thresholds are exact in 32 bits");

// Both passes of each set, indexed by a wave-uniform set (scalar loads).
static __constant__ SgrPasses kSgrPasses = sgr_passes();

// One tile of a list entry: row segment << 48 | column segment << 32 | item
// (LrPlane index) << 2 | unit type (kUnitNone, kUnitWiener, kUnitSelfGuided;
// kUnitNone also for planes without restoration). The type in the entry
// settles the workgroup's path before any descriptor is read; the unit's
// coefficients are read by the Wiener and self-guided paths with the staging
// loads (UnitStage).
struct Tile {
    const LrPlane *P;
    int kind;
    LrRowSeg rs;
    LrColSeg cs;
    uint32_t unit_ix;  // index of the tile's unit in P->units (kind != kUnitNone only)
    // The plane's byte fields, read as two words (scalar loads; byte loads
    // would be vector loads, each waited on before its first use).
    uint32_t b0, b1;
    __device__ __forceinline__ int bit_depth() const { return (int)(b0 & 255u); }
    __device__ __forceinline__ bool vec_ok() const { return ((b0 >> 16) & 255u) != 0; }
    __device__ __forceinline__ bool out_u8() const { return (b0 >> 24) != 0; }
    __device__ __forceinline__ bool out_vec() const { return (b1 & 255u) != 0; }
};
static_assert(offsetof(LrPlane, bit_depth) % 4 == 0 &&
              offsetof(LrPlane, ss_y) == offsetof(LrPlane, bit_depth) + 1 &&
              offsetof(LrPlane, vec_ok) == offsetof(LrPlane, bit_depth) + 2 &&
              offsetof(LrPlane, out_u8) == offsetof(LrPlane, bit_depth) + 3 &&
              offsetof(LrPlane, out_vec) == offsetof(LrPlane, bit_depth) + 4 &&
              offsetof(LrPlane, out_vec) + 4 <= sizeof(LrPlane),
              "LrPlane byte fields are read as two words");

// The descriptor words a tile needs before it can stage: its segment
// tables, the unit grid width and the byte fields.
struct PlaneHead {
    const LrRowSeg *rows;
    const LrColSeg *cols;
    uint32_t unit_cols, b0, b1;
    __device__ __forceinline__ explicit PlaneHead(const LrPlane &P)
        : rows(P.rows), cols(P.cols), unit_cols(P.unit_cols)
    {
        __builtin_memcpy(&b0, &P.bit_depth, 4);
        __builtin_memcpy(&b1, &P.out_vec, 4);
    }
};

// The tile of workgroup wg: list entry wg, or, without a list (tiles in
// plane order), tile wg of the n_items plane items: the item whose
// block_begin <= wg, its tile wg - block_begin in row-major order of its row
// and column segments, with the kind of the tile's unit. With a list, the
// entry names the segments (no division by the segment count before their
// loads).
__device__ __forceinline__ Tile fetch_tile(const LrPlane *planes, uint32_t n_items, const uint64_t *list,
                                           uint32_t wg)
{
    Tile t;
    uint32_t ri, ci;
    const LrRowSeg *rows;
    const LrColSeg *cols;
    uint32_t unit_cols;
    if (list) {
        const uint64_t e = list[wg];
        const PlaneHead h(planes[(uint32_t)e >> 2]);
        rows = h.rows;
        cols = h.cols;
        unit_cols = h.unit_cols;
        t.b0 = h.b0;
        t.b1 = h.b1;
        t.kind = (int)(e & 3u);
        t.P = &planes[(uint32_t)e >> 2];
        ri = (uint32_t)(e >> 48);
        ci = (uint32_t)(e >> 32) & 0xffffu;
    } else {
        t.P = &find_item(planes, (int)n_items, wg);
        const uint32_t local = wg - t.P->block_begin;
        ri = local / t.P->n_cols;
        ci = local - ri * t.P->n_cols;
        const PlaneHead h(*t.P);
        rows = h.rows;
        cols = h.cols;
        unit_cols = h.unit_cols;
        t.b0 = h.b0;
        t.b1 = h.b1;
    }
    t.rs = glb(rows)[ri];
    t.cs = glb(cols)[ci];
    t.unit_ix = t.rs.unit_row * unit_cols + t.cs.unit_col;
    if (!list) {
        const uint8_t k = t.P->units ? glb(t.P->units)[t.unit_ix].kind : kUnitNone;
        t.kind = k == kUnitWiener ? kUnitWiener : k == kUnitSelfGuided ? kUnitSelfGuided : kUnitNone;
    }
    return t;
}

// Output of tile t: its first sample's address (u16 or, for an 8-bit frame
// output, byte elements) and the row stride, so that per-sample addresses
// are 32-bit offsets from it.
struct TileOut {
    gptr<uint16_t> base16;
    gptr<uint8_t> base8;
    uint32_t stride;
    bool u8;
};

__device__ __forceinline__ TileOut tile_out(const Tile &t, int x0, int y0)
{
    const LrPlane &P = *t.P;
    const size_t o = (size_t)y0 * P.out_stride + x0;
    return TileOut{glb(P.out) + o, reinterpret_cast<gptr<uint8_t>>(glb(P.out)) + o, P.out_stride,
                   t.out_u8()};
}

// Store output samples v0, v1 at row r, columns c (even) and c + 1 of the
// tile as one store; the output must be 8-byte aligned with a row stride that
// is a multiple of 4 (Tile::out_vec), so the pair is aligned.
__device__ __forceinline__ void put2(const TileOut &o, int r, int c, int v0, int v1)
{
    const uint32_t i = __umul24((uint32_t)r, o.stride) + (uint32_t)c;
    if (o.u8)
        __builtin_nontemporal_store((uint16_t)(v0 | (v1 << 8)), reinterpret_cast<gptr<uint16_t>>(o.base8 + i));
    else
        __builtin_nontemporal_store((uint32_t)v0 | ((uint32_t)v1 << 16),
                                    reinterpret_cast<gptr<uint32_t>>(o.base16 + i));
}

// Store output sample v at row r, column c of the tile. Outputs are stored
// non-temporally: nothing in the launch reads them, and lines left dirty in
// the L2 would be written back by the release at the launch's end, which
// the next launch waits for.
__device__ __forceinline__ void put(const TileOut &o, int r, int c, int v)
{
    const uint32_t i = __umul24((uint32_t)r, o.stride) + (uint32_t)c;
    if (o.u8)
        __builtin_nontemporal_store((uint8_t)v, o.base8 + i);
    else
        __builtin_nontemporal_store((uint16_t)v, o.base16 + i);
}

// Stage Source::get (restoration.rs:233-244) over rows y0-3 .. y0+th+2 and
// columns x0-3 .. x0+tw+2 into S (row stride kS). Returns whether a sample
// above the bit depth was read. Source::get's row for a plane row y: y
// clamped to the plane, then, outside the stripe, clamped to the stripe's
// borders +-2 and read from the pre-CDEF image; inside it, from the CDEF
// output. Both clamps together are one clamp to [start - 2, end + 2]. Every
// thread issues all of its loads before it stores any of them; positions
// past the window load a position inside it and are not stored.
// Source window row of the folded superres staging (samples; 16-byte rows):
// the taps of 70 output columns span at most 78 source samples, because the
// step is below one source sample per output.
constexpr int kSrWinStride = 88;

// LDS the folded superres staging uses past the staged source: the source
// window of every staged row, the same shifted by one sample, and the filter
// table.
template <int ROWS>
constexpr int sr_stage_lds() { return 2 * (ROWS + 6) * kSrWinStride * 2 + 64 * 8 * 2; }

// Source::get over the upscaled images with superres folded into the
// staging: rows as stage_source picks them (pre-CDEF outside the stripe,
// CDEF output inside, both before upscaling); the sample at plane column X
// (clamped to the visible width) is upscale_plane's output there
// (superres.rs:173-214): eight taps at (src_x >> 14) - 3 .., edge-clamped to
// the padded width, weighted by the phase's filter, rounded by 7 bits and
// clipped. The workgroup stages, per row, the union of its columns' taps
// (the window lo .. hi, at most 78 samples, from at most 3 samples before
// lo) and a copy shifted by one sample
// in `scratch` (sr_stage_lds bytes), so that a column's eight taps are four
// aligned sample pairs summed with v_dot2; then each thread filters one
// column over several rows (position, phase and filter depend on the column
// only). Returns whether a source sample above maxv was read (the window
// holds exactly the samples the taps read).
template <int ROWS, int THREADS>
__device__ __forceinline__ bool stage_source_sr(uint16_t *S, const Tile &t, int maxv, uint8_t *scratch)
{
    const LrPlane &P = *t.P;
    const int W = P.width, H = P.height;
    const int x0 = t.cs.x0, y0 = t.rs.y0, ss = t.rs.stripe_start, se = t.rs.stripe_end;
    const int srows = t.rs.h + 6, scols = t.cs.w + 6;
    const int step = P.sr_step, start = P.sr_start, last = (int)P.sr_in_width - 1;
    uint16_t *W0 = reinterpret_cast<uint16_t *>(scratch);
    uint16_t *W1 = W0 + (ROWS + 6) * kSrWinStride;  // window index i at i + 1
    int16_t *tab = reinterpret_cast<int16_t *>(W1 + (ROWS + 6) * kSrWinStride);
    const gptr<const uint16_t> before = glb(P.before), after = glb(P.after);
    const uint32_t bst = P.before_stride, ast = P.after_stride;
    auto row_of = [&](int r) {
        const int yc = min(max(y0 - 3 + r, 0), H - 1);
        const bool out = yc < ss || yc > se;
        const int yr = min(max(yc, ss - 2), se + 2);
        return (out ? before : after) + (size_t)(uint32_t)yr * (out ? bst : ast);  // 0 <= yr < H
    };
    // -(1 << 14) + initial_subpel_x + x * step of staged column c
    auto pos = [&](int c) { return -(1 << 14) + start + min(max(x0 - 3 + c, 0), W - 1) * step; };
    const int lo = (pos(0) >> 14) - 3, hi = (pos(scols - 1) >> 14) + 4;  // hi - lo <= 77
    const int tid = (int)threadIdx.x;
    for (int i = tid; i < 64 * 8; i += THREADS)
        tab[i] = kUpscale[i >> 3][i & 7];
    // Window samples lo .. hi of every staged row (window index = sample -
    // org), every load of a thread issued before any is stored. Inside the
    // padded width with aligned rows, 8-byte loads of four samples from org =
    // lo rounded down to a multiple of 4 (at most 21 per row); else one
    // edge-clamped sample per load from org = lo.
    bool bad = false;
    const int org4 = lo & ~3, groups = (hi - org4) / 4 + 1;
    const bool vec = t.vec_ok() && lo >= 0 && org4 + 4 * groups - 1 <= last;
    const int org = vec ? org4 : lo;
    if (vec) {
        constexpr int kG4 = 21, N = (ROWS + 6) * kG4, K = (N + THREADS - 1) / THREADS;
        uint2 q[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = min(tid + k * THREADS, N - 1);
            const int r = min(e / kG4, srows - 1), g = min(e - (e / kG4) * kG4, groups - 1);
            const uint64_t w = *reinterpret_cast<gptr<const uint64_t>>(row_of(r) + org + 4 * g);
            q[k] = {(uint32_t)w, (uint32_t)(w >> 32)};
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = tid + k * THREADS;
            const int r = e / kG4, g = e - r * kG4;
            if (e >= N || r >= srows)
                break;
            if (g < groups) {
                const uint2 w = q[k];
                const int s0 = org + 4 * g;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int v = (int)(((u < 2 ? w.x : w.y) >> (16 * (u & 1))) & 0xffffu);
                    bad |= s0 + u >= lo && s0 + u <= hi && v > maxv;
                }
                uint16_t *d0 = static_cast<uint16_t *>(__builtin_assume_aligned(W0 + r * kSrWinStride + 4 * g, 8));
                *reinterpret_cast<uint2 *>(d0) = w;
                uint16_t *d1 = W1 + r * kSrWinStride + 4 * g + 1;  // window index i at i + 1
                d1[0] = (uint16_t)w.x;
                *reinterpret_cast<uint32_t *>(d1 + 1) = (w.x >> 16) | (w.y << 16);
                d1[3] = (uint16_t)(w.y >> 16);
            }
        }
    } else {
        const int wn = hi - lo + 1;  // <= 78
        constexpr int kWin = 80, N = (ROWS + 6) * kWin, K = (N + THREADS - 1) / THREADS;
        uint16_t q[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = min(tid + k * THREADS, N - 1);
            const int r = min(e / kWin, srows - 1), i = min(e - (e / kWin) * kWin, wn - 1);
            q[k] = row_of(r)[min(max(lo + i, 0), last)];
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = tid + k * THREADS;
            const int r = e / kWin, i = e - r * kWin;
            if (e >= N || r >= srows)
                break;
            if (i < wn) {
                bad |= q[k] > maxv;
                W0[r * kSrWinStride + i] = q[k];
                W1[r * kSrWinStride + i + 1] = q[k];
            }
        }
    }
    __syncthreads();
    constexpr int kC = 72, kR = THREADS / kC;  // column slots, row lanes
    const int c = tid % kC, rl = tid / kC;
    if (rl < kR && c < scols) {
        const int p = pos(c);
        const int b = (p >> 14) - 3 - org;  // window index of the first tap: 0 <= b, b + 7 <= hi - org
        const uint4 fq = *reinterpret_cast<const uint4 *>(&tab[((p & 0x3fff) >> 8) * 8]);
        const s16x2 f0 = __builtin_bit_cast(s16x2, fq.x), f1 = __builtin_bit_cast(s16x2, fq.y);
        const s16x2 f2 = __builtin_bit_cast(s16x2, fq.z), f3 = __builtin_bit_cast(s16x2, fq.w);
        const uint16_t *src = (b & 1) ? W1 + b + 1 : W0 + b;
        for (int r = rl; r < srows; r += kR) {
            const uint32_t *w = reinterpret_cast<const uint32_t *>(src + r * kSrWinStride);
            int sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[0]), f0, 0, false);
            sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[1]), f1, sum, false);
            sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[2]), f2, sum, false);
            sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2, w[3]), f3, sum, false);
            S[r * kS + c] = (uint16_t)min(max((sum + 64) >> 7, 0), maxv);
        }
    }
    return bad;
}

template <int ROWS, int THREADS>
__device__ __forceinline__ bool stage_source(uint16_t *S, const Tile &t, int maxv, uint8_t *scratch)
{
    if (t.P->sr_step)
        return stage_source_sr<ROWS, THREADS>(S, t, maxv, scratch);
    const LrPlane &P = *t.P;
    const int W = P.width, H = P.height;
    const int x0 = t.cs.x0, y0 = t.rs.y0, ss = t.rs.stripe_start, se = t.rs.stripe_end;
    const int srows = t.rs.h + 6, scols = t.cs.w + 6;
    const gptr<const uint16_t> before = glb(P.before), after = glb(P.after);
    const uint32_t bst = P.before_stride, ast = P.after_stride;
    auto row_of = [&](int r) {
        const int yc = min(max(y0 - 3 + r, 0), H - 1);
        const bool out = yc < ss || yc > se;
        const int yr = min(max(yc, ss - 2), se + 2);
        return (out ? before : after) + (size_t)(uint32_t)yr * (out ? bst : ast);  // 0 <= yr < H
    };
    bool bad = false;
    if (t.vec_ok() && t.cs.w == 64 && x0 >= 4 && x0 + 68 <= W) {
        // Columns x0-4 .. x0+67 all inside the plane: 8-byte loads of four
        // samples (the column left of the window is loaded and dropped).
        constexpr int V = 18, N = (ROWS + 6) * V, K = (N + THREADS - 1) / THREADS;
        uint2 q[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = min((int)threadIdx.x + k * THREADS, N - 1);
            const int r = min(e / V, srows - 1), v = e - (e / V) * V;
            const uint64_t w = *reinterpret_cast<gptr<const uint64_t>>(row_of(r) + x0 - 4 + 4 * v);
            q[k] = {(uint32_t)w, (uint32_t)(w >> 32)};
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int e = (int)threadIdx.x + k * THREADS;
            const int r = e / V, v = e - r * V;
            if (e >= N || r >= srows)
                break;
            const uint2 w = q[k];
            const uint32_t m = max(max(w.x & 0xffffu, w.x >> 16), max(w.y & 0xffffu, w.y >> 16));
            bad |= (int)m > maxv;
            uint16_t *d = S + r * kS + 4 * v - 1;  // column c of S is x0 - 3 + c
            if (v)
                d[0] = (uint16_t)w.x;
            d[1] = (uint16_t)(w.x >> 16);
            d[2] = (uint16_t)w.y;
            d[3] = (uint16_t)(w.y >> 16);
        }
        return bad;
    }
    constexpr int N = (ROWS + 6) * kS, K = (N + THREADS - 1) / THREADS;
    uint16_t q[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = min((int)threadIdx.x + k * THREADS, N - 1);
        const int r = min(e / kS, srows - 1), c = min(e - (e / kS) * kS, scols - 1);
        const int xc = min(max(x0 - 3 + c, 0), W - 1);
        q[k] = row_of(r)[xc];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int e = (int)threadIdx.x + k * THREADS;
        const int r = e / kS, c = e - r * kS;
        if (e >= N || r >= srows)
            break;
        if (c < scols) {
            bad |= q[k] > maxv;
            S[r * kS + c] = q[k];
        }
    }
    return bad;
}

// The tile's RestorationUnit, fetched alongside the staging loads: lanes
// 0-8 each load one 16-bit word of it (per-lane addresses, so nothing waits
// for them before the staging loads are issued), store it to LDS after the
// staged samples, and every thread reads the unit back after the staging
// barrier.
struct UnitStage {
    static constexpr int kWords = sizeof(LrUnit) / 2;
    static_assert(sizeof(LrUnit) % 2 == 0 && alignof(LrUnit) >= 2, "unit read as 16-bit words");
    uint32_t w = 0;
    __device__ __forceinline__ void load(const Tile &t)
    {
        if (threadIdx.x < kWords)
            w = reinterpret_cast<gptr<const uint16_t>>(glb(t.P->units) + t.unit_ix)[threadIdx.x];
    }
    __device__ __forceinline__ void store(uint16_t *su) const
    {
        if (threadIdx.x < kWords)
            su[threadIdx.x] = (uint16_t)w;
    }
};

// ---------------------------------------------------------------- copy --

// Workgroup size and LDS layout of k_lr for tiles of at most ROWS rows (64,
// 32 or 16; the host cuts tiles to the tuned height): 8 threads per tile
// row, so that the self-guided strips (64 columns x 8 rows) and the Wiener
// column blocks (32 column pairs x 4 rows) cover the tile with every thread.
template <int ROWS>
struct LrShape {
    static constexpr int kThreads = 8 * ROWS;
    static constexpr int kSrcBytes = (ROWS + 6) * kS * 2;                // staged source
    static constexpr int kWienerLds = kSrcBytes + (ROWS + 6) * 64 * 2;   // + row-filtered M
    static constexpr int kXBytes = (ROWS + 2) * kS * 3 / 2 * 4;          // VS + VQ, then A/B
    static constexpr int kSgrLds = kSrcBytes + kXBytes + 256 * 2;        // + x/(x+1) table
    static constexpr int kSrLds = kSrcBytes + sr_stage_lds<ROWS>();      // staging with superres folded in
    static constexpr int kMax2 = kSgrLds > kWienerLds ? kSgrLds : kWienerLds;
    // The staged unit (UnitStage) sits past every path's LDS.
    static constexpr int kUnitLds = ((kMax2 > kSrLds ? kMax2 : kSrLds) + 15) & ~15;
    static constexpr int kLds = kUnitLds + 32;  // the workgroup's LDS
};

template <int T>
__device__ __forceinline__ void lr_copy(const Tile &t)
{
    const LrPlane &P = *t.P;
    const int maxv = (1 << t.bit_depth()) - 1;
    const int tw = t.cs.w, th = t.rs.h, x0 = t.cs.x0, y0 = t.rs.y0;
    const TileOut o = tile_out(t, x0, y0);
    bool bad = false;
    if (t.vec_ok() && t.out_vec() && tw == 64 && (x0 & 3) == 0) {
        // Four samples per load and store, every group 8-byte aligned in
        // both planes.
#pragma unroll 2
        for (int e = threadIdx.x; e < 64 * 16; e += T) {
            const int r = e >> 4, c = (e & 15) * 4;
            if (r >= th)
                break;
            const uint64_t w = *reinterpret_cast<gptr<const uint64_t>>(
                glb(P.after) + (size_t)(y0 + r) * P.after_stride + x0 + c);
            const uint32_t lo = (uint32_t)w, hi = (uint32_t)(w >> 32);
            const uint32_t m = max(max(lo & 0xffffu, lo >> 16), max(hi & 0xffffu, hi >> 16));
            bad |= (int)m > maxv;
            const uint32_t i = __umul24((uint32_t)r, o.stride) + (uint32_t)c;
            if (o.u8)  // samples <= 255 when none is out of range
                __builtin_nontemporal_store(
                    (lo & 0xffu) | ((lo >> 8) & 0xff00u) | ((hi & 0xffu) << 16) | ((hi >> 16) << 24),
                    reinterpret_cast<gptr<uint32_t>>(o.base8 + i));
            else
                __builtin_nontemporal_store(w, reinterpret_cast<gptr<uint64_t>>(o.base16 + i));
        }
        if (__syncthreads_or(bad) && threadIdx.x == 0)
            raise_status(P.status, kFlagSampleRange);
        return;
    }
#pragma unroll 8
    for (int e = threadIdx.x; e < 64 * 64; e += T) {
        const int r = e >> 6, c = e & 63;
        if (r < th && c < tw) {
            const int v = glb(P.after)[(size_t)(y0 + r) * P.after_stride + x0 + c];
            bad |= v > maxv;
            put(o, r, c, min(v, maxv));
        }
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0)
        raise_status(P.status, kFlagSampleRange);
}

// -------------------------------------------------------------- Wiener --


template <int ROWS>
__device__ __forceinline__ void lr_wiener(const Tile &t, uint8_t *lds)
{
    using Sh = LrShape<ROWS>;
    constexpr int T = Sh::kThreads;
    uint16_t *S = reinterpret_cast<uint16_t *>(lds);
    int16_t *M = reinterpret_cast<int16_t *>(lds + Sh::kSrcBytes);
    const LrPlane &P = *t.P;
    const int bd = t.bit_depth(), maxv = (1 << bd) - 1;
    const int tw = t.cs.w, th = t.rs.h, x0 = t.cs.x0, y0 = t.rs.y0;
    const TileOut o = tile_out(t, x0, y0);
    const int tid = threadIdx.x;
    uint16_t *su = reinterpret_cast<uint16_t *>(lds + Sh::kUnitLds);
    UnitStage us;
    us.load(t);
    const bool bad = stage_source<ROWS, T>(S, t, maxv, lds + Sh::kSrcBytes);
    us.store(su);
    if (__syncthreads_or(bad) && tid == 0)
        raise_status(P.status, kFlagSampleRange);
    const LrUnit u = *reinterpret_cast<const LrUnit *>(su);

    // wiener, restoration.rs:264-283: rows, then columns. The intermediate
    // is clamped to [-2^(bd+6-rh), 2^(bd+8-rh) - 1 - 2^(bd+6-rh)], which
    // fits int16 for bd <= 12.
    const int round_h = bd == 12 ? 5 : 3;
    const int offset = 1 << (bd + 7 - round_h - 1);
    const int limit = (1 << (bd + 1 + 7 - round_h)) - 1;
    const int h0 = u.horizontal[0], h1 = u.horizontal[1], h2 = u.horizontal[2];
    const int h3 = 128 - 2 * (h0 + h1 + h2);
    const int v0 = u.vertical[0], v1 = u.vertical[1], v2 = u.vertical[2];
    const int v3 = 128 - 2 * (v0 + v1 + v2);
    const int srows = th + 6;
    // Rows: four adjacent outputs per item from ten staged samples (two
    // 8-byte and one 4-byte LDS read), stored as one 8-byte M write.
    for (int e = tid; e < srows * 16; e += T) {
        const int r = e >> 4, c = (e & 15) * 4;
        const uint16_t *s = &S[r * kS + c];
        const uint2 a = *reinterpret_cast<const uint2 *>(s);
        const uint2 b = *reinterpret_cast<const uint2 *>(s + 4);
        const uint32_t z = *reinterpret_cast<const uint32_t *>(s + 8);
        const int q[10] = {(int)(a.x & 0xffffu), (int)(a.x >> 16), (int)(a.y & 0xffffu), (int)(a.y >> 16),
                           (int)(b.x & 0xffffu), (int)(b.x >> 16), (int)(b.y & 0xffffu), (int)(b.y >> 16),
                           (int)(z & 0xffffu),   (int)(z >> 16)};
        uint32_t m[2];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int h = h0 * (q[j] + q[j + 6]) + h1 * (q[j + 1] + q[j + 5]) + h2 * (q[j + 2] + q[j + 4]) +
                    h3 * q[j + 3];
            h = round2i(h, round_h);
            const uint32_t v = (uint16_t)(int16_t)min(max(h, -offset), limit - offset);
            if (j & 1)
                m[j >> 1] |= v << 16;
            else
                m[j >> 1] = v;
        }
        *reinterpret_cast<uint2 *>(&M[r * 64 + c]) = make_uint2(m[0], m[1]);
    }
    __syncthreads();
    // Columns: a 4-row x 2-column block per thread (ROWS / 4 x 32 blocks
    // cover the tile) from ten 4-byte M reads, each holding both columns.
    {
        const int cb = tid & 31, rb = tid >> 5;
        const int c = 2 * cb, r0 = 4 * rb;
        const bool pair = t.out_vec();
        int lo[10], hi[10];
#pragma unroll
        for (int k = 0; k < 10; k++) {
            const uint32_t w = *reinterpret_cast<const uint32_t *>(&M[(r0 + k) * 64 + c]);
            lo[k] = (int16_t)(w & 0xffffu);
            hi[k] = (int16_t)(w >> 16);
        }
        if (c < tw) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = r0 + j;
                if (r >= th)
                    break;
                int s0 = v0 * (lo[j] + lo[j + 6]) + v1 * (lo[j + 1] + lo[j + 5]) +
                         v2 * (lo[j + 2] + lo[j + 4]) + v3 * lo[j + 3];
                int s1 = v0 * (hi[j] + hi[j + 6]) + v1 * (hi[j + 1] + hi[j + 5]) +
                         v2 * (hi[j + 2] + hi[j + 4]) + v3 * hi[j + 3];
                s0 = min(max(round2i(s0, 14 - round_h), 0), maxv);
                s1 = min(max(round2i(s1, 14 - round_h), 0), maxv);
                // one store per pair: a store instruction then writes whole
                // row segments, not every other sample
                if (pair && c + 1 < tw) {
                    put2(o, r, c, s0, s1);
                } else {
                    put(o, r, c, s0);
                    if (c + 1 < tw)
                        put(o, r, c + 1, s1);
                }
            }
        }
    }
}

// --------------------------------------------------------- self-guided --


// A/B of one box from its sums (restoration.rs:314-339), packed as
// b2 << 9 | a2 (a2 <= 256; b2 < 2^21 for 8/10/12-bit samples), in exact
// 32-bit arithmetic:
//  - sumsq <= 25 * 4095^2 < 2^31; a < 2^21 and d < 2^17, so a * N and d * d
//    are 24-bit products below 2^31;
//  - z >= 255 (a2 = 256) exactly when p >= thr = ceil((255 * 2^20 - 2^19) /
//    scale), so p is clamped to thr first: q * scale < 2^28 + scale;
//  - (256 - a2) * recip <= 255 * 455 < 2^24 and sum <= 25 * 4095 < 2^24, and
//    their product plus 2^11 stays below 2^32 (largest case 255 * 164 *
//    102375 + 2^11 = 4281324548).
template <int N, int BD>
__device__ __forceinline__ uint32_t ab_of(int sum, int sumsq, uint32_t scale, uint32_t thr,
                                          const uint16_t *xbx1)
{
    constexpr uint32_t recip = ((1u << 12) + N / 2) / N;
    const uint32_t a = (uint32_t)round2i(sumsq, 2 * (BD - 8));
    const uint32_t d = (uint32_t)round2i(sum, BD - 8);
    const int p = max((int)__umul24(a, (uint32_t)N) - (int)__umul24(d, d), 0);
    const uint32_t z = (min((uint32_t)p, thr) * scale + (1u << 19)) >> 20;
    const uint32_t a2 = z >= 255 ? 256u : xbx1[z];
    const uint32_t b2 = (__umul24((256 - a2) * recip, (uint32_t)sum) + (1u << 11)) >> 12;
    return (b2 << 9) | a2;
}

// Column sums over 2R+1 staged rows into VS/VQ, then per centre (y0-1+cr,
// x0-1+cc), cr < th+2, cc < tw+2, the packed A/B, written over VS/VQ as
// grid AB (row stride kG). Pass 0 (R = 2) needs odd rows y only
// (restoration.rs:347-356), so it walks only those centre rows.
//
// Each thread takes runs of 4 columns of one row, so that one 8-byte read
// brings 4 samples: the column sums read 2R+1 such groups of S; the box
// sums of 4 centres read VS columns 4j .. 4j+7 (two 8-byte reads) and VQ
// columns 4j .. 4j+7 (two 16-byte reads) and slide over them. Columns past the tile's
// (up to the row stride) are computed from stale LDS and never stored to
// the grid.
template <int R, int BD, int ROWS>
__device__ __forceinline__ void sgr_grid(const uint16_t *S, uint32_t *X, int th, int tw, int y0,
                                         uint32_t scale, uint32_t thr, const uint16_t *xbx1)
{
    constexpr int N = (2 * R + 1) * (2 * R + 1);
    constexpr int STEP = R == 2 ? 2 : 1;  // centre rows walked
    constexpr int T = LrShape<ROWS>::kThreads;
    constexpr int MAXR = R == 2 ? ROWS / 2 + 2 : ROWS + 2;  // centre rows walked, at most
    constexpr int Q1 = kS / 4;                               // 4-column groups of a VS row
    constexpr int Q2 = 17;                                   // 4-centre runs of a grid row (66 centres)
    constexpr int KC = (MAXR * Q2 + T - 1) / T;
    static_assert(kS % 8 == 0 && 4 * Q2 + 4 <= kS, "runs read VS / VQ columns 4j .. 4j+7 of one row");
    static_assert(kG % 4 == 0 && 4 * Q2 <= kG && (ROWS + 2) * kG * 4 <= LrShape<ROWS>::kXBytes,
                  "grid rows hold whole 16-byte runs inside the grid's LDS");
    const int crows = th + 2, ccols = tw + 2, scols = tw + 6;
    const int cr0 = R == 2 ? (y0 & 1) : 0;  // first centre row with y0-1+cr odd
    const int nr = (crows - cr0 + STEP - 1) / STEP;
    const int ng = (scols + 3) >> 2, nq = (ccols + 3) >> 2;
    const int tid = threadIdx.x;
    uint16_t *VS = reinterpret_cast<uint16_t *>(X);
    int *VQ = reinterpret_cast<int *>(X) + (ROWS + 2) * kS / 2;
#pragma unroll 2
    for (int e = tid; e < nr * Q1; e += T) {
        const int k = e / Q1, g = e - k * Q1, cr = cr0 + STEP * k;
        if (g >= ng)
            continue;
        int s[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
#pragma unroll
        for (int dy = -R; dy <= R; dy++) {
            const uint2 w = *reinterpret_cast<const uint2 *>(&S[(cr + 2 + dy) * kS + 4 * g]);
            const int v[4] = {(int)(w.x & 0xffffu), (int)(w.x >> 16), (int)(w.y & 0xffffu), (int)(w.y >> 16)};
#pragma unroll
            for (int m = 0; m < 4; m++) {
                s[m] += v[m];
                q[m] += v[m] * v[m];
            }
        }
        *reinterpret_cast<uint2 *>(&VS[cr * kS + 4 * g]) =
            make_uint2((uint32_t)s[0] | ((uint32_t)s[1] << 16), (uint32_t)s[2] | ((uint32_t)s[3] << 16));
        *reinterpret_cast<int4 *>(&VQ[cr * kS + 4 * g]) = make_int4(q[0], q[1], q[2], q[3]);
    }
    __syncthreads();
    uint32_t ab[KC][4];
#pragma unroll
    for (int i = 0; i < KC; i++) {
        const int e = tid + i * T;
        const int k = e / Q2, j = e - k * Q2, cr = cr0 + STEP * k;
#pragma unroll
        for (int m = 0; m < 4; m++)
            ab[i][m] = 0;
        if (k >= nr || j >= nq)
            continue;
        // VS rows are 8-byte aligned at column 4j, VQ rows 16-byte aligned
        const uint2 va = *reinterpret_cast<const uint2 *>(&VS[cr * kS + 4 * j]);
        const uint2 vb = *reinterpret_cast<const uint2 *>(&VS[cr * kS + 4 * j + 4]);
        const int4 q0 = *reinterpret_cast<const int4 *>(&VQ[cr * kS + 4 * j]);
        const int4 q1 = *reinterpret_cast<const int4 *>(&VQ[cr * kS + 4 * j + 4]);
        const int sv[8] = {(int)(va.x & 0xffffu), (int)(va.x >> 16), (int)(va.y & 0xffffu), (int)(va.y >> 16),
                           (int)(vb.x & 0xffffu), (int)(vb.x >> 16), (int)(vb.y & 0xffffu), (int)(vb.y >> 16)};
        const int qv[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
        for (int m = 0; m < 4; m++) {
            // centre column 4j+m is staged column 4j+m+2
            int s = 0, q = 0;
#pragma unroll
            for (int dx = -R; dx <= R; dx++) {
                s += sv[m + 2 + dx];
                q += qv[m + 2 + dx];
            }
            ab[i][m] = ab_of<N, BD>(s, q, scale, thr, xbx1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KC; i++) {
        const int e = tid + i * T;
        const int k = e / Q2, j = e - k * Q2, cr = cr0 + STEP * k;
        if (k >= nr || j >= nq)
            continue;
        uint32_t *dst = &X[cr * kG + 4 * j];  // 16-byte aligned
        if (4 * j + 4 <= ccols) {
            *reinterpret_cast<uint4 *>(dst) = make_uint4(ab[i][0], ab[i][1], ab[i][2], ab[i][3]);
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++)
                if (4 * j + m < ccols)
                    dst[m] = ab[i][m];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ int ab_a(uint32_t v) { return (int)(v & 511u); }
__device__ __forceinline__ int ab_b(uint32_t v) { return (int)(v >> 9); }

// guided_pass (restoration.rs:341-371), pass 0 (radius 2), for the 8 rows
// rs .. rs+7 of column c. Grid row j (tile row rs + j - 1) holds A/B only
// for odd plane rows: where (PAR + j) is even, PAR the parity of y0 + rs.
// An odd output row takes its own grid row with weights 5 6 5 and shift 8;
// an even one the odd rows above and below, shift 9.
template <int PAR>
__device__ __forceinline__ void sgr_pass0_strip(const uint32_t *X, int c, int rs, const int *uu, int *f0)
{
    int ta[10], tb[10];
#pragma unroll
    for (int j = 0; j < 10; j++) {
        if ((PAR + j) & 1)
            continue;
        const int k = (rs + j) * kG + c;
        const uint32_t p = X[k], q = X[k + 1], w = X[k + 2];
        ta[j] = 5 * (ab_a(p) + ab_a(w)) + 6 * ab_a(q);
        tb[j] = 5 * (ab_b(p) + ab_b(w)) + 6 * ab_b(q);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        if ((PAR + i) & 1)
            f0[i] = (ta[i + 1] * uu[i] + tb[i + 1] + (1 << 7)) >> 8;
        else
            f0[i] = ((ta[i] + ta[i + 2]) * uu[i] + tb[i] + tb[i + 2] + (1 << 8)) >> 9;
    }
}

// Pass 1 (radius 1) for the same strip: weights 4 on the cross, 3 on the
// diagonals, shift 9. Per grid row, T = 3 * (left + right) + 4 * centre
// (the row above or below) and U = 4 * (left + centre + right) (the own
// row); output row i sums T(i) + U(i + 1) + T(i + 2).
__device__ __forceinline__ void sgr_pass1_strip(const uint32_t *X, int c, int rs, const int *uu, int *f1)
{
    int ta[3], tb[3], ua[3], ub[3];  // rows j - 2, j - 1, j, rotating
#pragma unroll
    for (int j = 0; j < 10; j++) {
        const int k = (rs + j) * kG + c;
        const uint32_t p = X[k], q = X[k + 1], w = X[k + 2];
        const int sa = ab_a(p) + ab_a(w), sb = ab_b(p) + ab_b(w);
        const int qa = ab_a(q), qb = ab_b(q);
        ta[j % 3] = 3 * sa + 4 * qa;
        tb[j % 3] = 3 * sb + 4 * qb;
        ua[j % 3] = 4 * (sa + qa);
        ub[j % 3] = 4 * (sb + qb);
        if (j >= 2) {
            const int i = j - 2;
            const int a = ta[i % 3] + ua[(i + 1) % 3] + ta[j % 3];
            const int b = tb[i % 3] + ub[(i + 1) % 3] + tb[j % 3];
            f1[i] = (a * uu[i] + b + (1 << 8)) >> 9;
        }
    }
}

template <int ROWS>
__device__ __forceinline__ void lr_sgr(const Tile &t, uint8_t *lds)
{
    using Sh = LrShape<ROWS>;
    constexpr int T = Sh::kThreads;
    uint16_t *S = reinterpret_cast<uint16_t *>(lds);
    uint32_t *X = reinterpret_cast<uint32_t *>(lds + Sh::kSrcBytes);
    uint16_t *xbx1 = reinterpret_cast<uint16_t *>(lds + Sh::kSrcBytes + Sh::kXBytes);
    const LrPlane &P = *t.P;
    const int bd = t.bit_depth(), maxv = (1 << bd) - 1;
    const int tw = t.cs.w, th = t.rs.h, x0 = t.cs.x0, y0 = t.rs.y0;
    const TileOut o = tile_out(t, x0, y0);
    const int tid = threadIdx.x;
    uint16_t *su = reinterpret_cast<uint16_t *>(lds + Sh::kUnitLds);
    UnitStage us;
    us.load(t);
    const bool bad = stage_source<ROWS, T>(S, t, maxv, lds + Sh::kSrcBytes);
    us.store(su);
    if (__syncthreads_or(bad) && tid == 0)
        raise_status(P.status, kFlagSampleRange);
    const LrUnit u = *reinterpret_cast<const LrUnit *>(su);
    // x / (x + 1) table (restoration.rs:330-336), past the grids; written
    // after the staging (whose superres window may cover it), read after the
    // grid's first barrier.
    for (int z = tid; z < 256; z += T)
        xbx1[z] = z == 0 ? 1 : (uint16_t)(((z << 8) + z / 2) / (z + 1));

    // self_guided, restoration.rs:373-384, with guided_pass 341-371. The set
    // is the tile's (one unit), so its parameters come through scalar loads.
    const int set = __builtin_amdgcn_readfirstlane(u.set & 15);
    const SgrPass g0 = kSgrPasses.p[set][0], g1 = kSgrPasses.p[set][1];
    const int r0 = (int)g0.r, r1 = (int)g1.r;
    // Thread = column c and a strip of 8 output rows rs .. rs + 7 (its wave),
    // so each A/B grid row the strip needs is read once for all 8 rows.
    const int c = tid & 63, rs = (tid >> 6) * 8;
    int uu[8], f0[8], f1[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uu[i] = S[(rs + i + 3) * kS + c + 3];
        f0[i] = f1[i] = uu[i] << 4;
    }
    if (r0) {
        if (bd == 8)
            sgr_grid<2, 8, ROWS>(S, X, th, tw, y0, g0.scale, g0.thr, xbx1);
        else if (bd == 10)
            sgr_grid<2, 10, ROWS>(S, X, th, tw, y0, g0.scale, g0.thr, xbx1);
        else
            sgr_grid<2, 12, ROWS>(S, X, th, tw, y0, g0.scale, g0.thr, xbx1);
        if ((y0 + rs) & 1)
            sgr_pass0_strip<1>(X, c, rs, uu, f0);
        else
            sgr_pass0_strip<0>(X, c, rs, uu, f0);
    }
    if (r1) {
        if (r0)
            __syncthreads();  // the pass-0 grid is fully consumed
        if (bd == 8)
            sgr_grid<1, 8, ROWS>(S, X, th, tw, y0, g1.scale, g1.thr, xbx1);
        else if (bd == 10)
            sgr_grid<1, 10, ROWS>(S, X, th, tw, y0, g1.scale, g1.thr, xbx1);
        else
            sgr_grid<1, 12, ROWS>(S, X, th, tw, y0, g1.scale, g1.thr, xbx1);
        sgr_pass1_strip(X, c, rs, uu, f1);
    }
    const int w0 = u.xqd[0], w1 = u.xqd[1];
    if (c < tw) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = rs + i;
            if (r >= th)
                break;
            int v = w0 * f0[i] + w1 * (uu[i] << 4) + (128 - w0 - w1) * f1[i];
            v = (v + (1 << 10)) >> 11;
            put(o, r, c, min(max(v, 0), maxv));
        }
    }
}


// RestorationUnit::None on a plane with superres folded in: the upscaled
// CDEF samples of the tile (inside the stripe, so the staged rows are the
// CDEF image's), staged as the filters stage them.
template <int ROWS>
__device__ __forceinline__ void lr_copy_sr(const Tile &t, uint8_t *lds)
{
    using Sh = LrShape<ROWS>;
    constexpr int T = Sh::kThreads;
    uint16_t *S = reinterpret_cast<uint16_t *>(lds);
    const int maxv = (1 << t.bit_depth()) - 1;
    const int tw = t.cs.w, th = t.rs.h;
    const TileOut o = tile_out(t, t.cs.x0, t.rs.y0);
    const bool bad = stage_source_sr<ROWS, T>(S, t, maxv, lds + Sh::kSrcBytes);
    if (__syncthreads_or(bad) && threadIdx.x == 0)
        raise_status(t.P->status, kFlagSampleRange);
    for (int e = threadIdx.x; e < th * 64; e += T) {
        const int r = e >> 6, c = e & 63;
        if (c < tw)
            put(o, r, c, S[(r + 3) * kS + c + 3]);
    }
}

// One tile of any kind, in a workgroup of LrShape<ROWS>::kThreads threads
// over `lds` (LrShape<ROWS>::kLds bytes).
template <int ROWS>
__device__ __forceinline__ void lr_tile(const Tile &t, uint8_t *lds)
{
    if (t.kind == kUnitNone && t.P->sr_step)
        lr_copy_sr<ROWS>(t, lds);
    else if (t.kind == kUnitSelfGuided)
        lr_sgr<ROWS>(t, lds);
    else if (t.kind == kUnitWiener)
        lr_wiener<ROWS>(t, lds);
    else
        lr_copy<LrShape<ROWS>::kThreads>(t);
}

}  // namespace av1pf
