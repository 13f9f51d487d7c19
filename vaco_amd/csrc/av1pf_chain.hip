// av1pf_chain.hip - CDEF then loop restoration in one launch (per-frame
// path of a frame without superres).
//
// Reference semantics: the two stages of decode.rs:554-559 in order
// (apply_cdef, then apply_restoration reading the CDEF image inside each
// stripe and the pre-CDEF image across stripe borders, restoration.rs:
// 209-245); the stage bodies are cdef_sb (av1pf_cdef_body.h) and lr_tile
// (av1pf_lr_body.h), unchanged.
//
// Grid: workgroups 0 .. n_sb-1 run one CDEF superblock each, exactly as the
// CDEF kernel's single-frame form does; workgroups from lr_first (n_sb
// rounded up to a multiple of 8, so that workgroup b of either part runs on
// XCD b % 8) run one restoration tile each. A restoration tile reads the CDEF
// image of at most 2 x 4 superblocks (its stripe rows, its columns and the
// 3-sample filter margin, in luma superblock units), and waits for exactly
// those: each CDEF workgroup publishes its superblock with a release store of
// the launch's epoch, and the tile's first wave polls the flags and then
// acquires. The tile's descriptors come from the device mirror of the
// staging slot, which the CDEF workgroups fill (copying the staged bytes);
// the tile first waits for that copy the same way (a counter that the
// copying waves bump with release).
//
// Progress: workgroups are dispatched in index order, so when a restoration
// workgroup runs, every CDEF workgroup of the launch has been dispatched (and
// runs to its end without waiting on anything); the waits therefore end.
// They are bounded nonetheless: after cs.spin_limit polls a tile stops
// waiting, skips its work (reading neither descriptors that may not have
// arrived nor an unfinished CDEF image) and raises kFlagChainStall in the
// frame's status word, which the host reports as a device error.
//
// The restoration tiles then fill the machine while the last superblocks are
// filtered (one launch instead of two: no launch gap, no second ramp, and
// the CDEF launch's tail overlaps restoration work).
#include "av1pf_cdef_body.h"
#include "av1pf_lr_body.h"

namespace av1pf {

constexpr int kChainThreads = 512;  // 8-wave CDEF workgroups; restoration tiles of up to 64 rows
static_assert(LrShape<64>::kThreads == kChainThreads, "restoration tiles of 64 rows take 512 threads");

template <int SX, int SY>
constexpr size_t chain_lds()
{
    return sizeof(CdefLds<SX, SY>) > (size_t)LrShape<64>::kLds ? sizeof(CdefLds<SX, SY>)
                                                               : (size_t)LrShape<64>::kLds;
}

// Poll until every lane of the calling wave sees `ready()` (uniform after
// the vote), at most `limit` polls; then acquire at agent scope. Returns
// false if it gave up.
template <class Ready>
__device__ __forceinline__ bool wave_wait(Ready &&ready, uint32_t limit)
{
    bool ok = true;
    for (uint32_t n = 0;; n++) {
        if (__all(ready()))
            break;
        if (n >= limit) {
            ok = false;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok;
}

template <int SX, int SY, bool XCD>
__global__ __launch_bounds__(kChainThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_chain(
    const CdefFrame single, const LrPlane *planes, uint32_t n_items, const uint64_t *list, uint32_t n_sb,
    uint32_t lr_first,
    uint32_t n_tiles, uint32_t lr_xcd, const uint4 *__restrict__ up_src, uint4 *up_dst, uint32_t up_n16,
    const ChainSync cs)
{
    __shared__ __align__(16) uint8_t lds[chain_lds<SX, SY>()];
    const uint32_t b = blockIdx.x;
    if (b < n_sb) {
        cdef_sb<SX, SY, XCD, kChainThreads, true, true>(*reinterpret_cast<CdefLds<SX, SY> *>(lds), nullptr, 1,
                                                        single, up_src, up_dst, up_n16, b, n_sb, cs);
        return;
    }
    if (b < lr_first)
        return;  // padding up to a multiple of 8
    const int tid = threadIdx.x;
    const uint32_t j = b - lr_first;
    const uint32_t wg = lr_xcd ? xcd_remap(j, n_tiles) : j;
    // Whether the tile gave up waiting for the copy / for its superblocks
    // (one word each: a word is read by every thread between two barriers).
    __shared__ uint32_t no_copy, no_deps;

    // The descriptors (planes, list, row and column segments, units) are in
    // the device mirror once the CDEF workgroups' copy is complete. A tile
    // that gave up on the copy reads none of them (the mirror may still hold
    // another call's bytes): it only raises the flag.
    if (tid < 64) {
        const bool ok = wave_wait(
            [&]() {
                return __hip_atomic_load(cs.copied, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cs.copied_target;
            },
            cs.spin_limit);
        if (tid == 0)
            no_copy = !ok;
    }
    __syncthreads();
    if (no_copy) {
        if (tid == 0)
            raise_status(single.status, kFlagChainStall);
        return;
    }
    const Tile t = fetch_tile(planes, n_items, list, wg);

    // Superblocks whose CDEF output the tile reads: Source::get's rows
    // (restoration.rs:233-244) inside the stripe, within 3 rows of the tile;
    // columns within 4 of it (the staging's 8-byte groups start 4 columns
    // left of the tile); a copy tile reads its own rows and columns. Item
    // index = plane for a single frame (every plane has an LrPlane).
    if (tid < 64) {
        const LrPlane &P = *t.P;
        const bool chroma = t.P != planes;
        const int sx = chroma ? SX : 0, sy = chroma ? SY : 0;
        const int W = (int)P.width, H = (int)P.height;
        int r0 = t.rs.y0, r1 = t.rs.y0 + t.rs.h - 1, c0 = t.cs.x0, c1 = t.cs.x0 + t.cs.w - 1;
        if (t.kind != kUnitNone) {
            r0 = max(max(r0 - 3, t.rs.stripe_start), 0);
            r1 = min(min(r1 + 3, t.rs.stripe_end), H - 1);
            c0 = max(c0 - 4, 0);
            c1 = min(c1 + 4, W - 1);
        }
        const int sr0 = (r0 << sy) >> 6, sr1 = (r1 << sy) >> 6;
        const int sc0 = (c0 << sx) >> 6, sc1 = (c1 << sx) >> 6;
        const int sr = sr0 + (tid >> 2), sc = sc0 + (tid & 3);
        const bool need = tid < 8 && sr <= sr1 && sc <= sc1;
        const uint32_t *flag = cs.flags + (need ? (uint32_t)sr * single.sb_cols + (uint32_t)sc : 0u);
        const bool ok = wave_wait(
            [&]() { return !need || __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == cs.epoch; },
            cs.spin_limit);
        if (tid == 0)
            no_deps = !ok;
    }
    __syncthreads();
    if (no_deps) {  // the frame fails; its CDEF image is not complete
        if (tid == 0)
            raise_status(single.status, kFlagChainStall);
        return;
    }
    lr_tile<64>(t, lds);
}

template <int SX, int SY>
static void launch_chain_mode(const CdefFrame &f, bool xcd, const LrPlane *planes, uint32_t n_items,
                              const uint64_t *list,
                              uint32_t n_sb, uint32_t lr_first, uint32_t n_tiles, uint32_t lr_xcd,
                              const uint4 *us, uint4 *ud, uint32_t n16, const ChainSync &cs, hipStream_t s,
                              hipEvent_t ev0, hipEvent_t ev1)
{
    const dim3 g(lr_first + n_tiles), blk(kChainThreads);
    if (xcd)
        hipExtLaunchKernelGGL((k_chain<SX, SY, true>), g, blk, 0, s, ev0, ev1, 0, f, planes, n_items, list, n_sb,
                              lr_first, n_tiles, lr_xcd, us, ud, n16, cs);
    else
        hipExtLaunchKernelGGL((k_chain<SX, SY, false>), g, blk, 0, s, ev0, ev1, 0, f, planes, n_items, list,
                              n_sb, lr_first, n_tiles, lr_xcd, us, ud, n16, cs);
}

// One frame's CDEF and restoration in one launch. f: the frame's CDEF
// descriptor (side information read in place, as the single-frame CDEF
// launch does); planes / list: its n_items restoration items (one per
// plane, in plane order) and n_tiles tile entries (no list: tiles in plane
// order), in the device mirror that the launch fills from up_src (up_bytes). mode, xcd: as launch_cdef; lr_xcd: the
// list is in plane order, numbered XCD-contiguously. cs: flags for the
// frame's superblocks, the copy counter and its target, the epoch.
hipError_t launch_chain(const CdefFrame &f, int mode, bool xcd, const LrPlane *planes, uint32_t n_items,
                        const uint64_t *list, uint32_t n_tiles, bool lr_xcd, const void *up_src, void *up_dst,
                        size_t up_bytes,
                        const ChainSync &cs, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    const uint32_t n_sb = f.sb_cols * f.sb_rows;
    const uint32_t n16 = (uint32_t)((up_bytes + 15) / 16);
    if (n_sb == 0 || n_tiles == 0 || n16 == 0 || !cs.flags || !cs.copied || cs.epoch == 0)
        return hipErrorInvalidValue;
    const uint32_t lr_first = (n_sb + 7) & ~7u;
    const uint4 *us = static_cast<const uint4 *>(up_src);
    uint4 *ud = static_cast<uint4 *>(up_dst);
    switch (mode) {
    case 0:
        launch_chain_mode<1, 1>(f, xcd, planes, n_items, list, n_sb, lr_first, n_tiles, lr_xcd, us, ud, n16,
                                cs, s, ev0, ev1);
        break;
    case 1:
        launch_chain_mode<0, 0>(f, xcd, planes, n_items, list, n_sb, lr_first, n_tiles, lr_xcd, us, ud, n16,
                                cs, s, ev0, ev1);
        break;
    case 2:
        launch_chain_mode<1, 0>(f, xcd, planes, n_items, list, n_sb, lr_first, n_tiles, lr_xcd, us, ud, n16,
                                cs, s, ev0, ev1);
        break;
    default:
        launch_chain_mode<0, 1>(f, xcd, planes, n_items, list, n_sb, lr_first, n_tiles, lr_xcd, us, ud, n16,
                                cs, s, ev0, ev1);
        break;
    }
    return hipGetLastError();
}

// Copy-counter increments of one chained launch: the CDEF workgroups whose
// first wave copies part of the staged bytes (see cdef_sb).
uint32_t chain_copiers(uint32_t n_sb, size_t up_bytes)
{
    const uint32_t n16 = (uint32_t)((up_bytes + 15) / 16);
    const uint32_t w = (n16 + 63) / 64;
    return w < n_sb ? w : n_sb;
}

}  // namespace av1pf
