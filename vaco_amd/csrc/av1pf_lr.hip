// av1pf_lr.hip - loop restoration (AV1 7.17) for gfx950.
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
#include "av1pf_lr_body.h"

namespace av1pf {

// One launch over every tile of the plan, one workgroup per tile: list
// entry b, or without a list tile b of the n_items plane items in plane
// order (fetch_tile). XCD: workgroup b takes tile xcd_remap(b) (tiles in
// plane order). ROWS: no Wiener or self-guided tile is taller (copy tiles,
// which use no LDS, may be up to 64 rows).
template <bool XCD, int ROWS>
__global__ __launch_bounds__(LrShape<ROWS>::kThreads) void k_lr(const LrPlane *__restrict__ planes,
                                                               uint32_t n_items,
                                                               const uint64_t *__restrict__ list)
{
    __shared__ __align__(16) uint8_t lds[LrShape<ROWS>::kLds];
    const uint32_t wg = XCD ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    lr_tile<ROWS>(fetch_tile(planes, n_items, list, wg), lds);
}

template <int ROWS>
static void launch_lr_rows(const LrPlane *planes, uint32_t n_items, const uint64_t *list, uint32_t n, bool xcd,
                           hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    const dim3 g(n), b(LrShape<ROWS>::kThreads);
    if (xcd)
        hipExtLaunchKernelGGL((k_lr<true, ROWS>), g, b, 0, s, ev0, ev1, 0, planes, n_items, list);
    else
        hipExtLaunchKernelGGL((k_lr<false, ROWS>), g, b, 0, s, ev0, ev1, 0, planes, n_items, list);
}

// n tiles of n_items plane items; list: the tiles in launch order, or null
// for plane order; xcd: tiles in plane order, numbered XCD-contiguously.
// rows: the tallest Wiener or self-guided tile (64, 32 or 16; the workgroup
// takes 8 threads per row). ev0/ev1 (may be null) receive the dispatch's
// start and end.
hipError_t launch_lr(const LrPlane *planes, uint32_t n_items, const uint64_t *list, uint32_t n, bool xcd,
                     int rows, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (n == 0)
        return hipSuccess;
    if (rows <= 16)
        launch_lr_rows<16>(planes, n_items, list, n, xcd, s, ev0, ev1);
    else if (rows <= 32)
        launch_lr_rows<32>(planes, n_items, list, n, xcd, s, ev0, ev1);
    else
        launch_lr_rows<64>(planes, n_items, list, n, xcd, s, ev0, ev1);
    return hipGetLastError();
}

}  // namespace av1pf
