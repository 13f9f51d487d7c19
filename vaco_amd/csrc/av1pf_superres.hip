// av1pf_superres.hip - super-resolution upscaling (AV1 7.16) for gfx950.
//
// Reference semantics (matteopolak/vaco, crates/codec/vaco-codec-av1/src/
// superres.rs): subpel_start 145-152, upscale_plane 163-216. One workgroup
// per kSrSeg output samples of R rows (sr_block); the eight taps are read
// from the padded source row (edge-clamped to the padded width).
#include "av1pf_superres_body.h"

namespace av1pf {

template <int R>
__global__ __launch_bounds__(kSrThreads) void k_superres(const SrJob *__restrict__ jobs, int n_jobs)
{
    __shared__ __align__(16) uint8_t lds[sr_lds_bytes<R>()];
    const SrJob &j = find_item(jobs, n_jobs, blockIdx.x);
    sr_block<R>(j, blockIdx.x - j.block_begin, lds);
}

// rows: rows per workgroup (4, 8 or 16; the jobs' block_begin counts use
// the same value). ev0/ev1 (may be null) receive the dispatch's start and end.
hipError_t launch_superres(const SrJob *jobs, int n, uint32_t blocks, int rows, hipStream_t s,
                           hipEvent_t ev0, hipEvent_t ev1)
{
    if (blocks == 0)
        return hipSuccess;
    const dim3 g(blocks), b(kSrThreads);
    if (rows == 4)
        hipExtLaunchKernelGGL(k_superres<4>, g, b, 0, s, ev0, ev1, 0, jobs, n);
    else if (rows == 16)
        hipExtLaunchKernelGGL(k_superres<16>, g, b, 0, s, ev0, ev1, 0, jobs, n);
    else
        hipExtLaunchKernelGGL(k_superres<8>, g, b, 0, s, ev0, ev1, 0, jobs, n);
    return hipGetLastError();
}

}  // namespace av1pf
