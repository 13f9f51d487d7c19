// av1pf_stage.hip - the side-information upload of one call.
//
// Each call's side information (skip map, CDEF indices, restoration units,
// work lists, kernel descriptors) is written by the host into a pinned,
// coherent staging buffer. Instead of a DMA transfer, whose hand-over to the
// compute queue costs more than the copy itself for these tens of kilobytes,
// this kernel reads the buffer across the host link and writes its device
// mirror, in stream order with the kernels that read it.
#include <algorithm>

#include "av1pf_common.h"

namespace av1pf {

// At most 64 workgroups: the copy overlaps the previous call's kernels, and
// waves waiting on the host link would otherwise hold wave slots on every
// CU. Eight loads in flight per thread keep the link busy with few waves.
constexpr int kUpThreads = 256;
constexpr uint32_t kUpMaxBlocks = 64;
constexpr uint32_t kUpUnroll = 8;

// n16: 16-byte words to copy (both buffers hold at least n16 * 16 bytes).
__global__ __launch_bounds__(kUpThreads) void k_upload(uint4 *__restrict__ dst,
                                                       const uint4 *__restrict__ src, uint32_t n16)
{
    const uint32_t stride = gridDim.x * kUpThreads;
    uint32_t i = blockIdx.x * kUpThreads + threadIdx.x;
    for (; i + (kUpUnroll - 1) * stride < n16; i += kUpUnroll * stride) {
        uint4 v[kUpUnroll];
#pragma unroll
        for (uint32_t k = 0; k < kUpUnroll; k++)
            v[k] = src[i + k * stride];
#pragma unroll
        for (uint32_t k = 0; k < kUpUnroll; k++)
            dst[i + k * stride] = v[k];
    }
    for (; i < n16; i += stride)
        dst[i] = src[i];
}

// ev0/ev1 (may be null) receive the dispatch's start and end.
hipError_t launch_upload(void *dst, const void *src, size_t bytes, hipStream_t s, hipEvent_t ev0,
                         hipEvent_t ev1)
{
    const uint32_t n16 = (uint32_t)((bytes + 15) / 16);
    if (n16 == 0)
        return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>(kUpMaxBlocks, (n16 + kUpThreads - 1) / kUpThreads);
    hipExtLaunchKernelGGL(k_upload, dim3(blocks), dim3(kUpThreads), 0, s, ev0, ev1, 0,
                          (uint4 *)dst, (const uint4 *)src, n16);
    return hipGetLastError();
}

// w x h samples of a u16 plane to one byte each (8-bit frame output of a
// frame in which no stage runs). One thread per sample.
__global__ __launch_bounds__(256) void k_to_bytes(const uint16_t *__restrict__ src,
                                                  uint32_t src_stride, uint8_t *__restrict__ dst,
                                                  uint32_t dst_stride, uint32_t w, uint32_t h)
{
    const uint32_t x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x < w && y < h)
        dst[(size_t)y * dst_stride + x] = (uint8_t)src[(size_t)y * src_stride + x];
}

hipError_t launch_to_bytes(const uint16_t *src, uint32_t src_stride, uint8_t *dst,
                           uint32_t dst_stride, uint32_t w, uint32_t h, hipStream_t s)
{
    if (w == 0 || h == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_to_bytes, dim3((w + 255) / 256, h), dim3(256), 0, s, src, src_stride, dst,
                       dst_stride, w, h);
    return hipGetLastError();
}

}  // namespace av1pf
