// av1pf_common.h - device helpers shared by the stage's kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "av1pf_device.h"

namespace av1pf {

// Two 16-bit lanes of one 32-bit register (packed v_pk_* / v_dot2 operands).
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int ilog2u(uint32_t v) { return 31 - __clz(v); }

// Pointers into device memory. A pointer read out of a work-item struct is
// generic to the compiler, which then emits flat loads and stores; a flat
// access counts against the LDS wait counter as well, so an LDS wait issued
// after a flat store also waits for the store. Accesses through gptr are
// global_* instructions. (The host pass of a .hip file only parses device
// functions, and copies of address-space-qualified structs do not resolve
// there, so the alias is a plain pointer in that pass.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T *;
#else
template <class T>
using gptr = T *;
#endif

template <class T>
__device__ __forceinline__ gptr<T> glb(T *p)
{
    return (gptr<T>)p;
}

// Round2 of the AV1 specification (restoration.rs:251-257) for int32.
__device__ __forceinline__ int round2i(int v, int s)
{
    return s == 0 ? v : (v + (1 << (s - 1))) >> s;
}

// Raise `bits` in a frame status word. Status words may live in coherent
// pinned host memory, so the update is a system-scope vector atomic.
__device__ __forceinline__ void raise_status(uint32_t *status, uint32_t bits)
{
    atomicOr_system(status, bits);
}

// Wait until this wave's vector memory operations, stores included, have
// completed (s_waitcnt vmcnt(0); the expcnt and lgkmcnt fields at their
// maximum are not waited on).
__device__ __forceinline__ void wait_vm_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8) and each XCD has its own L2. The remap gives each XCD one
// contiguous range of the n work items, so neighbouring items, which share
// halo lines, are fetched into the same L2. It is a bijection of [0, n).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n)
{
    constexpr uint32_t kXcds = 8;
    const uint32_t per = n / kXcds, rem = n % kXcds, x = b % kXcds, l = b / kXcds;
    return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
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

}  // namespace av1pf
