// vc_common.h -- device helpers shared by the gfx950 kernels of libvc_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/vc_abi.h"
#include "../../include/vc_math.h"

#define VC_VE 0.0f      // reversal potentials, mV (src/CONST.h:7-9)
#define VC_VI (-75.0f)
#define VC_VL (-65.0f)

// error reporting for the C ABI (one message per host thread)
void vc_set_error(const char *fmt, ...);

#define VC_HIP_CHECK(expr)                                                   \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            vc_set_error("%s failed: %s", #expr, hipGetErrorString(e_));     \
            return VC_E_HIP;                                                 \
        }                                                                    \
    } while (0)

// number of set bits of m in lanes below the calling lane (wave64)
__device__ __forceinline__ uint32_t vc_mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t vc_lane() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t vc_readlane_u32(uint32_t x, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)l);
}
__device__ __forceinline__ uint64_t vc_readlane_u64(uint64_t x, uint32_t l) {
    uint32_t lo = vc_readlane_u32((uint32_t)x, l);
    uint32_t hi = vc_readlane_u32((uint32_t)(x >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ float vc_readlane_f32(float x, uint32_t l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), (int)l));
}

// Reference block-reduction tree restricted to one 32-lane group: the
// shuffle-down halving of src/util/cuda_util.cuh:22-44 (warps_reduce).
// Lane 0 of each 32-lane half of the wavefront ends with the group's sum.
__device__ __forceinline__ float vc_tree32(float x) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) x += __shfl_down(x, off, 32);
    return x;
}
