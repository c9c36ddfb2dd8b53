// Wavefront helpers shared by the device code (included by .hip files only).
#pragma once

#include <hip/hip_runtime.h>

namespace sprgpu {

// Value of lane l's v in every lane (l uniform across the wavefront).
__device__ __forceinline__ float lane_bcast(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_bcast(double v, int l)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(u)), l));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(u >> 32)), l));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// v, hidden from the optimiser (no code): keeps a per-lane value from being
// treated as loop-invariant, so comparisons with it stay where they are used.
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Index of the descriptor whose launch-local prefix range holds b.
__device__ __forceinline__ int find_desc(const int64_t *prefix, int n, int64_t b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

}  // namespace sprgpu
