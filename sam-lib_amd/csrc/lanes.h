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

// v of the lane that DPP control CTRL selects (quad_perm, row_half_mirror ...),
// per 32-bit half.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(u)), CTRL, 0xF, 0xF, false));
    const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(u >> 32)), CTRL, 0xF, 0xF, false));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// Sum of v over each aligned group of 8 lanes, in every lane of the group:
// lane pairs (xor 1), then quads (xor 2), then the two quads (each lane adds
// its mirror in the group, which holds the other quad's sum); a + b and
// b + a round alike, so all eight lanes get the same bits.
__device__ __forceinline__ double sum8(double v)
{
    v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    return v;
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
