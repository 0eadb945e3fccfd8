// Device helpers shared by the gfx950 kernels: wave64 reductions and the
// XCD-aware block numbering.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hdm {

// MI355X dispatches consecutive workgroups of a launch round-robin over its
// 8 XCDs, each with a private L2.  Renumbering so that XCD k executes one
// contiguous range of logical blocks keeps particles that are neighbours in
// index (and therefore in space) on the same L2.  Bijective for any grid size.
__device__ inline int64_t xcd_block_index(int64_t nBlocks) {
    const int64_t NX = 8;
    int64_t i = blockIdx.x;
    int64_t q = nBlocks / NX, r = nBlocks % NX;
    int64_t xcd = i % NX, local = i / NX;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

__device__ inline double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

__device__ inline int32_t wave_inclusive_sum_i32(int32_t v) {
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int32_t w = __shfl_up(v, o, 64);
        if (lane >= o) v += w;
    }
    return v;
}

}  // namespace hdm
