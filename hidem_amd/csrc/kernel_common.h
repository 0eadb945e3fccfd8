// Device helpers shared by the gfx950 kernels: wave64 reductions, the
// XCD-aware block numbering and the FP32 form's particle record.
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

// ---- CT side store -------------------------------------------------------------
// CT0..5 of the link side in tiled slot s (wave tile t, side k, lane:
// s = (t * maxLinks + k) * 64 + lane) sit in one block of 6 x 64 values per
// (t, k), interleaved so that a 64-lane access moves 16 B or 8 B per lane
// from one contiguous range: FP32 [c0..c3 of the 64 lanes as float4][c4, c5
// as float2], FP64 [c0, c1 as double2][c2, c3][c4, c5].  A side's CT is then
// 2 (FP32) or 3 (FP64) vector loads and as many stores instead of 6 of each;
// the fused step issues fewer memory instructions per side, which is what
// its address units spend their time on.
template <class CT>
__host__ __device__ inline int64_t ct_offset(int64_t s, int c) {
    const int64_t blk = (s >> 6) * 384, lane = s & 63;
    if constexpr (sizeof(CT) == 4) return c < 4 ? blk + lane * 4 + c : blk + 256 + lane * 2 + (c - 4);
    else return blk + (c >> 1) * 128 + lane * 2 + (c & 1);
}

typedef float ct_f4 __attribute__((ext_vector_type(4)));
typedef float ct_f2 __attribute__((ext_vector_type(2)));
typedef double ct_d2 __attribute__((ext_vector_type(2)));

// Non-temporal: a side's CT is read and written once per launch.
template <class CT>
__device__ __forceinline__ void ct_load(const CT *ct, int64_t s, CT v[6]) {
    const int64_t blk = (s >> 6) * 384, lane = s & 63;
    if constexpr (sizeof(CT) == 4) {
        const ct_f4 a = __builtin_nontemporal_load(reinterpret_cast<const ct_f4 *>(ct + blk + lane * 4));
        const ct_f2 b = __builtin_nontemporal_load(reinterpret_cast<const ct_f2 *>(ct + blk + 256 + lane * 2));
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
    } else {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const ct_d2 a = __builtin_nontemporal_load(reinterpret_cast<const ct_d2 *>(ct + blk + j * 128 + lane * 2));
            v[2 * j] = a.x;
            v[2 * j + 1] = a.y;
        }
    }
}

template <class CT>
__device__ __forceinline__ void ct_store(CT *ct, int64_t s, const CT v[6]) {
    const int64_t blk = (s >> 6) * 384, lane = s & 63;
    if constexpr (sizeof(CT) == 4) {
        const ct_f4 a = {v[0], v[1], v[2], v[3]};
        const ct_f2 b = {v[4], v[5]};
        __builtin_nontemporal_store(a, reinterpret_cast<ct_f4 *>(ct + blk + lane * 4));
        __builtin_nontemporal_store(b, reinterpret_cast<ct_f2 *>(ct + blk + 256 + lane * 2));
    } else {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const ct_d2 a = {v[2 * j], v[2 * j + 1]};
            __builtin_nontemporal_store(a, reinterpret_cast<ct_d2 *>(ct + blk + j * 128 + lane * 2));
        }
    }
}

// ---- FP32-form particle state ------------------------------------------------
// Gather record of a particle in the FP32 form (64 B: a link partner costs
// four 16 B loads): coordinates in FP64, the radius (NaN when not tame), the
// rotation bound, Diff rounded to FP32 and the low parts of Diff0 and Diff1.
// The contact test's fields (x, y, z, radius) fill the first 32 B, so an
// inert neighbour costs two 16 B loads.
struct __attribute__((aligned(16))) MixedRec {
    double x, y, z;
    float radius, rot;
    float d[6];
    float lo[2];
};
// The rest of the particle's own FP64 Diff (32 B, read only by the particle):
// Diff3..5 and the low part of Diff2.
struct __attribute__((aligned(16))) MixedTail {
    double dr[3];
    float lo2;
    float pad;
};

// The translation Diff of a particle in the FP32 form is stored as the
// record's FP32 part d[c] plus an FP32 low part lo = float(D - d[c]), rebuilt
// as double(d[c]) + double(lo).  It is the difference of two nearby FP64
// coordinates, so it seldom holds more than 48 significant bits: split_diff3
// checks the rebuilt values bit for bit and, when one does not come back (or
// is not finite, or is -0), sets lo[2] to NaN; the three FP64 components then
// go to an escape array that only such particles touch.  (The rotation Diff,
// the difference of small angles, often needs all 53 bits and stays FP64.)
// Returns true when exact.
__device__ __forceinline__ bool split_diff3(const double D[3], float hi[3], float lo[3]) {
    bool exact = true;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        hi[c] = (float)D[c];
        lo[c] = (float)(D[c] - (double)hi[c]);
        exact &= __double_as_longlong((double)hi[c] + (double)lo[c]) == __double_as_longlong(D[c]);
    }
    if (!exact) lo[2] = __builtin_nanf("");
    return exact;
}

// The FP64 Diff0..5 of particle p from its record, its tail and the escape
// array (3 doubles per particle).
__device__ __forceinline__ void join_diff6(const MixedRec &P, const MixedTail &T, const double *esc, int64_t p,
                                           double D[6]) {
    if (T.lo2 != T.lo2) {
#pragma unroll
        for (int c = 0; c < 3; c++) D[c] = esc[p * 3 + c];
    } else {
        D[0] = (double)P.d[0] + (double)P.lo[0];
        D[1] = (double)P.d[1] + (double)P.lo[1];
        D[2] = (double)P.d[2] + (double)T.lo2;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) D[3 + c] = T.dr[c];
}

}  // namespace hdm
