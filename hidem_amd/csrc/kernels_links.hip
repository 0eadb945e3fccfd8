// Per-particle link lists and the side (CT) layout for gfx950.
//
//   link_count   LinkAccelSide::getMaxLinksPerParticle (LinkAccelSide.h:69-90)
//   link_fill    LinkAccelSide::generateLinks (:92-143): each link id placed
//                in a slot of both endpoints, then every particle's slots are
//                sorted ascending (the reference bubble-sorts, :123-137), and
//                generatePerParticleIndexes (:43-67) resolves the other
//                endpoint and which end this particle is.
//   ct_import /  LinkAccelSide::sortSideData + LinkDataHandler::
//   ct_export    reorderSideDataPadded (LinkDataHandler.h:176-191): the CT
//                side arrays are permuted so that particle p's k-th link side
//                lives in wave-tile slot k of p (the reference's AOS_CT order
//                makes each particle's sides contiguous; the wave tile makes
//                slot k of 64 consecutive particles contiguous instead, which
//                is what a 64-lane wavefront reads in one access).  sidePos
//                plays the role of sideIndEndA/B (LinkData::reindexSideData,
//                LinkData.h:183-208).
#include <cstring>

#include "hdm_context.h"
#include "kernel_common.h"

namespace hdm {

namespace {

constexpr int TPB = 256;

__device__ inline int64_t slot_index(int64_t p, int32_t k, int32_t width) {
    return (p >> 6) * (int64_t)64 * width + (int64_t)k * 64 + (p & 63);
}

__global__ void __launch_bounds__(TPB) k_link_count(const int32_t *__restrict__ A, const int32_t *__restrict__ B,
                                                     int64_t L, int64_t n, uint32_t *__restrict__ count,
                                                     uint64_t *errFlag) {
    int64_t l = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (l >= L) return;
    int32_t a = A[l], b = B[l];
    if (a < 0 || a >= n || b < 0 || b >= n || a == b) {
        atomicOr((unsigned long long *)errFlag, 1ull);
        return;
    }
    atomicAdd(&count[a], 1u);
    atomicAdd(&count[b], 1u);
}

__global__ void __launch_bounds__(TPB) k_link_fill(const int32_t *__restrict__ A, const int32_t *__restrict__ B,
                                                    int64_t L, int32_t width, uint32_t *__restrict__ cursor,
                                                    int32_t *__restrict__ slotLink) {
    int64_t l = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (l >= L) return;
    int32_t a = A[l], b = B[l];
    uint32_t ka = atomicAdd(&cursor[a], 1u);
    uint32_t kb = atomicAdd(&cursor[b], 1u);
    slotLink[slot_index(a, (int32_t)ka, width)] = (int32_t)l;
    slotLink[slot_index(b, (int32_t)kb, width)] = (int32_t)l;
}

__global__ void __launch_bounds__(TPB) k_link_sort(const int32_t *__restrict__ A, const int32_t *__restrict__ B,
                                                    const uint8_t *__restrict__ state,
                                                    int64_t n, int32_t width, const uint32_t *__restrict__ count,
                                                    int32_t *__restrict__ slotLink, int32_t *__restrict__ slotOther,
                                                    int32_t *__restrict__ sidePos) {
    int64_t p = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    if (p >= n) return;
    int32_t cnt = (int32_t)count[p];
    int64_t base = slot_index(p, 0, width);
    for (int32_t i = 1; i < cnt; i++) {
        int32_t v = slotLink[base + (int64_t)i * 64];
        int32_t j = i - 1;
        while (j >= 0 && slotLink[base + (int64_t)j * 64] > v) {
            slotLink[base + (int64_t)(j + 1) * 64] = slotLink[base + (int64_t)j * 64];
            j--;
        }
        slotLink[base + (int64_t)(j + 1) * 64] = v;
    }
    for (int32_t k = 0; k < cnt; k++) {
        int64_t s = base + (int64_t)k * 64;
        int32_t l = slotLink[s];
        int32_t other = B[l];
        int32_t end = 0;
        if (other == (int32_t)p) {
            other = A[l];
            end = 1;
        }
        sidePos[2 * (int64_t)l + end] = (int32_t)s;
        // both slot words carry the link's broken flag in their sign bit, so
        // the force gather needs no per-link state lookup (the fused step
        // reads only the partner word)
        slotOther[s] = state[l] == 0 ? (other | SLOT_BROKEN) : other;
        if (state[l] == 0) slotLink[s] = l | SLOT_BROKEN;
    }
}

template <class CT>
__global__ void __launch_bounds__(TPB) k_ct_import(const int32_t *__restrict__ sidePos, int64_t nSides,
                                                    const double *__restrict__ s0, const double *__restrict__ s1,
                                                    const double *__restrict__ s2, const double *__restrict__ s3,
                                                    const double *__restrict__ s4, const double *__restrict__ s5,
                                                    CT *__restrict__ ct) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nSides) return;
    int64_t d = sidePos[i];
    ct[ct_offset<CT>(d, 0)] = (CT)s0[i];
    ct[ct_offset<CT>(d, 1)] = (CT)s1[i];
    ct[ct_offset<CT>(d, 2)] = (CT)s2[i];
    ct[ct_offset<CT>(d, 3)] = (CT)s3[i];
    ct[ct_offset<CT>(d, 4)] = (CT)s4[i];
    ct[ct_offset<CT>(d, 5)] = (CT)s5[i];
}

template <class CT>
__global__ void __launch_bounds__(TPB) k_ct_export(const int32_t *__restrict__ sidePos, int64_t nSides,
                                                    const CT *__restrict__ ct, double *__restrict__ s0,
                                                    double *__restrict__ s1, double *__restrict__ s2,
                                                    double *__restrict__ s3, double *__restrict__ s4,
                                                    double *__restrict__ s5) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nSides) return;
    int64_t d = sidePos[i];
    s0[i] = (double)ct[ct_offset<CT>(d, 0)];
    s1[i] = (double)ct[ct_offset<CT>(d, 1)];
    s2[i] = (double)ct[ct_offset<CT>(d, 2)];
    s3[i] = (double)ct[ct_offset<CT>(d, 3)];
    s4[i] = (double)ct[ct_offset<CT>(d, 4)];
    s5[i] = (double)ct[ct_offset<CT>(d, 5)];
}

// Break prefilter thresholds, in two passes after the link build.
// Pass 1 (per link): the largest L0 (as float bits, rounded up; +inf when
// some L0 is not finite), which bounds the bend term, and the largest
// threshold T = L0 (1 + maxLoad) with per-link maxLoad, else L0 (float bits,
// rounded up), which sets the scale of the codes.
__device__ __forceinline__ float float_up(double v) {
    const float f = (float)v;
    return (double)f < v ? __uint_as_float(__float_as_uint(f) + 1u) : f;  // next float up
}

__device__ __forceinline__ void block_max_bits(float v, uint64_t *out) {
    __shared__ uint32_t part[TPB / 64];
    // non-negative floats order as their bit patterns
    const uint32_t bits = wave_max_u32(__float_as_uint(v));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = bits;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = part[0];
        for (int w = 1; w < TPB / 64; w++) m = part[w] > m ? part[w] : m;
        if (m) atomicMax((unsigned long long *)out, (unsigned long long)m);
    }
}

__global__ void __launch_bounds__(TPB) k_threshold_max(int64_t L, const double *__restrict__ initialLen,
                                                        const double *__restrict__ maxLoad, uint64_t *maxBits) {
    float bigL0 = 0.0f, bigT = 0.0f;
    for (int64_t l = (int64_t)blockIdx.x * TPB + threadIdx.x; l < L; l += (int64_t)gridDim.x * TPB) {
        const double L0 = initialLen[l];
        if (L0 >= 1e-30 && L0 <= 1e300) {
            const double T = maxLoad ? L0 * (1.0 + maxLoad[l]) : L0;
            const float b = float_up(L0);
            bigL0 = b > bigL0 ? b : bigL0;
            if (T > 0.0 && T <= 1e300) {
                const float t = float_up(T);
                bigT = t > bigT ? t : bigT;
            }
        } else if (!(L0 < 1e-30)) {
            bigL0 = __builtin_inff();  // NaN or huge: no useful bound
        }
    }
    block_max_bits(bigL0, maxBits);
    block_max_bits(bigT, maxBits + 1);
}

// Pass 2 (per side, into its wave-tile slot): a 16-bit code with
// code * scale <= T, code = floor(T / scale) held to 65535, where scale is
// the largest T / 65535 (so codes resolve T to 1.5e-5 of the largest); 0
// (never passes) when L0 is below 1e-30, not finite, or no scale exists.
// The prefilter reads 2 B per side; the exact test still reads the FP64
// initialLen and maxLoad.
__global__ void __launch_bounds__(TPB) k_side_codes(const int32_t *__restrict__ sidePos, int64_t nSides,
                                                     const double *__restrict__ initialLen,
                                                     const double *__restrict__ maxLoad, double scale,
                                                     uint16_t *__restrict__ slotThr) {
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < nSides; i += (int64_t)gridDim.x * TPB) {
        const int64_t l = i >> 1;
        const double L0 = initialLen[l];
        const bool ok = L0 >= 1e-30 && L0 <= 1e300 && scale > 0.0;
        const double T = maxLoad ? L0 * (1.0 + maxLoad[l]) : L0;
        double q = ok ? floor(T / scale) : 0.0;
        if (!(q >= 0.0)) q = 0.0;  // also NaN
        if (q > 65535.0) q = 65535.0;
        slotThr[sidePos[i]] = (uint16_t)q;
    }
}

// Flags a particle whose link sides name one partner twice, or itself.
__global__ void __launch_bounds__(TPB) k_partner_check(const uint32_t *__restrict__ linkCount,
                                                        const int32_t *__restrict__ slotOther, int32_t width,
                                                        int64_t n, uint64_t *flag) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const uint32_t cnt = linkCount[p];
    const int64_t base = (p >> 6) * (int64_t)64 * width + (p & 63);
    bool dup = false;
    for (uint32_t k = 0; k < cnt && !dup; k++) {
        const int32_t q = slotOther[base + (int64_t)k * 64] & SLOT_LINK_MASK;
        dup = q == (int32_t)p;
        for (uint32_t j = 0; j < k && !dup; j++) dup = (slotOther[base + (int64_t)j * 64] & SLOT_LINK_MASK) == q;
    }
    if (dup) atomicOr((unsigned long long *)flag, 1ull);
}

__global__ void __launch_bounds__(TPB) k_count_broken(const uint8_t *__restrict__ state, int64_t L, uint64_t *out) {
    uint32_t c = 0;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < L; i += (int64_t)gridDim.x * TPB)
        c += state[i] == 0 ? 1u : 0u;
    c = wave_sum_u32(c);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd((unsigned long long *)out, (unsigned long long)c);
}

// computeActiveLinks and computeIteratedNeighbours (HiDEMLib.h:197-232): a
// particle's links with state != 0 (iterateParticleLinksCT over its slots),
// and the entries of the reference's neighbour list (the list as built plus
// the link partners it leaves out).  Sums over all particles (all are owned
// on one GPU) as finalizeTs reduces them (HiDEMLib.h:518-523).
__global__ void __launch_bounds__(TPB) k_particle_stats(const uint32_t *__restrict__ linkCount,
                                                         const int32_t *__restrict__ slotLink, int32_t width,
                                                         const uint8_t *__restrict__ state,
                                                         const int32_t *__restrict__ nbrCount,
                                                         const int32_t *__restrict__ nbrLinked, int64_t n,
                                                         int32_t *__restrict__ activeOut,
                                                         int32_t *__restrict__ nbrOut, uint64_t *sums) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    uint32_t act = 0, nb = 0;
    if (p < n) {
        const uint32_t cnt = linkCount[p];
        const int64_t base = (p >> 6) * (int64_t)64 * width + (p & 63);
        for (uint32_t k = 0; k < cnt; k++) act += state[slotLink[base + (int64_t)k * 64] & SLOT_LINK_MASK] ? 1u : 0u;
        nb = (uint32_t)nbrCount[p] + (nbrLinked ? (uint32_t)nbrLinked[p] : 0u);
        if (activeOut) activeOut[p] = (int32_t)act;
        if (nbrOut) nbrOut[p] = (int32_t)nb;
    }
    act = wave_sum_u32(act);
    nb = wave_sum_u32(nb);
    if ((threadIdx.x & 63) == 0) {
        if (act) atomicAdd((unsigned long long *)&sums[0], (unsigned long long)act);
        if (nb) atomicAdd((unsigned long long *)&sums[1], (unsigned long long)nb);
    }
}

inline int blocks_for(int64_t n) { return (int)((n + TPB - 1) / TPB); }

}  // namespace

hipError_t launch_particle_stats(Context &c, int32_t *activeLinks, int32_t *iteratedNeighbours, uint64_t *sums) {
    hipError_t e = launch_fill_u64(c.stream, sums, 0ull, 2);
    if (e != hipSuccess || c.n == 0) return e;
    hipLaunchKernelGGL(k_particle_stats, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.linkCount.as<uint32_t>(),
                       c.slotLink.as<int32_t>(), c.maxLinks, c.state.as<uint8_t>(), c.nbrCount.as<int32_t>(),
                       c.nbrExcludesLinked ? c.nbrLinked.as<int32_t>() : (const int32_t *)nullptr, c.n, activeLinks,
                       iteratedNeighbours, sums);
    return hipGetLastError();
}

hipError_t launch_link_count(Context &c) {
    hipError_t e = launch_fill_u32(c.stream, c.linkCount.as<uint32_t>(), 0u, c.n);
    if (e != hipSuccess) return e;
    e = launch_fill_u64(c.stream, c.red.as<uint64_t>() + 8, 0ull, 1);
    if (e != hipSuccess) return e;
    if (c.L > 0)
        hipLaunchKernelGGL(k_link_count, dim3(blocks_for(c.L)), dim3(TPB), 0, c.stream, c.la.as<int32_t>(),
                           c.lb.as<int32_t>(), c.L, c.n, c.linkCount.as<uint32_t>(), c.red.as<uint64_t>() + 8);
    return hipGetLastError();
}

hipError_t launch_link_fill(Context &c) {
    hipError_t e = launch_fill_u32(c.stream, c.linkCursor.as<uint32_t>(), 0u, c.n);
    if (e != hipSuccess) return e;
    if (c.L > 0)
        hipLaunchKernelGGL(k_link_fill, dim3(blocks_for(c.L)), dim3(TPB), 0, c.stream, c.la.as<int32_t>(),
                           c.lb.as<int32_t>(), c.L, c.maxLinks, c.linkCursor.as<uint32_t>(), c.slotLink.as<int32_t>());
    hipLaunchKernelGGL(k_link_sort, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.la.as<int32_t>(),
                       c.lb.as<int32_t>(), c.state.as<uint8_t>(), c.n, c.maxLinks, c.linkCount.as<uint32_t>(), c.slotLink.as<int32_t>(),
                       c.slotOther.as<int32_t>(), c.sidePos.as<int32_t>());
    return hipGetLastError();
}

hipError_t launch_ct_import(Context &c) {
    int64_t nSides = 2 * c.L;
    if (nSides == 0) return hipSuccess;
    const double *s[6];
    for (int k = 0; k < 6; k++) s[k] = c.ctStage[k].as<double>();
    if (c.prec == HDM_PREC_FP64)
        hipLaunchKernelGGL(k_ct_import<double>, dim3(blocks_for(nSides)), dim3(TPB), 0, c.stream,
                           c.sidePos.as<int32_t>(), nSides, s[0], s[1], s[2], s[3], s[4], s[5], c.ct.as<double>());
    else
        hipLaunchKernelGGL(k_ct_import<float>, dim3(blocks_for(nSides)), dim3(TPB), 0, c.stream,
                           c.sidePos.as<int32_t>(), nSides, s[0], s[1], s[2], s[3], s[4], s[5], c.ct.as<float>());
    return hipGetLastError();
}

hipError_t launch_ct_export(Context &c) {
    int64_t nSides = 2 * c.L;
    if (nSides == 0) return hipSuccess;
    double *s[6];
    for (int k = 0; k < 6; k++) s[k] = c.ctStage[k].as<double>();
    if (c.prec == HDM_PREC_FP64)
        hipLaunchKernelGGL(k_ct_export<double>, dim3(blocks_for(nSides)), dim3(TPB), 0, c.stream,
                           c.sidePos.as<int32_t>(), nSides, c.ct.as<double>(), s[0], s[1], s[2], s[3], s[4], s[5]);
    else
        hipLaunchKernelGGL(k_ct_export<float>, dim3(blocks_for(nSides)), dim3(TPB), 0, c.stream,
                           c.sidePos.as<int32_t>(), nSides, c.ct.as<float>(), s[0], s[1], s[2], s[3], s[4], s[5]);
    return hipGetLastError();
}

hipError_t launch_side_thresholds(Context &c, uint64_t *maxBits, float *l0Max, float *scale) {
    hipError_t e = launch_fill_u64(c.stream, maxBits, 0ull, 2);
    *l0Max = 0.0f;
    *scale = 0.0f;
    const int64_t nSides = 2 * c.L;
    if (e != hipSuccess || nSides == 0) return e;
    const int cap = c.numCU > 0 ? c.numCU * 8 : 2048;
    const double *ml = c.hasMaxLoad ? c.maxLoad.as<double>() : nullptr;
    int64_t nb = blocks_for(c.L);
    hipLaunchKernelGGL(k_threshold_max, dim3((unsigned)(nb < cap ? nb : cap)), dim3(TPB), 0, c.stream, c.L,
                       c.initialLen.as<double>(), ml, maxBits);
    uint64_t w[2] = {0, 0};
    e = hipMemcpyAsync(w, maxBits, sizeof(w), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return e;
    const uint32_t b0 = (uint32_t)w[0], b1 = (uint32_t)w[1];
    float tMax;
    std::memcpy(l0Max, &b0, sizeof(float));
    std::memcpy(&tMax, &b1, sizeof(float));
    float s = tMax / 65535.0f;
    if (!(s > 0.0f && s < 3e38f)) s = 0.0f;
    *scale = s;
    nb = blocks_for(nSides);
    hipLaunchKernelGGL(k_side_codes, dim3((unsigned)(nb < cap ? nb : cap)), dim3(TPB), 0, c.stream,
                       c.sidePos.as<int32_t>(), nSides, c.initialLen.as<double>(), ml, (double)s,
                       c.slotThr.as<uint16_t>());
    return hipGetLastError();
}

hipError_t launch_partner_check(Context &c, uint64_t *flag) {
    hipError_t e = launch_fill_u64(c.stream, flag, 0ull, 1);
    if (e != hipSuccess || c.n == 0) return e;
    hipLaunchKernelGGL(k_partner_check, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.linkCount.as<uint32_t>(),
                       c.slotOther.as<int32_t>(), c.maxLinks, c.n, flag);
    return hipGetLastError();
}

hipError_t launch_count_broken(Context &c, uint64_t *counter) {
    hipError_t e = launch_fill_u64(c.stream, counter, 0ull, 1);
    if (e != hipSuccess) return e;
    if (c.L == 0) return hipSuccess;
    int64_t nb = blocks_for(c.L);
    int cap = c.numCU > 0 ? c.numCU * 4 : 1024;
    int grid = (int)(nb < cap ? nb : cap);
    hipLaunchKernelGGL(k_count_broken, dim3(grid), dim3(TPB), 0, c.stream, c.state.as<uint8_t>(), c.L, counter);
    return hipGetLastError();
}

}  // namespace hdm
