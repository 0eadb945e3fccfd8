// Redistribution order on one GPU (HiDEMDistribution.h:150-181, single rank).
//
// Particles are ordered by their cell colour on four grids of spacing s, 2s,
// 4s and 8s (s = search distance; bounding box, dims and colours as the
// neighbour grid computes them, NeighbourGrid.cpp:72-161), coarsest grid
// first (ParallelSort::sortIndirectNLevel over grids[3..0], ParallelSort.h
// :84-89), which keeps particles that share a cell, a 2x2x2 block of cells,
// ... next to each other in memory: the fused step's partner gathers then
// hit the lines its neighbours just loaded.  Ties go by the old index (the
// reference's std::sort leaves them unspecified).  Links are then ordered
// by their old (A, B) (LinkData::getNewOrder, LinkData.h:166-181; ties by
// link index) and renamed to the new particle indices (reorderLinkTargets,
// :116-137); each link's two CT sides move with it.
//
// Both orders are least-significant-key-first passes of a stable radix sort
// (32-bit keys gathered through the permutation built so far), so the result
// is the lexicographic order with the old index as the last key.  Every
// array is then gathered once through the permutation into a scratch buffer
// and copied back.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "hdm_context.h"
#include "kernel_common.h"

namespace hdm {
namespace {

constexpr int TPB = 256;
inline int blocks_for(int64_t n) { return (int)((n + TPB - 1) / TPB); }

// Cell colour on one grid (k_colour_count's expressions), as a sort key.
__global__ void __launch_bounds__(TPB) k_reorder_colour(const double *__restrict__ x, const double *__restrict__ y,
                                                         const double *__restrict__ z, int64_t n, double xMin,
                                                         double yMin, double zMin, double s, int32_t X, int32_t Y,
                                                         const int32_t *__restrict__ perm, uint32_t *__restrict__ key) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
    const int ix = (int)floor((x[p] - xMin) / s);
    const int iy = (int)floor((y[p] - yMin) / s);
    const int iz = (int)floor((z[p] - zMin) / s);
    const int64_t c64 = (int64_t)ix + (int64_t)iy * (int64_t)X + (int64_t)iz * (int64_t)X * (int64_t)Y;
    key[i] = (uint32_t)(int32_t)c64;
}

__global__ void __launch_bounds__(TPB) k_iota(int32_t *__restrict__ v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) v[i] = (int32_t)i;
}

__global__ void __launch_bounds__(TPB) k_gather_key(const int32_t *__restrict__ src, const int32_t *__restrict__ perm,
                                                     int64_t n, uint32_t *__restrict__ key) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) key[i] = (uint32_t)src[perm[i]];
}

__global__ void __launch_bounds__(TPB) k_invert(const int32_t *__restrict__ perm, int64_t n, int32_t *__restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) inv[perm[i]] = (int32_t)i;
}

// out[i] = in[perm[i]], T-sized elements
template <class T>
__global__ void __launch_bounds__(TPB) k_gather(const T *__restrict__ in, const int32_t *__restrict__ perm, int64_t n,
                                                 T *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}

// New link j: endpoints of old link lo[j] renamed through inv.
__global__ void __launch_bounds__(TPB) k_link_targets(const int32_t *__restrict__ a, const int32_t *__restrict__ lo,
                                                       const int32_t *__restrict__ inv, int64_t L,
                                                       int32_t *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (j < L) out[j] = inv[a[lo[j]]];
}

// Side 2j + e of the new order = side 2 lo[j] + e of the old (checkpoint order).
__global__ void __launch_bounds__(TPB) k_gather_sides(const double *__restrict__ in, const int32_t *__restrict__ lo,
                                                       int64_t L, double *__restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (s < 2 * L) out[s] = in[2 * (int64_t)lo[s >> 1] + (s & 1)];
}

// ---- stable LSD radix sort of (32-bit key, index) pairs -----------------------
// One pass per 8-bit digit: k_radix_hist counts the digits of each tile of
// RADIX_TILE items (digit-major table hist[d * tiles + t]), an exclusive scan
// of that table gives every (digit, tile) its first output slot, and
// k_radix_scatter moves the tile's items there in input order: rounds of TPB
// items, each wave ranking its lanes among equal digits by ballots, the
// waves of a round in order, a running offset per digit across rounds.
constexpr int RADIX = 256;
constexpr int RADIX_ROUNDS = 16;
constexpr int RADIX_TILE = TPB * RADIX_ROUNDS;
static_assert(TPB == RADIX, "one thread per digit in the radix kernels");

__device__ __forceinline__ uint32_t radix_digit(uint32_t key, int shift, uint32_t mask) {
    return (key >> shift) & mask;
}

__global__ void __launch_bounds__(TPB) k_radix_hist(const uint32_t *__restrict__ key, int64_t n, int shift,
                                                     uint32_t mask, int32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[RADIX];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RADIX_TILE;
    for (int r = 0; r < RADIX_ROUNDS; r++) {
        const int64_t i = base + (int64_t)r * TPB + threadIdx.x;
        if (i < n) atomicAdd(&cnt[radix_digit(key[i], shift, mask)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = (int32_t)cnt[threadIdx.x];
}

__global__ void __launch_bounds__(TPB) k_radix_scatter(const uint32_t *__restrict__ key,
                                                        const int32_t *__restrict__ val, int64_t n, int shift,
                                                        uint32_t mask, const int32_t *__restrict__ offsets,
                                                        uint32_t *__restrict__ keyOut, int32_t *__restrict__ valOut) {
    constexpr int W = TPB / 64;
    __shared__ int32_t run[RADIX];     // next output slot of each digit
    __shared__ int32_t wcnt[W][RADIX]; // per wave: digit count, then its first slot
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    run[threadIdx.x] = offsets[(int64_t)threadIdx.x * gridDim.x + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * RADIX_TILE;
    for (int r = 0; r < RADIX_ROUNDS; r++) {
#pragma unroll
        for (int k = 0; k < W; k++) wcnt[k][threadIdx.x] = 0;
        __syncthreads();
        const int64_t i = base + (int64_t)r * TPB + threadIdx.x;
        const bool valid = i < n;
        const uint32_t kv = valid ? key[i] : 0u;
        const int32_t vv = valid ? val[i] : 0;
        const uint32_t d = radix_digit(kv, shift, mask);
        // lanes of this wave holding the same digit
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            same &= bit ? bal : ~bal;
        }
        const int rank = __popcll(same & below);
        if (valid && rank == 0) wcnt[w][d] = __popcll(same);
        __syncthreads();
        {
            int32_t acc = run[threadIdx.x];
#pragma unroll
            for (int k = 0; k < W; k++) {
                const int32_t cnt = wcnt[k][threadIdx.x];
                wcnt[k][threadIdx.x] = acc;
                acc += cnt;
            }
            run[threadIdx.x] = acc;
        }
        __syncthreads();
        if (valid) {
            const int32_t pos = wcnt[w][d] + rank;
            keyOut[pos] = kv;
            valOut[pos] = vv;
        }
        __syncthreads();
    }
}

int key_bits(int64_t cells) {
    int b = 1;
    while (b < 32 && ((int64_t)1 << b) < cells) b++;
    return b;
}

int64_t radix_tiles(int64_t n) { return (n + RADIX_TILE - 1) / RADIX_TILE; }

// Temporary storage of radix_pass over n items: a key and an index buffer
// for the passes between input and output, the digit table and its scan.
size_t radix_temp_bytes(int64_t n) {
    const size_t table = (size_t)RADIX * (size_t)radix_tiles(n) + 1;
    return 2 * sizeof(uint32_t) * (size_t)n + 2 * sizeof(int32_t) * table + 256;
}

// Stable sort of perm by the low `bits` bits of key (both length n) into
// (keyOut, permOut); tmp holds radix_temp_bytes(n) (sized beforehand, so no
// pass frees memory an earlier one still uses).
hipError_t radix_pass(Context &c, DevBuf &tmp, const uint32_t *key, uint32_t *keyOut, const int32_t *perm,
                      int32_t *permOut, int64_t n, int bits) {
    if (n == 0) return hipSuccess;
    if (n > INT32_MAX || bits < 1 || bits > 32 || tmp.bytes < radix_temp_bytes(n)) return hipErrorInvalidValue;
    const int64_t tiles = radix_tiles(n);
    const size_t table = (size_t)RADIX * (size_t)tiles + 1;
    uint32_t *tKey = tmp.as<uint32_t>();
    int32_t *tVal = reinterpret_cast<int32_t *>(tKey + n);
    int32_t *hist = tVal + n;
    int32_t *offs = hist + table;
    const int passes = (bits + 7) / 8;
    const uint32_t *srcKey = key;
    const int32_t *srcVal = perm;
    for (int ps = 0; ps < passes; ps++) {
        const int shift = 8 * ps;
        const int width = bits - shift < 8 ? bits - shift : 8;
        const uint32_t mask = (1u << width) - 1u;
        // the last pass writes the output; earlier ones alternate so that it does
        const bool toOut = ((passes - 1 - ps) & 1) == 0;
        uint32_t *dKey = toOut ? keyOut : tKey;
        int32_t *dVal = toOut ? permOut : tVal;
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)tiles), dim3(TPB), 0, c.stream, srcKey, n, shift, mask, hist);
        hipError_t e = launch_exclusive_scan_i32(c, hist, offs, (int64_t)RADIX * tiles);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)tiles), dim3(TPB), 0, c.stream, srcKey, srcVal, n, shift,
                           mask, offs, dKey, dVal);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        srcKey = dKey;
        srcVal = dVal;
    }
    return hipSuccess;
}

// Gathers `b` (n elements of T) through perm and copies the result back;
// scratch holds at least n elements of T.
template <class T>
hipError_t permute(Context &c, DevBuf &b, const int32_t *perm, int64_t n, DevBuf &scratch) {
    if (!b.ptr || n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather<T>, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, b.as<T>(), perm, n, scratch.as<T>());
    return hipMemcpyAsync(b.ptr, scratch.ptr, sizeof(T) * (size_t)n, hipMemcpyDeviceToDevice, c.stream);
}

}  // namespace

hipError_t launch_reorder(Context &c, double searchDistance, int64_t *order, int64_t *linkOrder) {
    const int64_t n = c.n, L = c.L;
    hipError_t e = sync_state(c);
    if (e != hipSuccess) return e;
    // bounding box of all particles (shared by the four grids)
    e = launch_bbox(c);
    if (e != hipSuccess) return e;
    uint64_t keys[6];
    e = hipMemcpyAsync(keys, c.red.ptr, sizeof(keys), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return e;
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) {
        mn[k] = ordered_value(keys[k]);
        mx[k] = ordered_value(keys[3 + k]);
    }
    const int64_t most = n > L ? n : L;
    DevBuf keyA, keyB, permA, permB, inv, tmp, scratch;
    for (DevBuf *b : {&keyA, &keyB, &permA, &permB, &inv}) {
        e = b->ensure(sizeof(uint32_t) * (size_t)most + 4);
        if (e != hipSuccess) return e;
    }
    // grid shape of each level, and the sort storage every pass needs
    int32_t dimX[4], dimY[4], levelBits[4];
    size_t tmpBytes = 0;
    for (int lvl = 0; lvl < 4; lvl++) {
        const double s = searchDistance * (double)(1 << lvl);
        const double ex = std::floor((mx[0] - mn[0]) / s), ey = std::floor((mx[1] - mn[1]) / s),
                     ez = std::floor((mx[2] - mn[2]) / s);
        if (!(ex < 1e9 && ey < 1e9 && ez < 1e9)) return hipErrorInvalidValue;
        dimX[lvl] = 1 + (int32_t)ex;
        dimY[lvl] = 1 + (int32_t)ey;
        const double cells = (double)dimX[lvl] * (double)dimY[lvl] * (double)(1 + (int32_t)ez);
        if (cells > (double)(1ll << 30)) return hipErrorInvalidValue;
        levelBits[lvl] = key_bits((int64_t)cells);
    }
    const int linkBits = key_bits(n);
    tmpBytes = radix_temp_bytes(most);
    if ((e = tmp.ensure(tmpBytes)) != hipSuccess) return e;
    if ((e = scratch.ensure(sizeof(double) * (size_t)(2 * L > n ? 2 * L : n))) != hipSuccess) return e;
    // particles: passes on the colour at s, 2s, 4s, 8s (the last pass is
    // the most significant key)
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, permA.as<int32_t>(), n);
    for (int lvl = 0; lvl < 4; lvl++) {
        const double s = searchDistance * (double)(1 << lvl);
        hipLaunchKernelGGL(k_reorder_colour, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                           c.y.as<double>(), c.z.as<double>(), n, mn[0], mn[1], mn[2], s, dimX[lvl], dimY[lvl],
                           permA.as<int32_t>(), keyA.as<uint32_t>());
        e = radix_pass(c, tmp, keyA.as<uint32_t>(), keyB.as<uint32_t>(), permA.as<int32_t>(), permB.as<int32_t>(), n,
                       levelBits[lvl]);
        if (e != hipSuccess) return e;
        swap_buf(permA, permB);
    }
    const int32_t *perm = permA.as<int32_t>();
    hipLaunchKernelGGL(k_invert, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, perm, n, inv.as<int32_t>());
    for (DevBuf *b : {&c.x, &c.y, &c.z, &c.tx, &c.ty, &c.tz, &c.radius}) {
        if ((e = permute<double>(c, *b, perm, n, scratch)) != hipSuccess) return e;
    }
    for (int k = 0; k < 6; k++) {
        if ((e = permute<double>(c, c.diff[k], perm, n, scratch)) != hipSuccess) return e;
        e = c.prec == HDM_PREC_FP64 ? permute<double>(c, c.force[k], perm, n, scratch)
                                    : permute<float>(c, c.force[k], perm, n, scratch);
        if (e != hipSuccess) return e;
    }
    if ((e = permute<uint8_t>(c, c.type, perm, n, scratch)) != hipSuccess) return e;
    if (c.nBlocks > 0 && (e = permute<int32_t>(c, c.blockId, perm, n, scratch)) != hipSuccess) return e;
    if (order) {
        std::vector<int32_t> h((size_t)n);
        e = hipMemcpyAsync(h.data(), perm, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) return e;
        for (int64_t i = 0; i < n; i++) order[i] = h[(size_t)i];
    }
    if (L == 0) return hipStreamSynchronize(c.stream);
    // links: passes on old B, then old A
    int32_t *lperm = permA.as<int32_t>();  // the particle permutation is no longer needed (inv holds it)
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(L)), dim3(TPB), 0, c.stream, lperm, L);
    for (DevBuf *src : {&c.lb, &c.la}) {
        hipLaunchKernelGGL(k_gather_key, dim3(blocks_for(L)), dim3(TPB), 0, c.stream, src->as<int32_t>(),
                           permA.as<int32_t>(), L, keyA.as<uint32_t>());
        e = radix_pass(c, tmp, keyA.as<uint32_t>(), keyB.as<uint32_t>(), permA.as<int32_t>(), permB.as<int32_t>(), L,
                       linkBits);
        if (e != hipSuccess) return e;
        swap_buf(permA, permB);
    }
    lperm = permA.as<int32_t>();
    for (DevBuf *b : {&c.la, &c.lb}) {
        hipLaunchKernelGGL(k_link_targets, dim3(blocks_for(L)), dim3(TPB), 0, c.stream, b->as<int32_t>(), lperm,
                           inv.as<int32_t>(), L, scratch.as<int32_t>());
        e = hipMemcpyAsync(b->ptr, scratch.ptr, sizeof(int32_t) * (size_t)L, hipMemcpyDeviceToDevice, c.stream);
        if (e != hipSuccess) return e;
    }
    if ((e = permute<uint8_t>(c, c.state, lperm, L, scratch)) != hipSuccess) return e;
    if ((e = permute<double>(c, c.initialLen, lperm, L, scratch)) != hipSuccess) return e;
    if (c.hasMaxLoad && (e = permute<double>(c, c.maxLoad, lperm, L, scratch)) != hipSuccess) return e;
    for (int k = 0; k < 6; k++) {
        if (!c.ctStage[k].ptr) continue;
        hipLaunchKernelGGL(k_gather_sides, dim3(blocks_for(2 * L)), dim3(TPB), 0, c.stream, c.ctStage[k].as<double>(),
                           lperm, L, scratch.as<double>());
        e = hipMemcpyAsync(c.ctStage[k].ptr, scratch.ptr, sizeof(double) * (size_t)(2 * L), hipMemcpyDeviceToDevice,
                           c.stream);
        if (e != hipSuccess) return e;
    }
    if (linkOrder) {
        std::vector<int32_t> h((size_t)L);
        e = hipMemcpyAsync(h.data(), lperm, sizeof(int32_t) * (size_t)L, hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) return e;
        for (int64_t j = 0; j < L; j++) linkOrder[j] = h[(size_t)j];
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c.stream);
}

}  // namespace hdm
