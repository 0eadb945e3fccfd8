// Neighbour grid and neighbour list kernels for gfx950.
//
//   bbox           ParticleData::getLimits (ParticleData.h:47-59): min/max of
//                  the coordinates, reduced per wavefront with DPP shuffles and
//                  merged with order-independent 64-bit atomics (exact, so the
//                  result does not depend on scheduling).
//   colour_count   NeighbourGrid::colourParticles (NeighbourGrid.cpp:116-161):
//                  cell id per particle + per-cell population.
//   scan           exclusive scan of the populations -> dense cellStart (the
//                  reference keeps a hash of occupied colours -> bucket; a
//                  dense array over the X*Y*Z box answers the same lookups).
//   scatter_sort   NeighbourGrid::generateIterationList (:163-184): particles
//                  grouped by cell; each cell's members then sorted by pInd so
//                  iterationOrder is exactly "sorted by (colour, pInd)".
//   cell_records   {x, y, z, radius} of the particles in iteration order
//                  and each particle's position in it.
//   neigh_count /  NeighAccel::generateNumNeigh / iterateBlocks
//   neigh_fill     (NeighAccel.h:85-137) over NeighbourGrid::
//                  iteratePossibleNeighbours (NeighbourGrid.h:85-105): 27
//                  linear colour offsets visited in z,y,x order, members in
//                  iterationOrder, self skipped, kept if
//                  sqrt(d.d) - (r_i + r_j) < interactDist (FP64), read as 9
//                  contiguous runs of cell records.  Output is a wave-tiled
//                  ELL list (see hdm_context.h); the fill pass alone suffices
//                  when the previous build's width holds every count.
#include "hdm_context.h"
#include "kernel_common.h"

namespace hdm {

namespace {

constexpr int TPB = 256;

__global__ void __launch_bounds__(TPB) k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(TPB) k_fill_u64(uint64_t *p, uint64_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i < n) p[i] = v;
}

// out[0..2] = min keys of x,y,z; out[3..5] = max keys.
__global__ void __launch_bounds__(TPB) k_bbox(const double *__restrict__ x, const double *__restrict__ y,
                                               const double *__restrict__ z, int64_t n, uint64_t *out) {
    double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            mn[c] = v[c] < mn[c] ? v[c] : mn[c];
            mx[c] = v[c] > mx[c] ? v[c] : mx[c];
        }
    }
    __shared__ double smn[3][TPB / 64], smx[3][TPB / 64];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        mn[c] = wave_min(mn[c]);
        mx[c] = wave_max(mx[c]);
    }
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) { smn[c][w] = mn[c]; smx[c][w] = mx[c]; }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int c = threadIdx.x;
        double a = smn[c][0], b = smx[c][0];
        for (int k = 1; k < TPB / 64; k++) {
            a = smn[c][k] < a ? smn[c][k] : a;
            b = smx[c][k] > b ? smx[c][k] : b;
        }
        atomicMin((unsigned long long *)&out[c], (unsigned long long)ordered_key(a));
        atomicMax((unsigned long long *)&out[3 + c], (unsigned long long)ordered_key(b));
    }
}

struct GridArgs {
    double xMin, yMin, zMin, s;
    int32_t X, Y;
    int64_t totalCells;
};

__global__ void __launch_bounds__(TPB) k_colour_count(const double *__restrict__ x, const double *__restrict__ y,
                                                       const double *__restrict__ z, int64_t n, GridArgs g,
                                                       int32_t *__restrict__ colour, uint32_t *__restrict__ cellCount,
                                                       uint64_t *errFlag) {
    int64_t p = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    if (p >= n) return;
    int ix = (int)floor((x[p] - g.xMin) / g.s);
    int iy = (int)floor((y[p] - g.yMin) / g.s);
    int iz = (int)floor((z[p] - g.zMin) / g.s);
    int64_t c64 = (int64_t)ix + (int64_t)iy * (int64_t)g.X + (int64_t)iz * (int64_t)g.X * (int64_t)g.Y;
    int c = (int)c64;
    colour[p] = c;
    if (c < 0 || (int64_t)c >= g.totalCells) {
        atomicOr((unsigned long long *)errFlag, 1ull);
        return;
    }
    atomicAdd(&cellCount[c], 1u);
}

// ---- exclusive scan (int32), 1024 elements per block --------------------------
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = TPB * SCAN_ITEMS;

__device__ inline int32_t block_exclusive_scan(int32_t v, int32_t *total) {
    __shared__ int32_t wsum[TPB / 64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t inc = wave_inclusive_sum_i32(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < TPB / 64; k++) {
        off += (k < w) ? wsum[k] : 0;
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(TPB) k_scan_partials(const int32_t *__restrict__ in, int64_t count,
                                                        int32_t *__restrict__ partials) {
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) s += (base + k < count) ? in[base + k] : 0;
    int32_t tot;
    (void)block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the per-tile partial sums, in place.
__global__ void __launch_bounds__(TPB) k_scan_partials_scan(int32_t *partials, int64_t nParts) {
    int32_t carry = 0;
    for (int64_t base = 0; base < nParts; base += TPB) {
        int64_t i = base + threadIdx.x;
        int32_t v = (i < nParts) ? partials[i] : 0;
        int32_t tot;
        int32_t ex = block_exclusive_scan(v, &tot);
        if (i < nParts) partials[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(TPB) k_scan_apply(const int32_t *__restrict__ in, int64_t count,
                                                     const int32_t *__restrict__ partials, int32_t *__restrict__ out) {
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < count) ? in[base + k] : 0;
        s += v[k];
    }
    int32_t tot;
    int32_t ex = block_exclusive_scan(s, &tot) + partials[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < count) out[base + k] = ex;
        ex += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == TPB - 1) out[count] = ex;
}

__global__ void __launch_bounds__(TPB) k_scatter(const int32_t *__restrict__ colour, int64_t n, int64_t totalCells,
                                                  int32_t *__restrict__ cursor, int32_t *__restrict__ order) {
    int64_t p = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    if (p >= n) return;
    int32_t c = colour[p];
    if (c < 0 || (int64_t)c >= totalCells) return;  // flagged by k_colour_count
    int32_t pos = atomicAdd(&cursor[c], 1);
    order[pos] = (int32_t)p;
}

// Sorts each cell's members by particle index; reports occupied cells and
// the largest population (stats[0], stats[1]).
__global__ void __launch_bounds__(TPB) k_cell_sort(const int32_t *__restrict__ cellStart, int64_t totalCells,
                                                    int32_t *__restrict__ order, uint64_t *stats) {
    int64_t c = (int64_t)blockIdx.x * TPB + threadIdx.x;
    uint32_t occ = 0, mx = 0;
    if (c < totalCells) {
        int32_t s = cellStart[c], e = cellStart[c + 1];
        if (e > s) {
            occ = 1;
            mx = (uint32_t)(e - s);
            for (int32_t i = s + 1; i < e; i++) {
                int32_t v = order[i];
                int32_t j = i - 1;
                while (j >= s && order[j] > v) {
                    order[j + 1] = order[j];
                    j--;
                }
                order[j + 1] = v;
            }
        }
    }
    // one pair of atomics per workgroup: same-address atomics of every wave
    // serialise in L2
    __shared__ uint32_t wocc[TPB / 64], wmx[TPB / 64];
    occ = wave_sum_u32(occ);
    mx = wave_max_u32(mx);
    if ((threadIdx.x & 63) == 0) {
        wocc[threadIdx.x >> 6] = occ;
        wmx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < TPB / 64; w++) {
            occ += wocc[w];
            mx = wmx[w] > mx ? wmx[w] : mx;
        }
        if (occ) atomicAdd((unsigned long long *)&stats[0], (unsigned long long)occ);
        if (mx) atomicMax((unsigned long long *)&stats[1], (unsigned long long)mx);
    }
}

// Candidate records in iteration order: rec[i] = {x, y, z, radius} of
// particle order[i], and pos[order[i]] = i, so the list build scans each
// stencil row's members with one contiguous 32 B load per candidate.
__global__ void __launch_bounds__(TPB) k_cell_records(const double *__restrict__ x, const double *__restrict__ y,
                                                       const double *__restrict__ z, const double *__restrict__ radius,
                                                       const int32_t *__restrict__ order, int64_t n,
                                                       double4 *__restrict__ rec, float4 *__restrict__ recF,
                                                       int32_t *__restrict__ pos) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const int32_t q = order[i];
    const double4 r = make_double4(x[q], y[q], z[q], radius[q]);
    rec[i] = r;
    recF[i] = make_float4((float)r.x, (float)r.y, (float)r.z, (float)r.w);
    pos[q] = (int32_t)i;
}

struct NeighArgs {
    int64_t rowOffsets[9];  // yD*X + zD*X*Y for (zD, yD) in z, y order
    int64_t totalCells;
    double interactDist;
    int32_t width;  // ELL width (fill pass)
    // Link partners left out of the list (nullptr: none): the FP32 force
    // evaluates their contact law in its link loop instead
    // (Context::nbrExcludesLinked).
    const uint32_t *linkCount;
    const int32_t *slotOther;
    int32_t linkWidth;
};

// Longest candidate stream whose link partners are marked in a bit mask
// (longer streams compare each near candidate with the partners instead).
constexpr int32_t NEIGH_MASK_BITS = 128;  // two 64-bit words
// Candidate records a lane loads together in the list build.
constexpr int NEIGH_BATCH = 4;

// One lane per particle.  NeighbourGrid::iteratePossibleNeighbours visits
// the 27 cells col + xD + yD*X + zD*X*Y in z, y, x order, skipping targets
// outside [0, totalCells) and wrapping across rows as the linear offsets do;
// for each (zD, yD) the three cells xD = -1, 0, 1 are consecutive colours,
// so their members are one contiguous run of the iteration order
// [cellStart[lo], cellStart[hi + 1]) with lo..hi the row's targets clipped to
// the grid: the same candidates in the same order, read as cell records.
// The decisions are NeighAccel's FP64 ones: an FP32 test on squared
// distances decides wherever its rounding bound allows, NeighAccel's FP64
// expressions elsewhere.  FILL writes the first `width` entries to the
// wave-tiled list; every pass writes the count and the skipped link
// partners.  Five waves per SIMD (the software pipeline's registers fit
// without spills).
template <bool FILL>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(5)))
k_neigh(const double4 *__restrict__ crec, const float4 *__restrict__ crecF, const int32_t *__restrict__ cpos,
        int64_t n, const int32_t *__restrict__ colour, const int32_t *__restrict__ cellStart,
        const int32_t *__restrict__ order, NeighArgs a, int32_t *__restrict__ count,
        int32_t *__restrict__ linked, int32_t *__restrict__ list) {
    int64_t p = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    if (p >= n) return;
    const int32_t self = cpos[p];
    const double4 P = crec[self];
    const double xCord = P.x, yCord = P.y, zCord = P.z, rad = P.w;
    const float4 Pf = crecF[self];
    const float idf = (float)a.interactDist;
    const float Rp = idf + Pf.w;
    const float Cp = fabsf(Pf.x) + fabsf(Pf.y) + fabsf(Pf.z) + 4.0f * (fabsf(idf) + fabsf(Pf.w));
    const int64_t col = colour[p];
    // The 9 stencil rows as runs [s, s + len) of the iteration order, all
    // bounds loaded at once.  Each lane walks its rows as one stream,
    // NEIGH_BATCH candidates per step: a step takes what is left of the
    // current row and fills up from the next one, so a wave steps about
    // sum(len) / NEIGH_BATCH times for its longest lane rather than the sum
    // over rows of each row's longest lane.  Candidate order is unchanged.
    __shared__ int2 rowRun[9][TPB];
    int32_t rs[9], rl[9];
#pragma unroll
    for (int row = 0; row < 9; ++row) {
        int64_t lo = col + a.rowOffsets[row] - 1, hi = col + a.rowOffsets[row] + 1;
        lo = lo < 0 ? 0 : lo;
        hi = hi >= a.totalCells ? a.totalCells - 1 : hi;
        const bool ok = lo <= hi;
        rs[row] = cellStart[ok ? lo : 0];
        rl[row] = ok ? cellStart[hi + 1] - rs[row] : 0;
    }
#pragma unroll
    for (int row = 0; row < 9; ++row) rowRun[row][threadIdx.x] = make_int2(rs[row], rl[row]);
    // p's link partners (list exclusion): each partner's iteration position
    // is found in the runs and its index in the candidate stream marked in
    // a bit mask (once per run holding it, as a position can recur when rows
    // of a grid under 3 cells wide overlap); a near candidate is a partner
    // when its stream index is marked.  Streams longer than the mask compare
    // near candidates with the partners' indices from global memory.
    uint32_t pcnt = 0;
    uint64_t pmaskLo = 0, pmaskHi = 0;  // stream indices 0-63, 64-127
    int32_t streamLen = 0;
#pragma unroll
    for (int row = 0; row < 9; ++row) streamLen += rl[row];
    const bool useMask = streamLen <= NEIGH_MASK_BITS;
    const int64_t pbase = (p >> 6) * (int64_t)64 * a.linkWidth + (p & 63);
    if (a.linkCount) {
        pcnt = a.linkCount[p];
        if (useMask) {
            // four partners at a time: their slot words, then their
            // positions, are loaded together
            for (uint32_t k0 = 0; k0 < pcnt; k0 += 4) {
                int32_t w[4], v[4];
#pragma unroll
                for (int t = 0; t < 4; t++)
                    w[t] = a.slotOther[pbase + (int64_t)(k0 + t < pcnt ? k0 + t : k0) * 64] & SLOT_LINK_MASK;
#pragma unroll
                for (int t = 0; t < 4; t++) v[t] = cpos[w[t]];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (k0 + t >= pcnt) break;
                    int32_t pre = 0;
#pragma unroll
                    for (int row = 0; row < 9; ++row) {
                        const uint32_t off = (uint32_t)(v[t] - rs[row]);
                        if (off < (uint32_t)rl[row]) {
                            const uint32_t j = (uint32_t)pre + off;
                            if (j < 64)
                                pmaskLo |= 1ull << j;
                            else
                                pmaskHi |= 1ull << (j - 64);
                        }
                        pre += rl[row];
                    }
                }
            }
        }
    }
    auto is_partner = [&](int32_t q) {
        for (uint32_t k = 0; k < pcnt; k++)
            if ((a.slotOther[pbase + (int64_t)k * 64] & SLOT_LINK_MASK) == q) return true;
        return false;
    };
    int32_t neighbours = 0, skipped = 0;
    int32_t *out = nullptr;
    if (FILL) out = list + (p >> 6) * (int64_t)64 * a.width + (p & 63);
    int row = 0;
    int32_t jNext = 0;  // stream index of the next step's first candidate
    int2 cur = rowRun[0][threadIdx.x], nxt = rowRun[1][threadIdx.x];
    // One step of the stream: up to NEIGH_BATCH candidates (-1: none) and
    // the stream index of the first; false once the runs are used up.
    auto take = [&](int32_t (&cand)[NEIGH_BATCH], int32_t &jStart) {
        jStart = jNext;
        if (row >= 9) {
#pragma unroll
            for (int k = 0; k < NEIGH_BATCH; k++) cand[k] = -1;
            return false;
        }
        const int32_t fromCur = cur.y < NEIGH_BATCH ? cur.y : NEIGH_BATCH;
        const int32_t room = NEIGH_BATCH - fromCur;
        const int32_t fromNext = nxt.y < room ? nxt.y : room;
#pragma unroll
        for (int k = 0; k < NEIGH_BATCH; k++) {
            const int32_t c = k < fromCur ? cur.x + k : nxt.x + (k - fromCur);
            cand[k] = k < fromCur + fromNext ? c : -1;
        }
        jNext += fromCur + fromNext;
        if (cur.y > NEIGH_BATCH) {
            cur.x += NEIGH_BATCH;
            cur.y -= NEIGH_BATCH;
        } else {
            cur = make_int2(nxt.x + fromNext, nxt.y - fromNext);
            ++row;
            nxt = row + 1 < 9 ? rowRun[row + 1][threadIdx.x] : make_int2(0, 0);
        }
        return true;
    };
    // Software-pipelined: a step's records and indices are loaded while the
    // previous step is tested, so the wait at a step is for loads issued a
    // step earlier (and the list stores, which gfx9 counts with the loads,
    // come after them).
    int32_t cand[NEIGH_BATCH], j0;
    bool live = take(cand, j0);
    float4 Q[NEIGH_BATCH];
    int32_t idx[NEIGH_BATCH];  // particle indices, loaded with the records (no wait when appending)
#pragma unroll
    for (int k = 0; k < NEIGH_BATCH; k++) Q[k] = crecF[cand[k] >= 0 ? cand[k] : self];
#pragma unroll
    for (int k = 0; k < NEIGH_BATCH; k++) idx[k] = order[cand[k] >= 0 ? cand[k] : self];
    while (live) {
        int32_t candN[NEIGH_BATCH], jN;
        const bool liveN = take(candN, jN);
        float4 QN[NEIGH_BATCH];
        int32_t idxN[NEIGH_BATCH];
#pragma unroll
        for (int k = 0; k < NEIGH_BATCH; k++) QN[k] = crecF[candN[k] >= 0 ? candN[k] : self];
#pragma unroll
        for (int k = 0; k < NEIGH_BATCH; k++) idxN[k] = order[candN[k] >= 0 ? candN[k] : self];
        {
#pragma unroll
            for (int k = 0; k < NEIGH_BATCH; k++) {
                const int32_t ik = cand[k];
                if (ik < 0 || ik == self) continue;
                // FP32 first test on squared values: with a = Pf - Qf, s2 =
                // |a|^2 (two FMAs and a product) and T = fl(fl(idf + Pf.w) +
                // Qf.w), the pair is near when T - E > 0 and s2 < (T - E)^2,
                // far when s2 >= (T + E)^2 (finite), and otherwise takes the
                // FP64 test.  In units of u = 2^-24: each record component is
                // within u of its FP64 value (plus 2^-150 below the normal
                // range), so |a - (P - Q)| <= u (sum|P| + sum|Q|) + sqrt(3) u
                // |a|; s2 is within 3.01 u of |a|^2, so sqrt(s2) within 1.51 u
                // of |a|; T is within 2.01 u (|idf| + |Pw| + |Qw|) + u |T| of
                // interactDist + r_P + r_Q; rounding the squared bounds adds u
                // more.  The FP32 decision is the FP64 one whenever E >= u
                // (sum|P| + sum|Q| + 5.9 |T| + 2.01 (|idf| + |Pw| + |Qw|)),
                // plus 2^-74 for underflowed squares (the FP64 evaluation's own
                // roundings are within 2^-50 of its operands).  E is four times
                // that plus 1e-20.  Non-finite or overflowing values make E, the
                // bounds or the comparisons non-finite or false, so those
                // candidates take the FP64 test.
                const float fx = Pf.x - Q[k].x, fy = Pf.y - Q[k].y, fz = Pf.z - Q[k].z;
                const float s2 = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz));
                const float T = Rp + Q[k].w;
                const float E = __builtin_fmaf(Cp + fabsf(Q[k].x) + fabsf(Q[k].y) + fabsf(Q[k].z) +
                                                   4.0f * fabsf(Q[k].w) + 8.0f * fabsf(T),
                                               0x1p-22f, 1e-20f);
                const float tlo = T - E, thi = T + E;
                bool near;
                if (tlo > 0.0f && s2 < tlo * tlo) {
                    near = true;
                } else if (s2 >= thi * thi && thi * thi <= 3.4028234663852886e38f) {
                    near = false;
                } else {
                    const double4 Qd = crec[ik];
                    double xDist = xCord - Qd.x;
                    double xDist2 = xDist * xDist;
                    double yDist = yCord - Qd.y;
                    double yDist2 = yDist * yDist;
                    double zDist = zCord - Qd.z;
                    double zDist2 = zDist * zDist;
                    double particleDistance2 = xDist2 + yDist2 + zDist2;
                    double particleDistance = sqrt(particleDistance2);
                    double radSum = rad + Qd.w;
                    double dist = particleDistance - radSum;
                    near = dist < a.interactDist;
                }
                if (near) {
                    if (pcnt) {
                        bool partner;
                        if (useMask) {
                            const uint32_t j = (uint32_t)(j0 + k);
                            partner = ((j < 64 ? pmaskLo >> j : pmaskHi >> (j - 64)) & 1) != 0;
                        } else {
                            partner = is_partner(idx[k]);
                        }
                        if (partner) {
                            skipped++;
                            continue;
                        }
                    }
                    if (FILL) {
                        if (neighbours < a.width) out[(int64_t)neighbours * 64] = idx[k];
                    }
                    neighbours++;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NEIGH_BATCH; k++) {
            cand[k] = candN[k];
            Q[k] = QN[k];
            idx[k] = idxN[k];
        }
        j0 = jN;
        live = liveN;
    }
    count[p] = neighbours;
    linked[p] = skipped;
}

__global__ void __launch_bounds__(TPB) k_max_i32(const int32_t *__restrict__ v, int64_t n, uint64_t *out) {
    uint32_t m = 0;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        uint32_t a = (uint32_t)(v[i] < 0 ? 0 : v[i]);
        m = a > m ? a : m;
    }
    m = wave_max_u32(m);
    if ((threadIdx.x & 63) == 0 && m) atomicMax((unsigned long long *)out, (unsigned long long)m);
}

inline int blocks_for(int64_t n) { return (int)((n + TPB - 1) / TPB); }

}  // namespace

hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(TPB), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t v, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n)), dim3(TPB), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_bbox(Context &c) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    uint64_t *r = c.red.as<uint64_t>();
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(TPB), 0, c.stream, r, ordered_key(__builtin_inf()), (int64_t)3);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(TPB), 0, c.stream, r + 3, ordered_key(-__builtin_inf()), (int64_t)3);
    int64_t nb = blocks_for(c.n);
    int cap = c.numCU > 0 ? c.numCU * 4 : 1024;
    int grid = (int)(nb < cap ? nb : cap);
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_bbox, dim3(grid), dim3(TPB), 0, c.stream, c.x.as<double>(), c.y.as<double>(),
                       c.z.as<double>(), c.n, r);
    return hipGetLastError();
}

hipError_t launch_colour_count(Context &c) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    GridArgs g{c.grid.xMin, c.grid.yMin, c.grid.zMin, c.grid.spacing, c.grid.X, c.grid.Y, c.grid.totalCells};
    hipError_t e = launch_fill_u32(c.stream, c.cellCursor.as<uint32_t>(), 0u, c.grid.totalCells);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(TPB), 0, c.stream, c.red.as<uint64_t>() + 8, 0ull, (int64_t)1);
    hipLaunchKernelGGL(k_colour_count, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                       c.y.as<double>(), c.z.as<double>(), c.n, g, c.colour.as<int32_t>(),
                       c.cellCursor.as<uint32_t>(), c.red.as<uint64_t>() + 8);
    return hipGetLastError();
}

hipError_t launch_exclusive_scan_i32(Context &c, const int32_t *in, int32_t *out, int64_t count) {
    int64_t nParts = (count + SCAN_TILE - 1) / SCAN_TILE;
    if (nParts < 1) nParts = 1;
    hipError_t e = c.scanPartials.ensure(sizeof(int32_t) * (size_t)nParts);
    if (e != hipSuccess) return e;
    int32_t *parts = c.scanPartials.as<int32_t>();
    hipLaunchKernelGGL(k_scan_partials, dim3((unsigned)nParts), dim3(TPB), 0, c.stream, in, count, parts);
    hipLaunchKernelGGL(k_scan_partials_scan, dim3(1), dim3(TPB), 0, c.stream, parts, nParts);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nParts), dim3(TPB), 0, c.stream, in, count, parts, out);
    return hipGetLastError();
}

hipError_t launch_scatter_sort(Context &c) {
    hipError_t e = hipMemcpyAsync(c.cellCursor.ptr, c.cellStart.ptr, sizeof(int32_t) * (size_t)c.grid.totalCells,
                                  hipMemcpyDeviceToDevice, c.stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.colour.as<int32_t>(), c.n,
                       c.grid.totalCells, c.cellCursor.as<int32_t>(), c.order.as<int32_t>());
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(TPB), 0, c.stream, c.red.as<uint64_t>() + 9, 0ull, (int64_t)2);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks_for(c.grid.totalCells)), dim3(TPB), 0, c.stream,
                       c.cellStart.as<int32_t>(), c.grid.totalCells, c.order.as<int32_t>(), c.red.as<uint64_t>() + 9);
    e = hipGetLastError();
    return e == hipSuccess ? launch_cell_records(c) : e;
}

hipError_t launch_cell_records(Context &c) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    hipError_t e = c.cellRec.ensure(sizeof(double4) * (size_t)c.n);
    if (e == hipSuccess) e = c.cellRecF.ensure(sizeof(float4) * (size_t)c.n);
    if (e == hipSuccess) e = c.cellPos.ensure(sizeof(int32_t) * (size_t)c.n);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cell_records, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                       c.y.as<double>(), c.z.as<double>(), c.radius.as<double>(), c.order.as<int32_t>(), c.n,
                       c.cellRec.as<double4>(), c.cellRecF.as<float4>(), c.cellPos.as<int32_t>());
    c.grid.recEpoch = c.stateEpoch;
    c.grid.recTs = c.ts;
    return hipGetLastError();
}

static NeighArgs neigh_args(Context &c, double interactDist) {
    NeighArgs a;
    int i = 0;
    for (int zD = -1; zD < 2; zD++)
        for (int yD = -1; yD < 2; yD++)
            a.rowOffsets[i++] = (int64_t)yD * c.grid.X + (int64_t)zD * c.grid.X * (int64_t)c.grid.Y;
    a.totalCells = c.grid.totalCells;
    a.interactDist = interactDist;
    a.width = c.maxNeigh;
    const bool exclude = merge_linked_contacts(c);
    a.linkCount = exclude ? c.linkCount.as<uint32_t>() : nullptr;
    a.slotOther = exclude ? c.slotOther.as<int32_t>() : nullptr;
    a.linkWidth = c.maxLinks;
    return a;
}

hipError_t launch_neigh_count(Context &c, double interactDist) {
    NeighArgs a = neigh_args(c, interactDist);
    hipLaunchKernelGGL(k_neigh<false>, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream,
                       c.cellRec.as<double4>(), c.cellRecF.as<float4>(), c.cellPos.as<int32_t>(), c.n,
                       c.colour.as<int32_t>(),
                       c.cellStart.as<int32_t>(), c.order.as<int32_t>(), a, c.nbrCount.as<int32_t>(),
                       c.nbrLinked.as<int32_t>(), (int32_t *)nullptr);
    return hipGetLastError();
}

hipError_t launch_neigh_fill(Context &c, double interactDist) {
    NeighArgs a = neigh_args(c, interactDist);
    hipLaunchKernelGGL(k_neigh<true>, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream,
                       c.cellRec.as<double4>(), c.cellRecF.as<float4>(), c.cellPos.as<int32_t>(), c.n,
                       c.colour.as<int32_t>(),
                       c.cellStart.as<int32_t>(), c.order.as<int32_t>(), a, c.nbrCount.as<int32_t>(),
                       c.nbrLinked.as<int32_t>(), c.nbr.as<int32_t>());
    return hipGetLastError();
}

hipError_t launch_max_i32(Context &c, const int32_t *v, int64_t n, uint64_t *out) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(TPB), 0, c.stream, out, 0ull, (int64_t)1);
    int64_t nb = blocks_for(n);
    int cap = c.numCU > 0 ? c.numCU * 4 : 1024;
    int grid = (int)(nb < cap ? nb : cap);
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_max_i32, dim3(grid), dim3(TPB), 0, c.stream, v, n, out);
    return hipGetLastError();
}

}  // namespace hdm
