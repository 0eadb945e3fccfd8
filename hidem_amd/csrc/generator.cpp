// Seeded FCC block and link generator (setup, host side).
//
// The reference generators are non-deterministic (std::random_device seeding,
// FCCParticleGenerator.h:172-173; OpenMP-critical append order, :200-206;
// atomic link order, LinkGenerator.h:101), so a run cannot be reproduced from
// a seed.  This generator is deterministic and order-canonical:
//   * FCC unit cell basis of FCCParticleGenerator::createUnitCell
//     (FCCParticleGenerator.h:47-50): (0,0,0), (0,a/2,a/2), (a/2,0,a/2),
//     (a/2,a/2,0);
//   * particle index = global key = rank(unit cell)*4 + basis, where the unit
//     cells are ranked in lattice order ((k*ny + j)*nx + i) or along a Morton
//     (Z-order) curve, the cell-local order the reference's redistribution
//     sort produces (HiDEMDistribution.h:150-169);
//   * jitter: each coordinate + jitter*(2u-1), u from a counter-based
//     SplitMix64 hash of (seed, 3*site + component), site = lattice index;
//   * links: pairs with dist <= (r_i + r_j)*linkGenScale, dist computed as in
//     LinkGenerator::genNewLinks (LinkGenerator.h:83-89); end A is the smaller
//     key (:64), links sorted by (A, B); initialLen = dist (:99);
//   * per-link fracture threshold threshold*(1 + spread*(2u-1)), u hashed
//     from (seed, A, B) so it does not depend on link order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hidem/hdm_abi.h"

namespace {

inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double uniform01(uint64_t seed, uint64_t counter) {
    uint64_t z = splitmix64(seed ^ splitmix64(counter));
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Interleaves the low 21 bits of x, y, z (x in bit 0): Z-order curve key.
uint64_t spread_bits3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
uint64_t morton3(uint64_t x, uint64_t y, uint64_t z) { return spread_bits3(x) | (spread_bits3(y) << 1) | (spread_bits3(z) << 2); }

int nthreads() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    if (h > 16) h = 16;
    return (int)h;
}

template <class F>
void parallel_for(int64_t n, F f) {
    int T = nthreads();
    if (n < 4096) T = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++) {
        int64_t b = n * t / T, e = n * (t + 1) / T;
        th.emplace_back([=]() { f(t, b, e); });
    }
    for (auto &x : th) x.join();
}

struct CellList {
    double xMin, yMin, zMin, s;
    int64_t X, Y, Z;
    std::vector<int64_t> start;  // X*Y*Z + 1
    std::vector<int32_t> items;

    void build(const hdm_particles_host *p, double cell) {
        int64_t n = p->n;
        xMin = yMin = zMin = INFINITY;
        double xMax = -INFINITY, yMax = -INFINITY, zMax = -INFINITY;
        for (int64_t i = 0; i < n; i++) {
            xMin = std::min(xMin, p->xCord[i]); xMax = std::max(xMax, p->xCord[i]);
            yMin = std::min(yMin, p->yCord[i]); yMax = std::max(yMax, p->yCord[i]);
            zMin = std::min(zMin, p->zCord[i]); zMax = std::max(zMax, p->zCord[i]);
        }
        s = cell;
        X = 1 + (int64_t)std::floor((xMax - xMin) / s);
        Y = 1 + (int64_t)std::floor((yMax - yMin) / s);
        Z = 1 + (int64_t)std::floor((zMax - zMin) / s);
        start.assign((size_t)(X * Y * Z + 1), 0);
        std::vector<int64_t> cellOf((size_t)n);
        for (int64_t i = 0; i < n; i++) {
            int64_t c = cell_of(p->xCord[i], p->yCord[i], p->zCord[i]);
            cellOf[(size_t)i] = c;
            start[(size_t)c + 1]++;
        }
        for (size_t c = 1; c < start.size(); c++) start[c] += start[c - 1];
        items.assign((size_t)n, 0);
        std::vector<int64_t> cur(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; i++) items[(size_t)cur[(size_t)cellOf[(size_t)i]]++] = (int32_t)i;
    }
    int64_t clampi(int64_t v, int64_t hi) const { return v < 0 ? 0 : (v >= hi ? hi - 1 : v); }
    int64_t cell_of(double x, double y, double z) const {
        int64_t ix = clampi((int64_t)std::floor((x - xMin) / s), X);
        int64_t iy = clampi((int64_t)std::floor((y - yMin) / s), Y);
        int64_t iz = clampi((int64_t)std::floor((z - zMin) / s), Z);
        return ix + X * (iy + Y * iz);
    }
};

// Visits candidate partners q > p within the link distance, ascending q.
template <class F>
void for_link_partners(const hdm_particles_host *p, const CellList &cl, double linkGenScale, int64_t i,
                       std::vector<int32_t> &scratch, F f) {
    double x = p->xCord[i], y = p->yCord[i], z = p->zCord[i];
    int64_t ix = cl.clampi((int64_t)std::floor((x - cl.xMin) / cl.s), cl.X);
    int64_t iy = cl.clampi((int64_t)std::floor((y - cl.yMin) / cl.s), cl.Y);
    int64_t iz = cl.clampi((int64_t)std::floor((z - cl.zMin) / cl.s), cl.Z);
    scratch.clear();
    for (int64_t dz = -1; dz <= 1; dz++)
        for (int64_t dy = -1; dy <= 1; dy++)
            for (int64_t dx = -1; dx <= 1; dx++) {
                int64_t jx = ix + dx, jy = iy + dy, jz = iz + dz;
                if (jx < 0 || jy < 0 || jz < 0 || jx >= cl.X || jy >= cl.Y || jz >= cl.Z) continue;
                int64_t c = jx + cl.X * (jy + cl.Y * jz);
                for (int64_t k = cl.start[(size_t)c]; k < cl.start[(size_t)c + 1]; k++) {
                    int32_t q = cl.items[(size_t)k];
                    if (q <= i) continue;
                    double xDist = p->xCord[q] - p->xCord[i];
                    double yDist = p->yCord[q] - p->yCord[i];
                    double zDist = p->zCord[q] - p->zCord[i];
                    double dist = std::sqrt((xDist * xDist) + (yDist * yDist) + (zDist * zDist));
                    double radiusSum = p->radius[q] + p->radius[i];
                    if (dist > radiusSum * linkGenScale) continue;
                    scratch.push_back(q);
                }
            }
    std::sort(scratch.begin(), scratch.end());
    for (int32_t q : scratch) f(q);
}

double max_radius(const hdm_particles_host *p) {
    double r = 0;
    for (int64_t i = 0; i < p->n; i++) r = std::max(r, p->radius[i]);
    return r;
}

}  // namespace

extern "C" {

int hdm_gen_fcc_count(const hdm_fcc_spec *spec, int64_t *nParticles) {
    if (!spec || !nParticles || spec->nx <= 0 || spec->ny <= 0 || spec->nz <= 0) return HDM_E_INVALID;
    *nParticles = (int64_t)spec->nx * spec->ny * spec->nz * 4;
    return HDM_OK;
}

int hdm_gen_fcc_fill(const hdm_fcc_spec *spec, hdm_particles_host *out) {
    int64_t n = 0;
    int r = hdm_gen_fcc_count(spec, &n);
    if (r != HDM_OK) return r;
    if (!out || out->n != n || !out->xCord || !out->yCord || !out->zCord || !out->radius) return HDM_E_INVALID;
    const double a = spec->a, h = spec->a / 2.0;
    const double basis[4][3] = {{0, 0, 0}, {0, h, h}, {h, 0, h}, {h, h, 0}};
    const int64_t nx = spec->nx, ny = spec->ny, nz = spec->nz;
    // rank -> unit cell (lattice order: x fastest; Morton: sorted by Z-order code)
    std::vector<int64_t> cellOfRank;
    if (spec->order == HDM_ORDER_MORTON) {
        cellOfRank.resize((size_t)(nx * ny * nz));
        std::vector<std::pair<uint64_t, int64_t>> keyed((size_t)(nx * ny * nz));
        parallel_for(nx * ny * nz, [&](int, int64_t b, int64_t e) {
            for (int64_t c = b; c < e; c++) {
                int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
                keyed[(size_t)c] = {morton3((uint64_t)i, (uint64_t)j, (uint64_t)k), c};
            }
        });
        std::sort(keyed.begin(), keyed.end());
        for (size_t r = 0; r < keyed.size(); r++) cellOfRank[r] = keyed[r].second;
    } else if (spec->order != HDM_ORDER_LATTICE) {
        return HDM_E_INVALID;
    }
    parallel_for(n, [&](int, int64_t b, int64_t e) {
        for (int64_t idx = b; idx < e; idx++) {
            int64_t cell = idx >> 2;
            if (!cellOfRank.empty()) cell = cellOfRank[(size_t)cell];
            int bi = (int)(idx & 3);
            int64_t i = cell % nx, j = (cell / nx) % ny, k = cell / (nx * ny);
            double pos[3] = {spec->origin[0] + (double)i * a + basis[bi][0],
                             spec->origin[1] + (double)j * a + basis[bi][1],
                             spec->origin[2] + (double)k * a + basis[bi][2]};
            for (int c = 0; c < 3; c++) {
                // keyed by the lattice site, so both storage orders hold the same system
                double u = uniform01(spec->seed, (uint64_t)(cell * 4 + bi) * 3u + (uint64_t)c);
                pos[c] += spec->jitter * (2.0 * u - 1.0);
            }
            out->xCord[idx] = pos[0];
            out->yCord[idx] = pos[1];
            out->zCord[idx] = pos[2];
            out->radius[idx] = spec->radius;
            double *zeros[] = {out->xTheta, out->yTheta, out->zTheta, out->xDiff, out->yDiff, out->zDiff,
                               out->uDiff, out->vDiff, out->wDiff, out->xForce, out->yForce, out->zForce,
                               out->uForce, out->vForce, out->wForce};
            for (double *z : zeros)
                if (z) z[idx] = 0.0;
            if (out->type) out->type[idx] = (pos[0] < spec->clampXBelow) ? (uint8_t)HDM_TYPE_FIXED : (uint8_t)0;
        }
    });
    return HDM_OK;
}

int hdm_gen_links_count(const hdm_particles_host *p, double linkGenScale, int64_t *nLinks) {
    if (!p || !nLinks || p->n < 0 || !(linkGenScale > 0)) return HDM_E_INVALID;
    if (p->n == 0) {
        *nLinks = 0;
        return HDM_OK;
    }
    CellList cl;
    cl.build(p, 2.0 * max_radius(p) * linkGenScale);
    int T = nthreads();
    std::vector<int64_t> counts((size_t)T, 0);
    parallel_for(p->n, [&](int t, int64_t b, int64_t e) {
        std::vector<int32_t> scratch;
        int64_t cnt = 0;
        for (int64_t i = b; i < e; i++) for_link_partners(p, cl, linkGenScale, i, scratch, [&](int32_t) { cnt++; });
        counts[(size_t)t] = cnt;
    });
    int64_t tot = 0;
    for (auto c : counts) tot += c;
    *nLinks = tot;
    return HDM_OK;
}

int hdm_gen_links_fill(const hdm_particles_host *p, double linkGenScale, double threshold, double spread,
                       uint64_t seed, hdm_links_host *out) {
    if (!p || !out || !(linkGenScale > 0)) return HDM_E_INVALID;
    if (out->n > 0 && (!out->localPartIndA || !out->localPartIndB || !out->state || !out->initialLen))
        return HDM_E_INVALID;
    if (p->n == 0) return out->n == 0 ? HDM_OK : HDM_E_INVALID;
    CellList cl;
    cl.build(p, 2.0 * max_radius(p) * linkGenScale);
    int T = nthreads();
    if (p->n < 4096) T = 1;
    std::vector<std::vector<int32_t>> pa((size_t)T), pb((size_t)T);
    std::vector<std::vector<double>> len((size_t)T);
    parallel_for(p->n, [&](int t, int64_t b, int64_t e) {
        std::vector<int32_t> scratch;
        for (int64_t i = b; i < e; i++)
            for_link_partners(p, cl, linkGenScale, i, scratch, [&](int32_t q) {
                double xDist = p->xCord[q] - p->xCord[i];
                double yDist = p->yCord[q] - p->yCord[i];
                double zDist = p->zCord[q] - p->zCord[i];
                pa[(size_t)t].push_back((int32_t)i);
                pb[(size_t)t].push_back(q);
                len[(size_t)t].push_back(std::sqrt((xDist * xDist) + (yDist * yDist) + (zDist * zDist)));
            });
    });
    int64_t tot = 0;
    for (auto &v : pa) tot += (int64_t)v.size();
    if (tot != out->n) return HDM_E_INVALID;
    int64_t pos = 0;
    for (int t = 0; t < T; t++) {
        for (size_t k = 0; k < pa[(size_t)t].size(); k++, pos++) {
            int32_t A = pa[(size_t)t][k], B = pb[(size_t)t][k];
            out->localPartIndA[pos] = A;
            out->localPartIndB[pos] = B;
            out->state[pos] = 1;
            out->initialLen[pos] = len[(size_t)t][k];
            if (out->maxLoad) {
                double u = uniform01(seed ^ 0x5851F42D4C957F2Dull, ((uint64_t)(uint32_t)A << 32) | (uint32_t)B);
                out->maxLoad[pos] = threshold * (1.0 + spread * (2.0 * u - 1.0));
            }
        }
    }
    for (int c = 0; c < 6; c++)
        if (out->CT[c]) std::memset(out->CT[c], 0, sizeof(double) * 2 * (size_t)out->n);
    return HDM_OK;
}

}  // extern "C"

// ---- batched form ------------------------------------------------------------
extern "C" {

int hdm_gen_batch_layout(const hdm_batch_spec *spec, hdm_batch_block *blocks, int64_t *nParticles) {
    if (!spec || !blocks || !nParticles || spec->nBlocks <= 0 || spec->minParticles < 4 ||
        spec->maxParticles < spec->minParticles || !(spec->a > 0))
        return HDM_E_INVALID;
    const int B = spec->nBlocks;
    const double lo = std::log((double)spec->minParticles), hi = std::log((double)spec->maxParticles);
    for (int b = 0; b < B; b++) {
        double u = uniform01(spec->seed ^ 0xB10CB10CB10Cull, (uint64_t)b * 4 + 0);
        double target = std::exp(lo + (hi - lo) * u);
        // nx = 2 ny = 2 nz unit cells, 4 particles each: 8 ny^3 = target/... -> ny = cbrt(target / 8)
        int ny = std::max(2, (int)std::lround(std::cbrt(target / 8.0)));
        hdm_batch_block &k = blocks[b];
        k.nx = 2 * ny;
        k.ny = ny;
        k.nz = ny;
        k.particleCount = (int64_t)k.nx * k.ny * k.nz * 4;
        double ut = uniform01(spec->seed ^ 0xB10CB10CB10Cull, (uint64_t)b * 4 + 1);
        double uw = uniform01(spec->seed ^ 0xB10CB10CB10Cull, (uint64_t)b * 4 + 2);
        k.threshold = spec->thrLo + (spec->thrHi - spec->thrLo) * ut;
        k.waterLevel = spec->waterLo + (spec->waterHi - spec->waterLo) * uw;  // fraction; made absolute below
    }
    // shelf packing: 16 columns; column width / row depth = largest block in it
    const int cols = 16, rows = (B + cols - 1) / cols;
    std::vector<double> colW((size_t)cols, 0.0), rowD((size_t)rows, 0.0);
    for (int b = 0; b < B; b++) {
        colW[(size_t)(b % cols)] = std::max(colW[(size_t)(b % cols)], blocks[b].nx * spec->a);
        rowD[(size_t)(b / cols)] = std::max(rowD[(size_t)(b / cols)], blocks[b].ny * spec->a);
    }
    int64_t start = 0;
    for (int b = 0; b < B; b++) {
        hdm_batch_block &k = blocks[b];
        double x = 0, y = 0;
        for (int c = 0; c < b % cols; c++) x += colW[(size_t)c] + spec->gap;
        for (int r = 0; r < b / cols; r++) y += rowD[(size_t)r] + spec->gap;
        k.origin[0] = x;
        k.origin[1] = y;
        k.origin[2] = 0.0;
        double height = k.nz * spec->a;
        k.waterLevel = k.origin[2] + k.waterLevel * height;
        k.particleStart = start;
        start += k.particleCount;
    }
    *nParticles = start;
    return HDM_OK;
}

int hdm_gen_batch_fill(const hdm_batch_spec *spec, const hdm_batch_block *blocks, hdm_particles_host *out,
                       int32_t *blockId) {
    if (!spec || !blocks || !out) return HDM_E_INVALID;
    for (int b = 0; b < spec->nBlocks; b++) {
        const hdm_batch_block &k = blocks[b];
        hdm_fcc_spec fs{};
        fs.nx = k.nx;
        fs.ny = k.ny;
        fs.nz = k.nz;
        fs.a = spec->a;
        for (int c = 0; c < 3; c++) fs.origin[c] = k.origin[c];
        fs.radius = spec->radius;
        fs.jitter = spec->jitter;
        fs.seed = spec->seed + 1000003ull * (uint64_t)(b + 1);
        fs.clampXBelow = k.origin[0] + spec->clampX;
        fs.order = HDM_ORDER_MORTON;
        hdm_particles_host sub{};
        int64_t off = k.particleStart;
        sub.n = k.particleCount;
        auto at = [&](double *p) { return p ? p + off : nullptr; };
        sub.xCord = at(out->xCord); sub.yCord = at(out->yCord); sub.zCord = at(out->zCord);
        sub.xTheta = at(out->xTheta); sub.yTheta = at(out->yTheta); sub.zTheta = at(out->zTheta);
        sub.xDiff = at(out->xDiff); sub.yDiff = at(out->yDiff); sub.zDiff = at(out->zDiff);
        sub.uDiff = at(out->uDiff); sub.vDiff = at(out->vDiff); sub.wDiff = at(out->wDiff);
        sub.radius = at(out->radius);
        sub.xForce = at(out->xForce); sub.yForce = at(out->yForce); sub.zForce = at(out->zForce);
        sub.uForce = at(out->uForce); sub.vForce = at(out->vForce); sub.wForce = at(out->wForce);
        sub.type = out->type ? out->type + off : nullptr;
        int r = hdm_gen_fcc_fill(&fs, &sub);
        if (r != HDM_OK) return r;
        if (blockId)
            for (int64_t i = 0; i < k.particleCount; i++) blockId[off + i] = b;
    }
    return HDM_OK;
}

int hdm_gen_batch_thresholds(const hdm_batch_spec *spec, const hdm_batch_block *blocks, const int32_t *blockId,
                             hdm_links_host *links) {
    if (!spec || !blocks || !blockId || !links || !links->maxLoad) return HDM_E_INVALID;
    for (int64_t l = 0; l < links->n; l++) {
        int32_t b = blockId[links->localPartIndA[l]];
        if (b < 0 || b >= spec->nBlocks || blockId[links->localPartIndB[l]] != b) return HDM_E_INVALID;
        links->maxLoad[l] *= blocks[b].threshold;
    }
    return HDM_OK;
}

}  // extern "C"
