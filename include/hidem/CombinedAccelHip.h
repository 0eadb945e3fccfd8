// CombinedAccelHip.h -- drop-in accelerator for HiDEMLib on MI355X (gfx950).
//
// Header-only C++ over the C ABI of hdm_abi.h.  It offers the method set
// HiDEMLib calls on its accelerator object (HiDEMLib.h:54-60 selects
// CombinedAccel / CombinedAccelSmem; calls at :265, :332, :347, :603-622,
// :815-818), plus the PLiB shapes the hot path exposes: Real3, Domain
// (PLiB/Domain.h:15-80), World (PLiB/World.h:18-43, a single domain owned by
// rank 0) and NeighbourGrid (PLiB/NeighbourGrid.h:23-107, read-back view).
//
// The reference passes arbitrary device lambdas into iterateCombined*/
// iterateLinks.  A C ABI cannot carry them, so this accelerator accepts the
// default functors of HiDEMDefaultFunctions.h (HiDEMDefNeighForceFunc,
// HiDEMDefLinkForceFunc, HiDEMDefLinkBreakFunc) as parameter structs with the
// same names and constructor argument order; any other functor type fails to
// compile (there is no generic fallback).
//
// Failures follow the reference's convention (DevCall.h:69-91): print the
// message and exit(1).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "hdm_abi.h"

namespace hidem {

inline void hdm_check(hdm_ctx *ctx, int rc, const char *what) {
    if (rc == HDM_OK) return;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, ctx ? hdm_last_error(ctx) : "no context");
    std::exit(1);
}

// ---- PLiB shapes -------------------------------------------------------------
struct Real3 {
    double x = 0, y = 0, z = 0;
    Real3() = default;
    Real3(double a, double b, double c) : x(a), y(b), z(c) {}
    explicit Real3(double a) : x(a), y(a), z(a) {}
    Real3 operator+(const Real3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    Real3 operator-(const Real3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    Real3 operator+(double s) const { return {x + s, y + s, z + s}; }
    Real3 operator-(double s) const { return {x - s, y - s, z - s}; }
};

// PLiB/Domain.h:15-80.
struct Domain {
    Real3 upper, lower, upperHalo, lowerHalo;
    int owner = 0;
    double load = 0;
    bool within(Real3 c) const {
        if (c.x < lower.x || c.x >= upper.x) return false;
        if (c.y < lower.y || c.y >= upper.y) return false;
        if (c.z < lower.z || c.z >= upper.z) return false;
        return true;
    }
    bool withinHalo(Real3 c) const {
        if (c.x < lowerHalo.x || c.x >= upperHalo.x) return false;
        if (c.y < lowerHalo.y || c.y >= upperHalo.y) return false;
        if (c.z < lowerHalo.z || c.z >= upperHalo.z) return false;
        return true;
    }
    Real3 getSize() const { return upper - lower; }
    void setHalo(double h) {
        upperHalo = upper + h;
        lowerHalo = lower - h;
    }
};

// PLiB/World.h:18-43 on one GPU: one rank, one domain.
struct World {
    std::vector<Domain> domains;
    int self = 0;
    int size = 1;
    Domain getOwnDomain() const {
        for (auto &d : domains)
            if (d.owner == self) return d;
        Domain d;
        d.owner = -1;
        return d;
    }
};

// ---- default functors (HiDEMDefaultFunctions.h) as parameter structs --------
struct HiDEMDefNeighForceFunc {
    double invDt, attractDistance, attractScaleForce, damping, dampingRotation, scaleE;
    HiDEMDefNeighForceFunc(double invDt_, double attractDistance_, double attractScaleForce_, double damping_,
                           double dampingRotation_, double scaleE_)
        : invDt(invDt_), attractDistance(attractDistance_), attractScaleForce(attractScaleForce_),
          damping(damping_), dampingRotation(dampingRotation_), scaleE(scaleE_) {}
};

struct HiDEMDefLinkForceFunc {
    double invDt, damping, dampingRotation, E, W2, W4, G;
    HiDEMDefLinkForceFunc(double invDt_, double damping_, double dampingRotation_, double E_, double W2_,
                          double W4_, double G_)
        : invDt(invDt_), damping(damping_), dampingRotation(dampingRotation_), E(E_), W2(W2_), W4(W4_), G(G_) {}
};

struct HiDEMDefLinkBreakFunc {
    double bendBreak, maxLoad;
    HiDEMDefLinkBreakFunc(double bendBreak_, double maxLoad_) : bendBreak(bendBreak_), maxLoad(maxLoad_) {}
};

inline hdm_force_params make_force_params(const HiDEMDefNeighForceFunc &n, const HiDEMDefLinkForceFunc &l) {
    hdm_force_params f;
    f.invDt = n.invDt;
    f.attractDistance = n.attractDistance;
    f.attractScaleForce = n.attractScaleForce;
    f.damping = n.damping;
    f.dampingRotation = n.dampingRotation;
    f.scaleE = n.scaleE;
    f.E = l.E;
    f.W2 = l.W2;
    f.W4 = l.W4;
    f.G = l.G;
    return f;
}

// ---- NeighbourGrid read-back view (PLiB/NeighbourGrid.h:23-107) --------------
struct NeighbourGrid {
    hdm_ctx *ctx = nullptr;
    int xGridSize = 0, yGridSize = 0, zGridSize = 0;
    double xMin = 0, yMin = 0, zMin = 0, xMax = 0, yMax = 0, zMax = 0;
    int64_t totalCells = 0;
    double gridSpacing = 0;
    std::vector<int64_t> colours;
    std::vector<int32_t> iterationOrder;

    int64_t getGridColor(int x, int y, int z) const {
        return (int64_t)x + (int64_t)y * (int64_t)xGridSize + (int64_t)z * (int64_t)xGridSize * (int64_t)yGridSize;
    }
    std::tuple<double, double, double> getGridId(double x, double y, double z) const {
        return std::make_tuple((int)std::floor((x - xMin) / gridSpacing), (int)std::floor((y - yMin) / gridSpacing),
                               (int)std::floor((z - zMin) / gridSpacing));
    }
    // Copies the device grid (dimensions, colours, iterationOrder) to the host.
    void refresh(int64_t nParticles) {
        hdm_grid_info g;
        hdm_check(ctx, hdm_grid_info_get(ctx, &g), "hdm_grid_info_get");
        xGridSize = g.X; yGridSize = g.Y; zGridSize = g.Z;
        xMin = g.xMin; yMin = g.yMin; zMin = g.zMin;
        xMax = g.xMax; yMax = g.yMax; zMax = g.zMax;
        totalCells = g.totalCells;
        gridSpacing = g.spacing;
        colours.resize((size_t)nParticles);
        iterationOrder.resize((size_t)nParticles);
        hdm_check(ctx, hdm_grid_download(ctx, colours.data(), iterationOrder.data()), "hdm_grid_download");
    }
    // NeighbourGrid::toVTKHDF (a debugging dump of the helper structures,
    // HiDEMLib.h:605-607 with output.writeHelperStructures) is not part of
    // this accelerator: asking for it ends the run the reference's way.
    void toVTKHDF(const std::string &path) const {
        std::fprintf(stderr, "NeighbourGrid::toVTKHDF(%s): helper-structure output is not supported on this "
                             "accelerator (unset output.writeHelperStructures)\n", path.c_str());
        std::exit(1);
    }
};

// ---- the link accelerator's members HiDEMLib touches ---------------------------
// LinkAccelSide (PLiB/Accel/LinkAccelSide.h): keepInactive (HiDEMLib.h:814)
// and sortSideData (HiDEMLib.h:621).  Broken links always stay in the lists
// here (their CT keeps integrating, as the reference's does), so
// keepInactive changes nothing.  The link set does not change between
// uploads on this path, so the lists and the side permutation built once
// stay valid across neighbour rebuilds: `current` says they are, and the
// reference-shaped calls that the reference makes at every rebuild
// (generateLinks + sortSideData, HiDEMLib.h:620-621) build only when not.
struct LinkAccelHip {
    hdm_ctx *ctx = nullptr;
    bool keepInactive = false;
    bool current = false;
    // LinkAccelSide::sortSideData (LinkAccelSide.h:159-417): the CT side
    // store is permuted when the link lists are built (hdm_links_build).
    void sortSideData() {
        hdm_check(ctx, hdm_links_build(ctx), "hdm_links_build");
        current = true;
    }
    template <class PData, class LData>
    void sortSideData(PData &, LData &) {
        if (!current) sortSideData();
    }
};

namespace detail {
template <class F, class Want>
constexpr bool is_default_functor = std::is_same<typename std::decay<F>::type, Want>::value;
}

// ---- the accelerator ------------------------------------------------------------
class CombinedAccelHip {
public:
    explicit CombinedAccelHip(int32_t precision = HDM_PREC_FP64, int32_t device = 0) {
        hdm_check(nullptr, hdm_ctx_create(precision, device, &ctx), "hdm_ctx_create");
        grid.ctx = ctx;
        lAccel.ctx = ctx;
    }
    ~CombinedAccelHip() { hdm_ctx_destroy(ctx); }
    CombinedAccelHip(const CombinedAccelHip &) = delete;
    CombinedAccelHip &operator=(const CombinedAccelHip &) = delete;

    hdm_ctx *ctx = nullptr;
    NeighbourGrid grid;
    LinkAccelHip lAccel;

    // Binding of the data handlers' fields (ParticleDataHandler / LinkDataHandler).
    void uploadParticles(const hdm_particles_host &p) { hdm_check(ctx, hdm_particles_upload(ctx, &p), "hdm_particles_upload"); }
    void uploadLinks(const hdm_links_host &l) {
        hdm_check(ctx, hdm_links_upload(ctx, &l), "hdm_links_upload");
        lAccel.current = false;
    }
    void downloadParticles(hdm_particles_host &p) { hdm_check(ctx, hdm_particles_download(ctx, &p), "hdm_particles_download"); }
    void downloadLinks(hdm_links_host &l) { hdm_check(ctx, hdm_links_download(ctx, &l), "hdm_links_download"); }

    // CombinedAccel::generatePrePostCommIndexes (CombinedAccel.h:53-96): one
    // GPU owns every particle, there is no pre/post-communication split.
    template <class... T> void generatePrePostCommIndexes(T &&...) {}

    // CombinedAccel::generateNeighbours (CombinedAccel.h:32-41).
    void generateNeighbours(double searchDistance, double interactDist) {
        hdm_check(ctx, hdm_grid_build(ctx, searchDistance), "hdm_grid_build");
        hdm_check(ctx, hdm_neighbours_build(ctx, interactDist), "hdm_neighbours_build");
    }
    // CombinedAccel::generateLinks (CombinedAccel.h:43-48) and
    // LinkAccelSide::sortSideData (LinkAccelSide.h:159-417): one call builds
    // the per-particle link lists and permutes the CT side store.
    void generateLinks() { lAccel.sortSideData(); }
    void sortSideData() { lAccel.sortSideData(); }
    // CombinedAccelSmem::generateIndexes (CombinedAccelSmem.h:345-524): the
    // wave-tiled lists need no further index generation.
    void generateIndexes() {}

    // CombinedAccel::iterateCombinedPre/Post (CombinedAccel.h:98-119).
    void iterateCombinedPre(const HiDEMDefNeighForceFunc &nFunc, const HiDEMDefLinkForceFunc &lFunc) {
        hdm_force_params f = make_force_params(nFunc, lFunc);
        hdm_check(ctx, hdm_force(ctx, &f), "hdm_force");
    }
    void iterateCombinedPost(const HiDEMDefNeighForceFunc &, const HiDEMDefLinkForceFunc &) {}

    // CombinedAccel::iterateLinks (CombinedAccel.h:200-235) with the default
    // break functor; returns the number of links broken by this call.
    int64_t iterateLinks(const HiDEMDefLinkBreakFunc &bFunc, bool countBroken = false) {
        hdm_break_params b{bFunc.bendBreak, bFunc.maxLoad};
        int64_t nb = 0;
        hdm_check(ctx, hdm_break(ctx, &b, countBroken ? &nb : nullptr), "hdm_break");
        return nb;
    }

    // ---- the reference's call expressions -----------------------------------
    // HiDEMLib calls its accelerator with the data handlers, an execution
    // policy and the packed ManVec argument tuples (HiDEMLib.h:265, 332,
    // 347, 603-622, 814-818).  The context holds device copies of every
    // field those packs name (uploadParticles / uploadLinks bind them once,
    // downloadParticles / downloadLinks return them), so these overloads
    // take the reference's argument lists and pass on only what the native
    // kernels need: the distances and the functors' parameters.  Only the
    // default functors compile; any other functor type is a compile error.
    template <class PData, class LData>
    void generateNeighbours(PData &, LData &, double searchDistance, double interactDist) {
        generateNeighbours(searchDistance, interactDist);
    }
    template <class PData, class LData>
    void generateLinks(PData &, LData &, double, double) {
        if (!lAccel.current) generateLinks();
    }
    template <class PData, class LData>
    void generateIndexes(PData &, LData &, double, double) {}

    template <class Policy, class PData, class LData, class NFunc, class LFunc, class Pack>
    void iterateCombinedPre(Policy, PData &, LData &, const NFunc &nFunc, const LFunc &lFunc, const Pack &) {
        static_assert(detail::is_default_functor<NFunc, HiDEMDefNeighForceFunc>,
                      "CombinedAccelHip runs the default neighbour force functor only (HiDEMDefNeighForceFunc)");
        static_assert(detail::is_default_functor<LFunc, HiDEMDefLinkForceFunc>,
                      "CombinedAccelHip runs the default link force functor only (HiDEMDefLinkForceFunc)");
        iterateCombinedPre(nFunc, lFunc);
    }
    template <class Policy, class PData, class LData, class NFunc, class LFunc, class Pack>
    void iterateCombinedPost(Policy, PData &, LData &, const NFunc &nFunc, const LFunc &lFunc, const Pack &) {
        static_assert(detail::is_default_functor<NFunc, HiDEMDefNeighForceFunc> &&
                          detail::is_default_functor<LFunc, HiDEMDefLinkForceFunc>,
                      "CombinedAccelHip runs the default force functors only");
        iterateCombinedPost(nFunc, lFunc);
    }
    template <class Policy, class PData, class LData, class BFunc, class Pack>
    void iterateLinks(Policy, PData &, LData &, const BFunc &bFunc, const Pack &) {
        static_assert(detail::is_default_functor<BFunc, HiDEMDefLinkBreakFunc>,
                      "CombinedAccelHip runs the default link break functor only (HiDEMDefLinkBreakFunc)");
        iterateLinks(bFunc);
    }

    // applyForce (HiDEMLib.h:303-314) with a native integrator.
    void applyForce(const hdm_integ_params &ip) { hdm_check(ctx, hdm_integrate(ctx, &ip), "hdm_integrate"); }

    void sync() { hdm_check(ctx, hdm_sync(ctx), "hdm_sync"); }

    // The redistribution sort (HiDEMDistribution.h:150-181) on this rank's
    // data; order / linkOrder (may be null) receive the old index of each new
    // particle / link, in place of the reference's global-key remapping.
    void redistribute(double searchDistance, int64_t *order = nullptr, int64_t *linkOrder = nullptr) {
        hdm_check(ctx, hdm_reorder(ctx, searchDistance, order, linkOrder), "hdm_reorder");
    }

    // GlacierDem map inputs (InputBed's bed and friction, ExternalForces):
    // HDM_MAP_COUNT slots, null for the analytic input of hdm_integ_params.
    void setMaps(const hdm_map2d *const maps[HDM_MAP_COUNT]) {
        hdm_check(ctx, hdm_set_maps(ctx, maps), "hdm_set_maps");
    }
};

}  // namespace hidem
