// Per-timestep kernels for gfx950: force gather, time integration, link
// break and the rebuild (skin) reductions.
//
// Force gather (CombinedAccel::iterateCombined, CombinedAccel.h:121-197):
// one lane per particle, 64 consecutive particles per wavefront.  A lane
// zeroes its force, applies the contact law to every neighbour-list entry in
// list order, then the beam law to each of its link sides in ascending link
// id, updating that side's CT0..CT5 in place, and finally overwrites its
// force.  Each particle owns its sides, so there are no atomics and the
// summation order is fixed: results are deterministic and, in the FP64 form,
// bit-identical to the sequential CPU order (this translation unit is built
// with -ffp-contract=off; FP64 sqrt and division are correctly rounded).
//
// The neighbour list and the link sides live in wave tiles: slot k of the 64
// particles of a tile is one contiguous run, so every list/CT access of a
// wavefront is a single coalesced 256 B / 512 B transaction.  Partner state
// (coordinates, Diff, radius) is gathered through L2; particles are stored in
// spatial (Morton) order and blocks are renumbered per XCD, so partners of a
// tile were usually fetched by a neighbouring tile on the same XCD.  A link's
// broken flag rides in the sign bit of its slot words (no per-link gather),
// and contact pairs beyond the attraction range are skipped when that provably
// leaves the force unchanged (contact_is_inert).
#include <cmath>
#include <type_traits>

#include "hdm_context.h"
#include "kernel_common.h"

namespace hdm {

namespace {

constexpr int TPB = 256;

template <class R>
struct ForceParams {
    R invDt, attractDistance, attractScaleForce, damping, dampingRotation, scaleE;
    R E, W2, W4, G;
    int32_t skipInert;  // contact constants small enough for contact_is_inert's argument
};

// The inert-contact skip relies on every term of the contact law staying
// finite in R for radius sums up to 1e6; these bounds on the law's constants
// guarantee it (Cnst <= sqrt(1e18) * 1e15 or 1e12 * 1e6 * 1e15 < FLT_MAX).
inline bool inert_skip_allowed(const hdm_force_params &f) {
    return std::fabs(f.attractDistance) <= 1e6 && std::fabs(f.attractScaleForce) <= 1e12 &&
           std::fabs(f.scaleE) <= 1e15;
}

template <class R>
ForceParams<R> cast_params(const hdm_force_params &f) {
    return ForceParams<R>{(R)f.invDt, (R)f.attractDistance, (R)f.attractScaleForce, (R)f.damping,
                          (R)f.dampingRotation, (R)f.scaleE, (R)f.E, (R)f.W2, (R)f.W4, (R)f.G,
                          inert_skip_allowed(f) ? 1 : 0};
}

// Diff differences of a pair, s - o: in the FP64 form exactly the law's own
// (self[c] - other[c]).  In the FP32 form the partner's Diff is the FP32 value
// its gather record carries, and the difference with the particle's own FP64
// Diff is formed in FP64 and rounded once: neighbouring Diffs are nearly
// equal, and rounding both to FP32 before subtracting would cancel most of the
// significant bits of their difference.
template <class R>
__device__ __forceinline__ void diff_delta(const double s[6], const double o[6], R d[6]) {
#pragma unroll
    for (int c = 0; c < 6; c++) d[c] = (R)(s[c] - (sizeof(R) == 4 ? (double)(float)o[c] : o[c]));
}

// HiDEMDefNeighForceFunc::operator() (HiDEMDefaultFunctions.h:38-86); the
// damping terms take the Diff differences d (diff_delta).
template <class R>
__device__ __forceinline__ void contact_law(R mask, R xDist, R yDist, R zDist, const R self[7], const R other[7],
                                            const R d[6], const ForceParams<R> &fp, R F[6]) {
    R xDist2 = xDist * xDist;
    R yDist2 = yDist * yDist;
    R zDist2 = zDist * zDist;
    R particleDistance2 = xDist2 + yDist2 + zDist2;
    R radSum = self[6] + other[6];
    R activeDist2 = (radSum + fp.attractDistance) * (radSum + fp.attractDistance);
    mask = (particleDistance2 >= activeDist2) ? R(0) : mask;
    R particleDistance = sqrt(particleDistance2);
    R invDist = R(1) / particleDistance;
    R xComp = xDist * invDist;
    R yComp = yDist * invDist;
    R zComp = zDist * invDist;
    R Cnst = 0;
    R dimDist = radSum - particleDistance;
    Cnst = (particleDistance < radSum) ? R(sqrt(dimDist * dimDist * dimDist)) : Cnst;
    Cnst = (particleDistance >= radSum && particleDistance < radSum + fp.attractDistance)
               ? fp.attractScaleForce * dimDist : Cnst;
    Cnst *= fp.scaleE;
    F[0] += mask * ((-fp.damping) * (d[0] * fp.invDt) + (Cnst * xComp));
    F[1] += mask * ((-fp.damping) * (d[1] * fp.invDt) + (Cnst * yComp));
    F[2] += mask * ((-fp.damping) * (d[2] * fp.invDt) + (Cnst * zComp));
    F[3] += mask * ((-fp.dampingRotation) * (d[3] * fp.invDt));
    F[4] += mask * ((-fp.dampingRotation) * (d[4] * fp.invDt));
    F[5] += mask * ((-fp.dampingRotation) * (d[5] * fp.invDt));
}

template <class R> __device__ __forceinline__ R huge_value();
template <> __device__ __forceinline__ double huge_value<double>() { return 1e300; }
template <> __device__ __forceinline__ float huge_value<float>() { return 1e37f; }

// True when contact_law would add exactly +-0 to every force component, so the
// pair can be skipped without changing a bit of the result: beyond
// radSum + attractDistance the law multiplies its terms by mask = 0, and while
// those terms are finite the product is +-0.  Adding +-0 leaves F unchanged,
// because F starts at +0 and round-to-nearest sums never yield -0.  The terms
// are finite when the distance is positive and at most sqrt(huge), the radius
// sum at most 1e6, the constants within inert_skip_allowed and the damping
// terms at most huge (recomputed here in the law's own expression order).
// Other pairs take the full path, so NaN/inf propagate exactly as in the
// sequential order of CombinedAccel::iterateCombined.
template <class R>
__device__ __forceinline__ bool contact_is_inert(R xDist, R yDist, R zDist, const R self[7], const R other[7],
                                                 const R d[6], const ForceParams<R> &fp) {
    R particleDistance2 = xDist * xDist + yDist * yDist + zDist * zDist;
    R radSum = self[6] + other[6];
    R activeDist2 = (radSum + fp.attractDistance) * (radSum + fp.attractDistance);
    if (!fp.skipInert || !(particleDistance2 >= activeDist2) || !(particleDistance2 <= huge_value<R>()) ||
        !(particleDistance2 > R(0)) || !(fabs(radSum) <= R(1e6)))
        return false;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 3; c++) ok &= fabs((-fp.damping) * (d[c] * fp.invDt)) <= huge_value<R>();
#pragma unroll
    for (int c = 3; c < 6; c++) ok &= fabs((-fp.dampingRotation) * (d[c] * fp.invDt)) <= huge_value<R>();
    return ok;
}

// HiDEMDefLinkForceFunc::operator() (HiDEMDefaultFunctions.h:105-203).  The
// FP64 form evaluates the reference expressions in their order.  The FP32
// form groups the same terms by pair differences, B0 - B6 = T0 . (s - o) etc.,
// with s - o = d from diff_delta, so the CT increments keep the relative
// motion of the two ends instead of the difference of two rounded values.
template <class R>
__device__ __forceinline__ void beam_law(R mask, R xDist, R yDist, R zDist, const R s[7], const R o[7],
                                         const R d[6], const ForceParams<R> &fp, R ct[6], R F[6]) {
    F[0] += (-fp.damping * (d[0] * fp.invDt)) * mask;
    F[1] += (-fp.damping * (d[1] * fp.invDt)) * mask;
    F[2] += (-fp.damping * (d[2] * fp.invDt)) * mask;
    F[3] += (-fp.dampingRotation * (d[3] * fp.invDt)) * mask;
    F[4] += (-fp.dampingRotation * (d[4] * fp.invDt)) * mask;
    F[5] += (-fp.dampingRotation * (d[5] * fp.invDt)) * mask;

    R xDist2 = xDist * xDist;
    R yDist2 = yDist * yDist;
    if (xDist == R(0) && yDist == R(0)) {
        xDist2 = R(1e-10);
        yDist2 = R(1e-10);
    }
    R zDist2 = zDist * zDist;
    R distSum = xDist2 + yDist2 + zDist2;
    R invLI = R(1) / sqrt(distSum);
    R invLP = R(1) / sqrt(xDist2 + yDist2);
    R invLILP = invLI * invLP;
    R LI = invLI * distSum;

    R T00 = xDist * invLI, T01 = yDist * invLI, T02 = zDist * invLI;
    R T10 = -yDist * invLP, T11 = xDist * invLP;
    R T20 = -(zDist * xDist) * invLILP;
    R T21 = -(zDist * yDist) * invLILP;
    R T22 = (xDist2 + yDist2) * invLILP;

    R L2 = LI * LI;
    R L3 = LI * LI * LI;
    R LIDiv6 = (R(1) / R(6)) * invLI;

    R K0 = fp.E * fp.W2 * invLI;
    R K1 = fp.E * fp.W4 / L3;
    R K5 = fp.E * fp.W4 / (R(2) * L2);
    R K2 = fp.G * fp.W4 * LIDiv6;
    R K3 = fp.E * fp.W4 * LIDiv6 * R(2);
    R K4 = fp.E * fp.W4 * LIDiv6;

    R B4 = T10 * s[3] + T11 * s[4];
    R B5 = T20 * s[3] + T21 * s[4] + T22 * s[5];
    R B10 = T10 * o[3] + T11 * o[4];
    R B11 = T20 * o[3] + T21 * o[4] + T22 * o[5];

    if constexpr (sizeof(R) == 8) {
        R B0 = T00 * s[0] + T01 * s[1] + T02 * s[2];
        R B1 = T10 * s[0] + T11 * s[1];
        R B2 = T20 * s[0] + T21 * s[1] + T22 * s[2];
        R B6 = T00 * o[0] + T01 * o[1] + T02 * o[2];
        R B7 = T10 * o[0] + T11 * o[1];
        R B8 = T20 * o[0] + T21 * o[1] + T22 * o[2];

        ct[0] += K0 * B0 + (-K0) * B6;
        ct[1] += K1 * B1 + K5 * B5 + (-K1) * B7 + K5 * B11;
        ct[2] += K1 * B2 + (-K5) * B4 + (-K1) * B8 + (-K5) * B10;

        F[0] += (-(T00 * ct[0] + T10 * ct[1] + T20 * ct[2])) * mask;
        F[1] += (-(T01 * ct[0] + T11 * ct[1] + T21 * ct[2])) * mask;
        F[2] += (-(T02 * ct[0] + T22 * ct[2])) * mask;

        R B3 = T00 * s[3] + T01 * s[4] + T02 * s[5];
        R B9 = T00 * o[3] + T01 * o[4] + T02 * o[5];

        ct[3] += K2 * B3 + (-K2) * B9;
        ct[4] += (-K5) * B2 + K3 * B4 + K5 * B8 + K4 * B10;
        ct[5] += K5 * B1 + K3 * B5 + (-K5) * B7 + K4 * B11;
    } else {
        R D0 = T00 * d[0] + T01 * d[1] + T02 * d[2];   // B0 - B6
        R D1 = T10 * d[0] + T11 * d[1];                 // B1 - B7
        R D2 = T20 * d[0] + T21 * d[1] + T22 * d[2];   // B2 - B8
        R D3 = T00 * d[3] + T01 * d[4] + T02 * d[5];   // B3 - B9

        ct[0] += K0 * D0;
        ct[1] += K1 * D1 + K5 * (B5 + B11);
        ct[2] += K1 * D2 + (-K5) * (B4 + B10);

        F[0] += (-(T00 * ct[0] + T10 * ct[1] + T20 * ct[2])) * mask;
        F[1] += (-(T01 * ct[0] + T11 * ct[1] + T21 * ct[2])) * mask;
        F[2] += (-(T02 * ct[0] + T22 * ct[2])) * mask;

        ct[3] += K2 * D3;
        ct[4] += (-K5) * D2 + K3 * B4 + K4 * B10;
        ct[5] += K5 * D1 + K3 * B5 + K4 * B11;
    }

    F[3] += (-(T00 * ct[3] + T10 * ct[4] + T20 * ct[5])) * mask;
    F[4] += (-(T01 * ct[3] + T11 * ct[4] + T21 * ct[5])) * mask;
    F[5] += (-(T02 * ct[3] + T22 * ct[5])) * mask;
}

template <class R, class CT, class FO>
struct ForceArgs {
    using CTType = CT;
    const double *x, *y, *z, *d0, *d1, *d2, *d3, *d4, *d5, *radius;
    int64_t n;
    const int32_t *nbrCount, *nbr;
    int32_t nbrWidth;
    const uint32_t *linkCount;
    const int32_t *slotLink, *slotOther;
    int32_t linkWidth;
    CT *ct;  // side store (ct_offset layout)
    FO *f0, *f1, *f2, *f3, *f4, *f5;
    ForceParams<R> fp;
    // Fused link break of the previous timestep (fused step only; breakCheck
    // = 0 elsewhere): before its beam law, every intact side evaluates
    // HiDEMDefLinkBreakFunc on the current state, exactly as k_break does
    // after the previous integration.
    int32_t breakCheck;
    // The neighbour list leaves out link partners (Context::nbrExcludesLinked,
    // FP32 only): the link loop applies the contact law of each side's pair,
    // before its beam law.
    int32_t linkedContacts;
    int32_t *slotLinkW, *slotOtherW;
    const double *tx, *ty, *tz, *initialLen, *maxLoad;
    uint8_t *state;
    double bendBreak, scalarMaxLoad;
    uint64_t *brokenCount;  // fused break: links broken are added here (nullptr: not counted)
};

// Force on particle p (CombinedAccel::iterateCombined body): contacts in list
// order, then link sides in ascending link id, CT updated in place.
// HiDEMDefLinkBreakFunc (HiDEMDefaultFunctions.h:217-241) for link l seen
// from one end: d = this end's coordinates minus the other's, t = this end's
// rotations, q = the other end.  Both ends compute the same bits: IEEE
// subtraction is antisymmetric (a - b == -(b - a)), so the squared distances
// and the absolute rotation differences do not depend on which end asks.
template <class R, class CT, class FO>
__device__ __forceinline__ bool link_breaks(const ForceArgs<R, CT, FO> &a, int32_t l, int32_t q, double xDist,
                                            double yDist, double zDist, const double t[3]) {
    double stretchLength = sqrt((xDist * xDist) + (yDist * yDist) + (zDist * zDist));
    double bx = fabs(t[0] - a.tx[q]);
    double by = fabs(t[1] - a.ty[q]);
    double bz = fabs(t[2] - a.tz[q]);
    double inner = (by < bz) ? bz : by;
    double maxBend = (bx < inner) ? inner : bx;
    double initialLen = a.initialLen[l];
    double load = (stretchLength - initialLen) / initialLen + a.bendBreak * maxBend;
    double maxLoad = a.maxLoad ? a.maxLoad[l] : a.scalarMaxLoad;
    return load >= maxLoad;
}

template <class R, class CT, class FO>
__device__ __forceinline__ void force_particle(const ForceArgs<R, CT, FO> &a, int64_t p, R F[6]) {
    const double xCord = a.x[p], yCord = a.y[p], zCord = a.z[p];
    const double sd[6] = {a.d0[p], a.d1[p], a.d2[p], a.d3[p], a.d4[p], a.d5[p]};
    R self[7] = {(R)sd[0], (R)sd[1], (R)sd[2], (R)sd[3], (R)sd[4], (R)sd[5], (R)a.radius[p]};
#pragma unroll
    for (int c = 0; c < 6; c++) F[c] = 0;
    const int64_t lane = p & 63;
    {
        const int32_t cnt = a.nbrCount[p];
        const int32_t *nl = a.nbr + (p >> 6) * (int64_t)64 * a.nbrWidth + lane;
        for (int32_t k = 0; k < cnt; k++) {
            const int32_t q = nl[(int64_t)k * 64];
            const double od[6] = {a.d0[q], a.d1[q], a.d2[q], a.d3[q], a.d4[q], a.d5[q]};
            R other[7] = {(R)od[0], (R)od[1], (R)od[2], (R)od[3], (R)od[4], (R)od[5], (R)a.radius[q]};
            R d[6];
            diff_delta<R>(sd, od, d);
            R xDist = (R)(xCord - a.x[q]);
            R yDist = (R)(yCord - a.y[q]);
            R zDist = (R)(zCord - a.z[q]);
            if (contact_is_inert<R>(xDist, yDist, zDist, self, other, d, a.fp)) continue;
            contact_law<R>(R(1), xDist, yDist, zDist, self, other, d, a.fp, F);
        }
    }
    {
        const int32_t cnt = (int32_t)a.linkCount[p];
        const int64_t base = (p >> 6) * (int64_t)64 * a.linkWidth + lane;
        for (int32_t k = 0; k < cnt; k++) {
            const int64_t sidx = base + (int64_t)k * 64;
            const int32_t ow = a.slotOther[sidx];
            const int32_t q = ow & SLOT_LINK_MASK;
            const double od[6] = {a.d0[q], a.d1[q], a.d2[q], a.d3[q], a.d4[q], a.d5[q]};
            R other[7] = {(R)od[0], (R)od[1], (R)od[2], (R)od[3], (R)od[4], (R)od[5], R(0)};
            R d[6];
            diff_delta<R>(sd, od, d);
            R xDist = (R)(xCord - a.x[q]);
            R yDist = (R)(yCord - a.y[q]);
            R zDist = (R)(zCord - a.z[q]);
            R mask = 1;
            if (ow < 0) mask = 0;  // broken link (state == 0)
            CT cs[6];
            ct_load(a.ct, sidx, cs);
            R ct[6] = {(R)cs[0], (R)cs[1], (R)cs[2], (R)cs[3], (R)cs[4], (R)cs[5]};
            beam_law<R>(mask, xDist, yDist, zDist, self, other, d, a.fp, ct, F);
#pragma unroll
            for (int c = 0; c < 6; c++) cs[c] = (CT)ct[c];
            ct_store(a.ct, sidx, cs);
        }
        // the contacts of linked pairs (list without link partners), after
        // every beam, in side order (as the fused step)
        if (sizeof(R) == 4 && a.linkedContacts) {
            for (int32_t k = 0; k < cnt; k++) {
                const int32_t q = a.slotOther[base + (int64_t)k * 64] & SLOT_LINK_MASK;
                const double od[6] = {a.d0[q], a.d1[q], a.d2[q], a.d3[q], a.d4[q], a.d5[q]};
                R other[7] = {(R)od[0], (R)od[1], (R)od[2], (R)od[3], (R)od[4], (R)od[5], (R)a.radius[q]};
                R d[6];
                diff_delta<R>(sd, od, d);
                R xDist = (R)(xCord - a.x[q]);
                R yDist = (R)(yCord - a.y[q]);
                R zDist = (R)(zCord - a.z[q]);
                if (!contact_is_inert<R>(xDist, yDist, zDist, self, other, d, a.fp))
                    contact_law<R>(R(1), xDist, yDist, zDist, self, other, d, a.fp, F);
            }
        }
    }
}

template <class R, class CT, class FO>
__device__ __forceinline__ void store_force(const ForceArgs<R, CT, FO> &a, int64_t p, const R F[6]) {
    a.f0[p] = (FO)F[0];
    a.f1[p] = (FO)F[1];
    a.f2[p] = (FO)F[2];
    a.f3[p] = (FO)F[3];
    a.f4[p] = (FO)F[4];
    a.f5[p] = (FO)F[5];
}

template <class R, class CT, class FO>
__global__ void __launch_bounds__(TPB) k_force(ForceArgs<R, CT, FO> a) {
    const int64_t p = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    if (p >= a.n) return;
    R F[6];
    force_particle(a, p, F);
    store_force(a, p, F);
}

// ---- time integration --------------------------------------------------------
// Particle state read by the integrator (in) and written by it (out).  The
// standalone integrator updates in place (out == in); the fused step kernel
// writes the other buffer of a ping-pong pair, because other lanes of the same
// launch still gather the old state for their forces.
struct IntegArgs {
    const double *x, *y, *z, *tx, *ty, *tz, *d0, *d1, *d2, *d3, *d4, *d5;
    double *ox, *oy, *oz, *otx, *oty, *otz, *od0, *od1, *od2, *od3, *od4, *od5;
    const double *radius;
    const uint8_t *type;
    int64_t n;
    hdm_integ_params ip;
    // rebuild-test reductions of the new state (nullptr: none): one row of 4
    // doubles per workgroup, max(0, max Diff) per axis (HiDEMLib.h:424-461)
    // and max |x - x0|^2, folded by k_skin_finalize
    double *partials;
    // exact-test guard (nullptr: off; see StepGuard): a launch that finds
    // *guard (the trip word of its input records) set does nothing but set
    // *guardOut; otherwise a particle whose new position trips the exact test
    // sets *guardOut, and workgroup 0 counts the executed launch in
    // *stepsDone.  k_skin_finalize evaluates the same test for the
    // host-checked policies.
    const uint32_t *guard;
    uint32_t *guardOut, *stepsDone;
    double guardThr;
    const double *x0, *y0, *z0;   // nullptr: no displacement reduction
    const float *x0f, *y0f, *z0f; // the snapshot rounded to FP32 (guard_trips)
    const int32_t *blockId;       // batched form: waterline of the particle's block
    const double *blockWater;     // (nullptr: ip.waterLevel for every particle)
    const MapDev *maps;           // GlacierDem map inputs, slot k valid when bit k of mapMask
    int32_t mapMask;
};

// Damped central difference (HiDEMDefTimeIntegration, HiDEMDefaultFunctions.h
// :317-370; same expressions as the inline IceBlock form).
__device__ __forceinline__ void central_difference(double dragCoeff, double mass, double inertia,
                                                   const hdm_integ_params &ip, const double F[6],
                                                   const double diff[6], double cord[3], double cordPrev[3],
                                                   double theta[3], double thetaPrev[3], double vel[6]) {
    double dragInvDt2 = dragCoeff * ip.invDt2;
    double massDtSqr = mass * ip.invDtSqr;
    double pMassDtSqr = mass / ip.dtSqr;
    double pMemDtSqr = inertia / ip.dtSqr;
    double invC1 = 1.0 / (pMassDtSqr + dragCoeff * ip.invDt2);
    double invC2 = 1.0 / (pMemDtSqr + dragCoeff * ip.invDt2);
    double interDtSqr = inertia * ip.invDtSqr;
    double next[6];
#pragma unroll
    for (int c = 0; c < 3; c++) next[c] = (F[c] + massDtSqr * diff[c] + dragInvDt2 * -diff[c]) * invC1;
#pragma unroll
    for (int c = 3; c < 6; c++) next[c] = (F[c] + interDtSqr * diff[c] + dragInvDt2 * -diff[c]) * invC2;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cordPrev[c] = cord[c];
        thetaPrev[c] = theta[c];
        cord[c] = cord[c] + next[c];
        theta[c] = theta[c] + next[3 + c];
        vel[c] = next[c] * ip.invDt;
        vel[3 + c] = (theta[c] - thetaPrev[c]) * ip.invDt;
    }
}

// Map2d::getIndexes (Map2d.h:92-108) for one axis: floor of the scaled
// coordinate, held within [-1, n] before the int conversion (so NaN and huge
// coordinates stay defined; the clamp to [0, n - 1] below treats them as the
// reference does any value beyond the grid).
__device__ __forceinline__ int32_t map_floor(float f, int32_t n) {
    f = floorf(f);
    if (!(f >= -1.0f)) f = -1.0f;
    if (f > (float)n) f = (float)n;
    return (int32_t)f;
}

__device__ __forceinline__ int32_t map_clamp(int32_t v, int32_t lo, int32_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Map2d::interp (Map2d.h:110-145): FP32 bilinear interpolation in the
// reference's expression order; the grid value itself where both axes clamp
// to one point.
__device__ double map_interp(const MapDev &m, float x, float y) {
    x += m.xOffset;
    y += m.yOffset;
    const int32_t xPoint = map_floor(((x - m.xMin) * m.invXMaxMin) * (float)(m.xPoints - 1), m.xPoints);
    const int32_t yPoint = map_floor(((y - m.yMin) * m.invYMaxMin) * (float)(m.yPoints - 1), m.yPoints);
    const int32_t xLow = map_clamp(xPoint, 0, m.xPoints - 1), yLow = map_clamp(yPoint, 0, m.yPoints - 1);
    const int32_t xHigh = map_clamp(xPoint + 1, 0, m.xPoints - 1), yHigh = map_clamp(yPoint + 1, 0, m.yPoints - 1);
    if (xLow == xHigh && yLow == yHigh) return m.data[xLow + yLow * m.xPoints];
    const float v00 = (float)m.data[xLow + yLow * m.xPoints];
    const float v01 = (float)m.data[xLow + yHigh * m.xPoints];
    const float v10 = (float)m.data[xHigh + yLow * m.xPoints];
    const float v11 = (float)m.data[xHigh + yHigh * m.xPoints];
    const float x0 = x - m.xMin;
    const float y0 = y - m.yMin;
    const float x1 = (float)xLow * m.xRes;
    const float x2 = (float)xHigh * m.xRes;
    const float y1 = (float)yLow * m.yRes;
    const float y2 = (float)yHigh * m.yRes;
    const float invX21 = 1.0f / (x2 - x1);
    const float c1 = (x2 - x0) * invX21;
    const float c2 = (x0 - x1) * invX21;
    float fxy1 = (c1 * v00) + (c2 * v10);
    float fxy2 = (c1 * v01) + (c2 * v11);
    if (xLow == xHigh) {
        fxy1 = v00;
        fxy2 = v01;
    } else if (yLow == yHigh) {
        return fxy1;
    }
    const float invY21 = 1.0f / (y2 - y1);
    return (((y2 - y0) * invY21) * fxy1) + (((y0 - y1) * invY21) * fxy2);
}

// applyForce: IceBlock (MainIceBlock.cpp:299-400, clamp = HDM_TYPE_FIXED) or
// GlacierDem (MainDem.cpp:145-259: bed height, friction and external force
// from maps (bit k of mapMask: maps[k]) or the analytic bed and constant
// force of ip, HiDEMDefVelocityLimit and HiDEMDefBoundaryCondition), on one
// particle's state in registers: cord/theta/diff are replaced by the new
// coordinates, rotations and Diff.
__device__ __forceinline__ void integrate_state(const hdm_integ_params &ip, const MapDev *maps, int32_t mapMask,
                                                bool fixed, double radius, double waterLevel, const double Fin[6],
                                                double cord[3], double theta[3], double diff[6]) {
    if (ip.variant == HDM_INTEG_ICEBLOCK && fixed) {
#pragma unroll
        for (int c = 0; c < 6; c++) diff[c] = 0;
        return;
    }
    double F[6] = {Fin[0], Fin[1], Fin[2], Fin[3], Fin[4], Fin[5]};
    double cordPrev[3], thetaPrev[3], vel[6];
    const double diameter = radius * 2.0;
    const double zCord = cord[2];
    const double particleMass = ip.rhoIce * diameter * diameter * diameter;
    const double particleMomentInertia = particleMass * diameter * diameter / 6.0;
    double dragCoeff;
    if (ip.variant == HDM_INTEG_ICEBLOCK) {
        dragCoeff = (zCord > waterLevel) ? ip.airDrag : ip.waterDrag;
        if (zCord > waterLevel) F[2] += -ip.g * particleMass;
        if (zCord <= waterLevel) F[2] += ip.lift * particleMass;
        central_difference(dragCoeff, particleMass, particleMomentInertia, ip, F, diff, cord, cordPrev, theta,
                           thetaPrev, vel);
    } else {
        // InputBed::getHeight / getFric (InputGeometry.h:57-67): the
        // coordinates go to Map2d::interp as float
        const float fx = (float)cord[0], fy = (float)cord[1];
        const double bedzCord = (mapMask & (1 << HDM_MAP_BED)) ? map_interp(maps[HDM_MAP_BED], fx, fy)
                                                               : ((cord[0] < ip.bedXEnd) ? ip.bedZ : ip.deepZ);
        const double fric = (mapMask & (1 << HDM_MAP_FRICTION)) ? map_interp(maps[HDM_MAP_FRICTION], fx, fy)
                                                                : ((cord[0] < ip.frictionXEnd) ? 1.0 : 0.0);
        if ((zCord + radius) - bedzCord < ip.bedInteract) {
            double d = ip.bedDrag * fric;
            dragCoeff = (ip.frictionCap < d) ? ip.frictionCap : d;
        } else if (zCord > waterLevel) {
            dragCoeff = ip.airDrag;
        } else {
            dragCoeff = ip.waterDrag;
        }
        if (zCord > waterLevel) F[2] += -ip.g * particleMass;
        if (zCord <= waterLevel) F[2] += ip.lift * particleMass;
        if (zCord + radius < bedzCord) F[2] += ip.bedStiffness * (bedzCord - (zCord + radius));
        if (mapMask >> HDM_MAP_EXT) {
            for (int c = 0; c < 6; c++)
                F[c] += (mapMask & (1 << (HDM_MAP_EXT + c))) ? map_interp(maps[HDM_MAP_EXT + c], fx, fy) : ip.ext[c];
        } else {
#pragma unroll
            for (int c = 0; c < 6; c++) F[c] += ip.ext[c];
        }
        central_difference(dragCoeff, particleMass, particleMomentInertia, ip, F, diff, cord, cordPrev, theta,
                           thetaPrev, vel);
        if (ip.velocityLimit) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (vel[c] > ip.maxVel) cord[c] = ip.dt * ip.maxVel + cordPrev[c];
                if (vel[3 + c] > ip.maxRot) theta[c] = ip.dt * ip.maxRot + thetaPrev[c];
            }
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (vel[c] < -ip.maxVel) cord[c] = ip.dt * -ip.maxVel + cordPrev[c];
                if (vel[3 + c] < -ip.maxRot) theta[c] = ip.dt * -ip.maxRot + thetaPrev[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (cord[c] + radius < ip.lower[c] && ip.bound[2 * c]) cord[c] = cordPrev[c];
            if (cord[c] - radius > ip.upper[c] && ip.bound[2 * c + 1]) cord[c] = cordPrev[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        diff[c] = cord[c] - cordPrev[c];
        diff[3 + c] = theta[c] - thetaPrev[c];
    }
}

// Integrates particle i under force Fin from the IntegArgs state arrays;
// returns its new coordinates and translational Diff.
__device__ __forceinline__ void integrate_particle(const IntegArgs &a, int64_t i, const double Fin[6],
                                                   double newCord[3], double newDiff[3]) {
    double cord[3] = {a.x[i], a.y[i], a.z[i]};
    double theta[3] = {a.tx[i], a.ty[i], a.tz[i]};
    double diff[6] = {a.d0[i], a.d1[i], a.d2[i], a.d3[i], a.d4[i], a.d5[i]};
    const bool fixed = (a.type[i] & HDM_TYPE_FIXED) != 0;
    const double waterLevel = a.blockWater ? a.blockWater[a.blockId[i]] : a.ip.waterLevel;
    integrate_state(a.ip, a.maps, a.mapMask, fixed, a.radius[i], waterLevel, Fin, cord, theta, diff);
    a.ox[i] = cord[0];
    a.oy[i] = cord[1];
    a.oz[i] = cord[2];
    a.otx[i] = theta[0];
    a.oty[i] = theta[1];
    a.otz[i] = theta[2];
    a.od0[i] = diff[0];
    a.od1[i] = diff[1];
    a.od2[i] = diff[2];
    a.od3[i] = diff[3];
    a.od4[i] = diff[4];
    a.od5[i] = diff[5];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        newCord[c] = cord[c];
        newDiff[c] = diff[c];
    }
}

// Workgroup maxima of the rebuild-test quantities of the particles this
// workgroup integrated, written as one row of a.partials (every quantity is
// >= 0, so the maxima start from 0 and inactive lanes contribute 0); the rows
// are folded by k_skin_finalize after the launch.
template <int BLOCK = TPB>
__device__ __forceinline__ void skin_partials(const IntegArgs &a, bool active, int64_t i, const double cord[3],
                                              const double diff[3]) {
    __shared__ double rows[BLOCK / 64][4];
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    if (active) {
        for (int c = 0; c < 3; c++) m[c] = diff[c] > 0.0 ? diff[c] : 0.0;
        if (a.x0) {
            double ddx = cord[0] - a.x0[i], ddy = cord[1] - a.y0[i], ddz = cord[2] - a.z0[i];
            m[3] = ddx * ddx + ddy * ddy + ddz * ddz;
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) m[c] = wave_max(m[c]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 4; c++) rows[w][c] = m[c];
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = rows[0][threadIdx.x];
        for (int k = 1; k < BLOCK / 64; k++) v = rows[k][threadIdx.x] > v ? rows[k][threadIdx.x] : v;
        a.partials[(int64_t)blockIdx.x * 4 + threadIdx.x] = v;
    }
}

struct FinalizeArgs {
    const double *partials;
    int64_t rows;
    double *stage;        // [gridDim.x][4]
    uint32_t *ticket;
    uint64_t *skin;
};

// Folds the rows of the integrating launch: each workgroup reduces a slice
// into stage, and the last one (a ticket after a release fence: one fence
// per workgroup of this small launch) stores skin[0..2], skin[4] as ordered
// keys for the host-checked rebuild policies.
__global__ void __launch_bounds__(TPB) k_skin_finalize(FinalizeArgs a) {
    __shared__ double part[TPB / 64][4];
    __shared__ int last;
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t r = (int64_t)blockIdx.x * TPB + threadIdx.x; r < a.rows; r += (int64_t)gridDim.x * TPB)
        for (int c = 0; c < 4; c++) {
            const double v = a.partials[r * 4 + c];
            m[c] = v > m[c] ? v : m[c];
        }
#pragma unroll
    for (int c = 0; c < 4; c++) m[c] = wave_max(m[c]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 4; c++) part[w][c] = m[c];
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = part[0][threadIdx.x];
        for (int k = 1; k < TPB / 64; k++) v = part[k][threadIdx.x] > v ? part[k][threadIdx.x] : v;
        a.stage[blockIdx.x * 4 + threadIdx.x] = v;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(a.ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    double f[4] = {0.0, 0.0, 0.0, 0.0};
    const unsigned long long *st = reinterpret_cast<const unsigned long long *>(a.stage);
    for (uint32_t b = 0; b < gridDim.x; b++)
        for (int c = 0; c < 4; c++) {
            const double v = __longlong_as_double(
                (long long)__hip_atomic_load(st + b * 4 + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            f[c] = v > f[c] ? v : f[c];
        }
    for (int c = 0; c < 3; c++) a.skin[c] = ordered_key(f[c]);
    a.skin[4] = ordered_key(f[3]);
    *a.ticket = 0;
}

template <class FO>
__global__ void __launch_bounds__(TPB) k_integrate(IntegArgs a, const FO *__restrict__ f0, const FO *__restrict__ f1,
                                                   const FO *__restrict__ f2, const FO *__restrict__ f3,
                                                   const FO *__restrict__ f4, const FO *__restrict__ f5) {
    const int64_t i = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    const bool active = i < a.n;
    double cord[3] = {0, 0, 0}, diff[3] = {0, 0, 0};
    if (active) {
        const double F[6] = {(double)f0[i], (double)f1[i], (double)f2[i], (double)f3[i], (double)f4[i], (double)f5[i]};
        integrate_particle(a, i, F, cord, diff);
    }
    if (a.partials) skin_partials(a, active, i, cord, diff);
}

// ---- fused step ------------------------------------------------------------------
// Gather records: MixedRec (kernel_common.h) in the FP32 form; the FP64 form
// gathers rec = double4 {x, y, z, radius} and the FP64 Diff from drec.
template <class R> struct RecOf { using type = double4; };
template <> struct RecOf<float> { using type = MixedRec; };
__device__ __forceinline__ double rec_radius(const double4 &P) { return P.w; }
__device__ __forceinline__ float rec_radius(const MixedRec &P) { return P.radius; }

// Gather records (see Context::rec / drec) and the break prefilter inputs.
template <class R>
struct StepRecs {
    using Real = R;
    using Rec = typename RecOf<R>::type;
    const Rec *rec;            // current state
    const double *drec;        // FP64 form: every particle's FP64 Diff; FP32 form: the escapes
    const MixedTail *tail;     // FP32 form: rest of the own Diff (split_diff3)
    Rec *recOut;               // integrated state
    double *drecOut;
    MixedTail *tailOut;
    double tame;               // Diff bound behind the radius flag of rec
    // Break prefilter inputs: per side a 16-bit code, code * thrScale at most
    // L0 (1 + maxLoad) with per-link maxLoad (perLinkMaxLoad) else L0, 0 when
    // L0 < 1e-30 or not finite; the largest L0 of the state (rounded up)
    const uint16_t *slotThr;
    float thrScale;
    int32_t prefilter;         // 0: every intact side takes the exact break test
    int32_t perLinkMaxLoad;
    float scalarMaxLoad, bendBreak, l0Max;
    int32_t storeForce;        // write x..wForce (the last step of a run; earlier ones are overwritten unseen)
};

// Upper bound of max |theta_p - theta_q| over components, as a float: the sum
// of absolute values, NaN/inf when any component is (then the prefilter
// defers to the exact test).
__device__ __forceinline__ float rotation_bound(const double t[3]) {
    return (float)(fabs(t[0]) + fabs(t[1]) + fabs(t[2]));
}

// Doubles per particle in drec: the FP64 form's partners gather it (Diff0..5,
// rotation bound, 0); in the FP32 form it is the escape array of
// split_diff3 (Diff0..2, written and read only for particles whose
// translation Diff does not split exactly), the rest travelling in the
// MixedRec and the MixedTail.
inline int64_t drec_stride(int32_t prec) { return prec == HDM_PREC_FP64 ? 8 : 3; }

__device__ __forceinline__ void load_drec(const double *drec, int64_t q, double o[6], float &rot) {
    const double4 *v = reinterpret_cast<const double4 *>(drec + q * 8);
    const double4 a = v[0], b = v[1];
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; o[4] = b.x; o[5] = b.y;
    rot = (float)b.z;
}

template <class R>
__device__ __forceinline__ void store_records(const StepRecs<R> &r, int64_t p, const double cord[3],
                                              const double theta[3], const double diff[6], double radius) {
    bool tame = fabs(radius) <= 5e5;
#pragma unroll
    for (int c = 0; c < 6; c++) tame &= fabs(diff[c]) <= r.tame;
    if constexpr (sizeof(R) == 4) {
        MixedRec m;
        m.x = cord[0];
        m.y = cord[1];
        m.z = cord[2];
        float lo[3];
        const bool exact = split_diff3(diff, m.d, lo);
#pragma unroll
        for (int c = 3; c < 6; c++) m.d[c] = (float)diff[c];
        m.rot = rotation_bound(theta);
        m.radius = tame ? (float)radius : __builtin_nanf("");
        m.lo[0] = lo[0];
        m.lo[1] = lo[1];
        r.recOut[p] = m;
        MixedTail t;
#pragma unroll
        for (int c = 0; c < 3; c++) t.dr[c] = diff[3 + c];
        t.lo2 = lo[2];
        t.pad = 0.0f;
        r.tailOut[p] = t;
        if (!exact) {
#pragma unroll
            for (int c = 0; c < 3; c++) r.drecOut[p * 3 + c] = diff[c];
        }
    } else {
        r.recOut[p] = make_double4(cord[0], cord[1], cord[2], tame ? radius : __builtin_nan(""));
        double4 *d = reinterpret_cast<double4 *>(r.drecOut + p * 8);
        d[0] = make_double4(diff[0], diff[1], diff[2], diff[3]);
        d[1] = make_double4(diff[4], diff[5], (double)rotation_bound(theta), 0.0);
    }
}

// A partner's Diff (as the laws take it) and the pair differences, from its
// gathered record: FP32 from the record itself, FP64 from drec.
template <class R>
__device__ __forceinline__ void partner_diff(const StepRecs<R> &r, int32_t q, const typename RecOf<R>::type &Q,
                                             const double diff[6], R other[6], R d[6], float &rot) {
    if constexpr (sizeof(R) == 4) {
#pragma unroll
        for (int c = 0; c < 6; c++) {
            other[c] = Q.d[c];
            d[c] = (R)(diff[c] - (double)Q.d[c]);
        }
        rot = Q.rot;
    } else {
        double od[6];
        load_drec(r.drec, q, od, rot);
#pragma unroll
        for (int c = 0; c < 6; c++) other[c] = (R)od[c];
        diff_delta<R>(diff, od, d);
    }
}

// Break prefilter: true when the link certainly stays intact.  The exact test
// (s - L0) / L0 + bendBreak * maxBend >= maxLoad, with L0 > 0, is
// s + bendBreak * maxBend * L0 >= L0 (1 + maxLoad) =: T.  In float, with the
// rotation bound (at least maxBend) and the largest L0 standing in for the
// link's own, the left side can only grow, and T comes from a 16-bit code
// that rounds it down (k_side_codes); the link passes when that estimate is
// below T by more than 1e-5 of the magnitudes.  Float rounding of the inputs
// and operations is below 1e-6 of those magnitudes and bendBreak >= 0 (the
// host disables the prefilter otherwise), so a link the exact FP64 test would
// break never passes.  NaN/inf inputs and code 0 (L0 < 1e-30) fail the
// comparison and go to the exact test.
__device__ __forceinline__ bool certainly_intact(float T, float bendBreak, float l0Max, double dx, double dy,
                                                 double dz, float rotBound) {
    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
    const float s = sqrtf(fx * fx + fy * fy + fz * fz);
    const float lhs = s + bendBreak * rotBound * l0Max;
    const float margin = 1e-5f * (lhs + fabsf(T));
    return lhs + margin < T;
}

// Streaming accesses of the fused step: every CT value, slot word, list entry
// and per-particle SoA field is touched once per launch, so it is loaded and
// stored non-temporally and leaves L2 to the gathered partner records, which
// each particle's neighbours fetch again.
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
    return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void st_stream(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ float2 ld_stream(const float2 *p) {
    const unsigned long long u = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(p));
    float2 r;
    __builtin_memcpy(&r, &u, sizeof(r));
    return r;
}

// Contacts of the linked pairs flagged in `sides` (bit k: side k's pair is
// not provably inert), after every beam, in side order: outside the link
// loop, so that loop does not carry the contact law's registers.
template <class R, class CT>
__device__ __forceinline__ void contact_sides(const ForceArgs<R, CT, R> &f, const StepRecs<R> &r,
                                                        uint64_t sides, int64_t base, double xCord, double yCord,
                                                        double zCord, const double diff[6], const R self[7],
                                                        R F[6]) {
    while (sides) {
        const int k = __builtin_ctzll(sides);
        sides &= sides - 1;
        const int32_t q = f.slotOther[base + (int64_t)k * 64] & SLOT_LINK_MASK;
        const auto Q = r.rec[q];
        R other[7], d[6];
        float rot;
        partner_diff<R>(r, q, Q, diff, other, d, rot);
        other[6] = (R)f.radius[q];
        contact_law<R>(R(1), (R)(xCord - Q.x), (R)(yCord - Q.y), (R)(zCord - Q.z), self, other, d, f.fp, F);
    }
}

// One timestep of hdm_run in one launch: the link break of the previous
// timestep (when f.breakCheck), force on particle p from the current state
// (CombinedAccel::iterateCombined, as k_force), then its integration (as
// k_integrate), with the rebuild-test partials.
//
// The records are the state: a particle's coordinates and Diff are read from
// its own records (the lines its neighbours gather) and the integrated ones
// are written to the other record buffers only; the SoA coordinates and Diff
// are unpacked from the records when something else needs them (sync_state).
// The rotations stay SoA (ping-pong), read at integration and by the rare
// exact break test; the prefilter takes their bound from the records.
//
// Partner state comes from the records too.  FP64 form: one 32 B load of rec
// decides a contact pair that is provably inert (contact_is_inert: beyond
// radSum + attractDistance with every Diff bounded by tame, so every term the
// law adds is +-0); only other pairs, and link partners, load drec.  FP32
// form: one 64 B MixedRec per partner.  The break test of an intact side runs
// exactly (FP64, as k_break) unless the float prefilter proves the link stays
// intact.
// BLOCK (workgroup size), XCD (the per-XCD contiguous renumbering of
// workgroups) and OCC (one more wave per SIMD than the register count allows,
// bought with a few spills) are the autotuned launch shape; every shape
// computes the same per-particle arithmetic, so results do not depend on it.
template <class R>
constexpr int step_waves(bool occ) {
    return sizeof(R) == 4 ? (occ ? 5 : 4) : (occ ? 4 : 3);
}

// The device guard's test for particle p at its new position c: the exact
// test 2 |c - x0| >= guardThr in FP64, decided first from the snapshot
// rounded to FP32 (12 B instead of 24 B read per particle).  Rounding to
// nearest moves a component by less than 2^-23 of its FP32 value (and by
// less than 1e-30 below the normal range), so the FP64 distance lies within
// e = (|x0f| + |y0f| + |z0f|) 2^-23 + 1e-30 of the distance d to the FP32
// snapshot.  With a margin of 1e-12 (d + e) for the rounding of both
// evaluations, d outside [thr/2 - e, thr/2 + e] decides the test as the FP64
// evaluation would; only the rest (and non-finite values, which fail both
// comparisons) read the FP64 snapshot.
__device__ __forceinline__ bool guard_trips(const IntegArgs &g, int64_t p, const double c[3]) {
    const float fx = ld_stream(g.x0f + p), fy = ld_stream(g.y0f + p), fz = ld_stream(g.z0f + p);
    const double ax = c[0] - (double)fx, ay = c[1] - (double)fy, az = c[2] - (double)fz;
    const double d = sqrt(ax * ax + ay * ay + az * az);
    const double e = (fabs((double)fx) + fabs((double)fy) + fabs((double)fz)) * 0x1p-23 + 1e-30;
    const double m = (d + e) * 1e-12;
    if (2.0 * (d + e + m) < g.guardThr) return false;
    if (d - e - m > 0.0 && 2.0 * (d - e - m) >= g.guardThr) return true;
    const double ddx = c[0] - g.x0[p], ddy = c[1] - g.y0[p], ddz = c[2] - g.z0[p];
    return 2.0 * sqrt(ddx * ddx + ddy * ddy + ddz * ddz) >= g.guardThr;
}

// One particle of the fused step (the body of k_step and k_step_loop):
// break of the previous step, force, integration and the new records;
// cordN / diffN receive the integrated coordinates and translation Diff, and
// broke counts the links it broke at their smaller-index end.
template <class R, class CT>
__device__ __forceinline__ void step_particle(const ForceArgs<R, CT, R> &f, const IntegArgs &g, const StepRecs<R> &r,
                                              const int64_t p, double cordN[3], double diffN[3], uint32_t &broke) {
    {
        const auto P = r.rec[p];
        const double xCord = P.x, yCord = P.y, zCord = P.z;
        double diff[6];
        float rotSelf;
        if constexpr (sizeof(R) == 4) {
            join_diff6(P, r.tail[p], r.drec, p, diff);
            rotSelf = P.rot;
        } else {
            load_drec(r.drec, p, diff, rotSelf);
        }
        const double radius = ld_stream(g.radius + p);
        R self[7] = {(R)diff[0], (R)diff[1], (R)diff[2], (R)diff[3], (R)diff[4], (R)diff[5], (R)radius};
        R F[6] = {0, 0, 0, 0, 0, 0};
        const int64_t lane = p & 63;
        {
            const int32_t cnt = ld_stream(f.nbrCount + p);
            const int32_t *nl = f.nbr + (p >> 6) * (int64_t)64 * f.nbrWidth + lane;
            // each entry's index is loaded one iteration ahead, so only the
            // record gather waits inside the iteration
            int32_t qNext = cnt > 0 ? ld_stream(nl) : 0;
            for (int32_t k = 0; k < cnt; k++) {
                const int32_t q = qNext;
                if (k + 1 < cnt) qNext = ld_stream(nl + (int64_t)(k + 1) * 64);
                const auto Q = r.rec[q];
                const R xDist = (R)(xCord - Q.x);
                const R yDist = (R)(yCord - Q.y);
                const R zDist = (R)(zCord - Q.z);
                {
                    // contact_is_inert with contact_law's own expressions; the
                    // record radius is NaN unless both the Diff (damping terms)
                    // and the radius are within their bounds
                    const R particleDistance2 = xDist * xDist + yDist * yDist + zDist * zDist;
                    const R radSum = (R)rec_radius(P) + (R)rec_radius(Q);
                    const R activeDist = radSum + f.fp.attractDistance;
                    if (particleDistance2 >= activeDist * activeDist && particleDistance2 <= huge_value<R>() &&
                        particleDistance2 > R(0))
                        continue;
                }
                R other[7], d[6];
                float rotOther;
                partner_diff<R>(r, q, Q, diff, other, d, rotOther);
                other[6] = (R)f.radius[q];
                contact_law<R>(R(1), xDist, yDist, zDist, self, other, d, f.fp, F);
            }
        }
        {
            const int32_t cnt = (int32_t)ld_stream(f.linkCount + p);
            const int64_t base = (p >> 6) * (int64_t)64 * f.linkWidth + lane;
            uint64_t activeSides = 0;  // linked pairs in contact (FP32, merged list; cnt <= 64)
            // Software-pipelined: a side's partner word is loaded one
            // iteration ahead, and its record, CT and threshold code are
            // issued together, so a side waits for one memory latency (the
            // slowest of them) instead of a chain of dependent ones.
            int32_t owNext = cnt > 0 ? ld_stream(f.slotOther + base) : 0;  // partner; sign bit: broken
            // vmcnt(0) here: nothing loaded before the loop is still pending
            // at its header, so the header needs no wait (which, with the
            // previous side's CT stores outstanding, would drain them too).
            // gfx9 s_waitcnt immediate: vmcnt in bits 3:0 and 15:14 (0),
            // expcnt in 6:4 and lgkmcnt in 11:8 at their maxima (no wait).
            __builtin_amdgcn_s_waitcnt(0x0F70);
            for (int32_t k = 0; k < cnt; k++) {
                // FP64 form: the wave issues the side's loads at raised
                // priority, so the loads of all waves get in flight before
                // any returns to its arithmetic (measured 2-3% faster; the
                // FP32 form measured 1% slower with it)
                if constexpr (sizeof(R) == 8) __builtin_amdgcn_s_setprio(1);
                const int64_t sidx = base + (int64_t)k * 64;
                const int32_t ow = owNext;
                const int32_t q = ow & SLOT_LINK_MASK;
                const auto Q = r.rec[q];
                CT cs[6];
                ct_load(f.ct, sidx, cs);
                // both loads unconditional (slotThr has the slot arrays'
                // extent; the last side re-reads its own word), so no branch
                // splits the issue from the one wait
                const uint16_t code = ld_stream(r.slotThr + sidx);
                owNext = ld_stream(f.slotOther + (k + 1 < cnt ? sidx + 64 : sidx));
                if constexpr (sizeof(R) == 8) __builtin_amdgcn_s_setprio(0);
                R other[7], d[6];
                float rotOther;
                partner_diff<R>(r, q, Q, diff, other, d, rotOther);
                other[6] = R(0);  // no radius needed
                const double dx = xCord - Q.x, dy = yCord - Q.y, dz = zCord - Q.z;
                // The prefilter is evaluated for every side and folded into
                // one branch condition, so the code's load cannot be sunk
                // into the break branch (where its wait would come late).
                const float thr = (float)code * r.thrScale;
                const float T = r.perLinkMaxLoad ? thr : thr * (1.0f + r.scalarMaxLoad);
                const bool skip =
                    (r.prefilter != 0) & certainly_intact(T, r.bendBreak, r.l0Max, dx, dy, dz, rotSelf + rotOther);
                R mask = ow < 0 ? R(0) : R(1);  // broken link (state == 0)
                if ((ow >= 0) & (f.breakCheck != 0) & !skip) {
                    const int32_t word = f.slotLink[sidx];  // link id (intact: sign bit clear)
                    const double theta[3] = {g.tx[p], g.ty[p], g.tz[p]};
                    if (link_breaks(f, word, q, dx, dy, dz, theta)) {
                        // both ends reach the same verdict; each flags its own slot
                        mask = 0;
                        f.slotLinkW[sidx] = word | SLOT_BROKEN;
                        f.slotOtherW[sidx] = ow | SLOT_BROKEN;
                        f.state[word] = 0;
                        broke += p < q ? 1u : 0u;
                    }
                }
                R ct[6] = {(R)cs[0], (R)cs[1], (R)cs[2], (R)cs[3], (R)cs[4], (R)cs[5]};
                beam_law<R>(mask, (R)dx, (R)dy, (R)dz, self, other, d, f.fp, ct, F);
                if (sizeof(R) == 4 && f.linkedContacts) {
                    // the pair's contact test (the contact loop's record test);
                    // the rare active ones are applied after the loop, which
                    // keeps the contact law out of this loop's registers
                    // (placed after the beam law, where it does not delay the
                    // side's loads)
                    const R xDist = (R)dx, yDist = (R)dy, zDist = (R)dz;
                    const R particleDistance2 = xDist * xDist + yDist * yDist + zDist * zDist;
                    const R radSum = (R)rec_radius(P) + (R)rec_radius(Q);
                    const R activeDist = radSum + f.fp.attractDistance;
                    if (!(particleDistance2 >= activeDist * activeDist && particleDistance2 <= huge_value<R>() &&
                          particleDistance2 > R(0)))
                        activeSides |= 1ull << k;
                }
#pragma unroll
                for (int c = 0; c < 6; c++) cs[c] = (CT)ct[c];
                ct_store(f.ct, sidx, cs);
            }
            if (sizeof(R) == 4 && activeSides) contact_sides(f, r, activeSides, base, xCord, yCord, zCord, diff, self, F);
        }
        if (r.storeForce) store_force(f, p, F);
        const double Fd[6] = {(double)F[0], (double)F[1], (double)F[2], (double)F[3], (double)F[4], (double)F[5]};
        double cord[3] = {xCord, yCord, zCord};
        double theta[3] = {ld_stream(g.tx + p), ld_stream(g.ty + p), ld_stream(g.tz + p)};
        const bool fixed = (ld_stream(g.type + p) & HDM_TYPE_FIXED) != 0;
        const double waterLevel = g.blockWater ? g.blockWater[g.blockId[p]] : g.ip.waterLevel;
        integrate_state(g.ip, g.maps, g.mapMask, fixed, radius, waterLevel, Fd, cord, theta, diff);
        st_stream(g.otx + p, theta[0]);
        st_stream(g.oty + p, theta[1]);
        st_stream(g.otz + p, theta[2]);
        store_records(r, p, cord, theta, diff, radius);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            cordN[c] = cord[c];
            diffN[c] = diff[c];
        }
    }
}

template <class R, class CT, int BLOCK, bool XCD, bool OCC>
__global__ void __launch_bounds__(BLOCK, step_waves<R>(OCC)) k_step(ForceArgs<R, CT, R> f, IntegArgs g,
                                                                    StepRecs<R> r) {
    if (g.guard && *g.guard) {  // an earlier step found the neighbour list stale
        if (blockIdx.x == 0 && threadIdx.x == 0) *g.guardOut = 1u;
        return;
    }
    const int64_t wg = XCD ? xcd_block_index(gridDim.x) : (int64_t)blockIdx.x;
    uint32_t broke = 0;  // links this lane's particle broke (counted at its smaller-index end)
    const int64_t p = wg * BLOCK + threadIdx.x;
    const bool active = p < f.n;
    double cordN[3] = {0, 0, 0}, diffN[3] = {0, 0, 0};
    if (active) step_particle<R, CT>(f, g, r, p, cordN, diffN, broke);
    if (f.brokenCount) {
        broke = wave_sum_u32(broke);
        if ((threadIdx.x & 63) == 0 && broke) atomicAdd((unsigned long long *)f.brokenCount, (unsigned long long)broke);
    }
    if (g.guardOut) {
        if (active && guard_trips(g, p, cordN)) *g.guardOut = 1u;
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(g.stepsDone, 1u);
    } else if (g.partials) {
        skin_partials<BLOCK>(g, active, p, cordN, diffN);
    }
}

// The link break that ends a run in the FP32 form: linkUpdate of its last
// step (HiDEMLib.h:256-269, HiDEMDefLinkBreakFunc) applied per particle from
// the records, as the next fused step would apply it, so the SoA state need
// not be unpacked first.  Every intact side of p takes the float prefilter and,
// when that cannot prove the link intact, the exact FP64 test; both ends
// reach the same verdict (the pair differences are antisymmetric) and flag
// their own slot, and the end with the smaller index counts the link.  The
// links broken are those k_break would break.
template <class R>
__global__ void __launch_bounds__(TPB) k_break_recs(ForceArgs<R, R, R> f, StepRecs<R> r, uint64_t *counter) {
    static_assert(sizeof(R) == 4, "the FP64 form's run-end break is k_break_slots");
    const int64_t p = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    uint32_t broke = 0;
    if (p < f.n) {
        const auto P = r.rec[p];
        const float rotSelf = P.rot;
        const int32_t cnt = (int32_t)f.linkCount[p];
        const int64_t base = (p >> 6) * (int64_t)64 * f.linkWidth + (p & 63);
        // one intact side: prefilter, then the exact test where it cannot
        // prove the link intact
        auto side = [&](int64_t sidx, int32_t ow, double Qx, double Qy, double Qz, float rotOther, uint16_t code) {
            const int32_t q = ow & SLOT_LINK_MASK;
            const double dx = P.x - Qx, dy = P.y - Qy, dz = P.z - Qz;
            const float thr = (float)code * r.thrScale;
            const float T = r.perLinkMaxLoad ? thr : thr * (1.0f + r.scalarMaxLoad);
            if (r.prefilter != 0 && certainly_intact(T, r.bendBreak, r.l0Max, dx, dy, dz, rotSelf + rotOther))
                return;
            const int32_t word = f.slotLink[sidx];
            const double theta[3] = {f.tx[p], f.ty[p], f.tz[p]};
            if (link_breaks(f, word, q, dx, dy, dz, theta)) {
                f.slotLinkW[sidx] = word | SLOT_BROKEN;
                f.slotOtherW[sidx] = ow | SLOT_BROKEN;
                f.state[word] = 0;
                broke += p < q ? 1u : 0u;
            }
        };
        // Pipelined two deep: the partner word two sides ahead, the
        // partner's coordinates and rotation bound and the side's code
        // one side ahead, so a side waits for loads issued a side
        // earlier.  Past the last side the loads re-read the last side
        // (valid addresses, results unused), and a broken side's word
        // still names a valid partner.
        auto fetch = [&](int32_t ow, int64_t sidx, double &x, double &y, double &z, float &rot, uint16_t &c) {
            const int32_t q = ow & SLOT_LINK_MASK;
            x = r.rec[q].x;
            y = r.rec[q].y;
            z = r.rec[q].z;
            rot = r.rec[q].rot;
            c = ld_stream(r.slotThr + sidx);
        };
        int32_t owCur = cnt > 0 ? ld_stream(f.slotOther + base) : 0;
        int32_t owNext = cnt > 1 ? ld_stream(f.slotOther + base + 64) : owCur;
        double nx = 0, ny = 0, nz = 0;
        float nrot = 0;
        uint16_t ncode = 0;
        if (cnt > 0) fetch(owCur, base, nx, ny, nz, nrot, ncode);
        for (int32_t k = 0; k < cnt; k++) {
            const int64_t sidx = base + (int64_t)k * 64;
            const int32_t ow = owCur;
            const double Qx = nx, Qy = ny, Qz = nz;
            const float rotOther = nrot;
            const uint16_t code = ncode;
            const bool more = k + 1 < cnt;
            fetch(more ? owNext : ow, more ? sidx + 64 : sidx, nx, ny, nz, nrot, ncode);
            owCur = owNext;
            owNext = ld_stream(f.slotOther + (k + 2 < cnt ? sidx + 128 : sidx));
            if (ow >= 0) side(sidx, ow, Qx, Qy, Qz, rotOther, code);  // ow < 0: broken already
        }
    }
    broke = wave_sum_u32(broke);
    if ((threadIdx.x & 63) == 0 && broke) {
        if (counter) atomicAdd((unsigned long long *)counter, (unsigned long long)broke);
        if (f.brokenCount) atomicAdd((unsigned long long *)f.brokenCount, (unsigned long long)broke);
    }
}

// The same run-end break with one thread per link slot instead of one per
// particle: thread s is tiled slot s = (t * linkWidth + k) * 64 + lane, side k
// of particle t * 64 + lane, so a wave holds one side of 64 consecutive
// particles (its slot words, codes and own records coalesced) and no thread
// waits on more than one side.  Same tests, flags and counting as
// k_break_recs; the FP64 form's run-end break (0.68 against 0.85 ms for the
// per-particle walk on the 8M slab, which stays the faster for the FP32
// records: 0.62 against 0.81 ms).
template <class R>
__global__ void __launch_bounds__(TPB) k_break_slots(ForceArgs<R, R, R> f, StepRecs<R> r, int64_t nSlots,
                                                     uint64_t *counter) {
    const int64_t s = xcd_block_index(gridDim.x) * TPB + threadIdx.x;
    uint32_t broke = 0;
    if (s < nSlots) {
        const int64_t w64 = (int64_t)f.linkWidth * 64;
        const int64_t t = s / w64;
        const int32_t k = (int32_t)((s - t * w64) >> 6);
        const int64_t p = t * 64 + (s & 63);
        if (p < f.n && k < (int32_t)f.linkCount[p]) {
            const int32_t ow = ld_stream(f.slotOther + s);
            if (ow >= 0) {  // ow < 0: broken already
                const int32_t q = ow & SLOT_LINK_MASK;
                const auto P = r.rec[p];
                const auto Q = r.rec[q];
                const uint16_t code = ld_stream(r.slotThr + s);
                float rotSelf, rotOther;
                if constexpr (sizeof(R) == 4) {
                    rotSelf = P.rot;
                    rotOther = Q.rot;
                } else {
                    rotSelf = (float)r.drec[p * 8 + 6];
                    rotOther = (float)r.drec[(int64_t)q * 8 + 6];
                }
                const double dx = P.x - Q.x, dy = P.y - Q.y, dz = P.z - Q.z;
                const float thr = (float)code * r.thrScale;
                const float T = r.perLinkMaxLoad ? thr : thr * (1.0f + r.scalarMaxLoad);
                if (!(r.prefilter != 0 && certainly_intact(T, r.bendBreak, r.l0Max, dx, dy, dz, rotSelf + rotOther))) {
                    const int32_t word = f.slotLink[s];
                    const double theta[3] = {f.tx[p], f.ty[p], f.tz[p]};
                    if (link_breaks(f, word, q, dx, dy, dz, theta)) {
                        f.slotLinkW[s] = word | SLOT_BROKEN;
                        f.slotOtherW[s] = ow | SLOT_BROKEN;
                        f.state[word] = 0;
                        broke = p < q ? 1u : 0u;
                    }
                }
            }
        }
    }
    broke = wave_sum_u32(broke);
    if ((threadIdx.x & 63) == 0 && broke) {
        if (counter) atomicAdd((unsigned long long *)counter, (unsigned long long)broke);
        if (f.brokenCount) atomicAdd((unsigned long long *)f.brokenCount, (unsigned long long)broke);
    }
}

// ---- persistent timestep ------------------------------------------------------
// Grid barrier of a co-resident (cooperative) launch, in two levels so no
// word takes more than a few dozen arrivals: workgroup b arrives at group
// b % GRID_GROUPS, the last arrival of a group at the top word, and the last
// arrival there re-arms the counters and advances the generation the others
// wait on.  The words sit 128 B apart (their own L2 lines).  Only the first
// thread of a workgroup fences: the workgroup barrier before it has every
// wave's stores in its XCD's L2, which its release fence writes back; its
// acquire fence invalidates the L2 and the CU's L1 that the workgroup's
// waves read through after the closing workgroup barrier.  (Agent-scope
// fences write back and invalidate a whole XCD's L2, so one per workgroup
// and barrier rather than one per wave.)  A wait that outlasts maxSpins sets
// *error and returns, so a launch that could not make every workgroup
// resident ends instead of hanging (the host then fails the run).
constexpr int GRID_GROUPS = 64;
constexpr int BARRIER_STRIDE = 32;   // u32 words between counters
constexpr int BARRIER_WORDS = (GRID_GROUPS + 2) * BARRIER_STRIDE;

__device__ __forceinline__ void grid_barrier(uint32_t *bar, uint32_t *error, uint32_t maxSpins) {
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t *top = bar + GRID_GROUPS * BARRIER_STRIDE;
        uint32_t *gen = top + BARRIER_STRIDE;
        const uint32_t g = blockIdx.x % GRID_GROUPS;
        const uint32_t groups = gridDim.x < GRID_GROUPS ? gridDim.x : GRID_GROUPS;
        const uint32_t members = gridDim.x / GRID_GROUPS + (g < gridDim.x % GRID_GROUPS ? 1u : 0u);
        // the generation cannot advance before this workgroup arrives
        const uint32_t seen = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (atomicAdd(bar + g * BARRIER_STRIDE, 1u) == members - 1) {
            __hip_atomic_store(bar + g * BARRIER_STRIDE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == groups - 1) {
                __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        uint32_t spins = 0;
        while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == seen) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > maxSpins) {
                atomicOr(error, 1u);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Tile of the k-th visit of workgroup b in a grid of G workgroups (G a
// multiple of 8) over T tiles: visit v = b + k G is dispatched on XCD v % 8,
// and each XCD's visits cover one contiguous range of tiles (the numbering
// of xcd_block_index over the T visits).
__device__ __forceinline__ int64_t xcd_tile(int64_t v, int64_t T) {
    const int64_t q = T / 8, rem = T % 8, xcd = v % 8, local = v / 8;
    return (xcd < rem) ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

struct LoopArgs {
    int32_t steps;        // steps to run; fewer when one trips the exact test
    uint32_t *barrier;    // BARRIER_WORDS, zero before the first launch
    uint32_t *error;      // set when a barrier wait timed out
    uint32_t maxSpins;
};

// The arguments of one step of k_step_loop.  Two sets sit in device memory:
// the first step's and the one with the ping-pong roles swapped, and step s
// reads set s % 2.  Loading a field from the set of the step's parity at its
// point of use keeps it out of the registers between uses; roles swapped in
// registers stay live across the whole loop and spill (and two sets passed
// as kernel arguments and selected by parity are copied to scratch).
template <class R, class CT>
struct LoopSet {
    ForceArgs<R, CT, R> f;
    IntegArgs g;
    StepRecs<R> r;
};

// Several fused steps in one co-resident launch (HDM_REBUILD_LEARNED with
// hdm_exec_params.persistent): every workgroup runs the particles of its
// tiles through step_particle, the grid synchronises, and the next step
// starts on the other argument set -- unless the step just done tripped the
// exact test (its output trip word, read after the barrier by every
// workgroup alike), in which case the launch ends there as the guarded
// per-step launches would have ended.  Same arithmetic per particle as
// k_step, so the same bits.
template <class R, class CT, int BLOCK>
__global__ void __launch_bounds__(BLOCK, step_waves<R>(false)) k_step_loop(const LoopSet<R, CT> *__restrict__ sets,
                                                                            LoopArgs la) {
    const int64_t T = (sets[0].f.n + BLOCK - 1) / BLOCK;
    for (int32_t s = 0; s < la.steps; s++) {
        const LoopSet<R, CT> &a = sets[s & 1];
        if (__hip_atomic_load(a.g.guard, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) *a.g.guardOut = 1u;
            return;
        }
        uint32_t broke = 0;
        bool trip = false;
        for (int64_t v = blockIdx.x; v < T; v += gridDim.x) {
            const int64_t p = xcd_tile(v, T) * BLOCK + threadIdx.x;
            if (p < a.f.n) {
                double cordN[3], diffN[3];
                step_particle<R, CT>(a.f, a.g, a.r, p, cordN, diffN, broke);
                trip |= guard_trips(a.g, p, cordN);
            }
        }
        if (trip) *a.g.guardOut = 1u;
        if (a.f.brokenCount) {
            broke = wave_sum_u32(broke);
            if ((threadIdx.x & 63) == 0 && broke)
                atomicAdd((unsigned long long *)a.f.brokenCount, (unsigned long long)broke);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.g.stepsDone, 1u);
        grid_barrier(la.barrier, la.error, la.maxSpins);
        if (__hip_atomic_load(la.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    }
}

// Stores the two argument sets of a k_step_loop launch.
template <class R, class CT>
__global__ void k_store_loop_sets(LoopSet<R, CT> even, LoopSet<R, CT> odd, LoopSet<R, CT> *out) {
    if (threadIdx.x == 0) {
        out[0] = even;
        out[1] = odd;
    }
}

// The argument set of the step after `a`: the integrated state becomes the
// input, the trip words change roles.
template <class R, class CT>
LoopSet<R, CT> swapped_roles(const LoopSet<R, CT> &a) {
    LoopSet<R, CT> b = a;
    b.r.rec = a.r.recOut;
    b.r.recOut = const_cast<typename StepRecs<R>::Rec *>(a.r.rec);
    b.r.drec = a.r.drecOut;
    b.r.drecOut = const_cast<double *>(a.r.drec);
    b.r.tail = a.r.tailOut;
    b.r.tailOut = const_cast<MixedTail *>(a.r.tail);
    b.g.tx = a.g.otx;
    b.g.ty = a.g.oty;
    b.g.tz = a.g.otz;
    b.g.otx = const_cast<double *>(a.g.tx);
    b.g.oty = const_cast<double *>(a.g.ty);
    b.g.otz = const_cast<double *>(a.g.tz);
    b.f.tx = b.g.tx;
    b.f.ty = b.g.ty;
    b.f.tz = b.g.tz;
    b.g.guard = a.g.guardOut;
    b.g.guardOut = const_cast<uint32_t *>(a.g.guard);
    return b;
}

// SoA coordinates and Diff from the records (sync_state).
template <class Rec>
__global__ void __launch_bounds__(TPB) k_unpack_records(const Rec *__restrict__ rec,
                                                         const double *__restrict__ drec,
                                                         const MixedTail *__restrict__ tail, int64_t n, double *x,
                                                         double *y, double *z, double *d0, double *d1, double *d2,
                                                         double *d3, double *d4, double *d5) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const Rec P = rec[p];
    x[p] = P.x;
    y[p] = P.y;
    z[p] = P.z;
    double v[6];
    if constexpr (std::is_same<Rec, MixedRec>::value) {
        join_diff6(P, tail[p], drec, p, v);
    } else {
#pragma unroll
        for (int c = 0; c < 6; c++) v[c] = drec[p * 8 + c];
    }
    d0[p] = v[0];
    d1[p] = v[1];
    d2[p] = v[2];
    d3[p] = v[3];
    d4[p] = v[4];
    d5[p] = v[5];
}

// Gather records of the current SoA state (hdm_run entry).
template <class R>
__global__ void __launch_bounds__(TPB) k_pack_records(const double *__restrict__ x, const double *__restrict__ y,
                                                       const double *__restrict__ z, const double *__restrict__ tx,
                                                       const double *__restrict__ ty, const double *__restrict__ tz,
                                                       const double *__restrict__ d0, const double *__restrict__ d1,
                                                       const double *__restrict__ d2, const double *__restrict__ d3,
                                                       const double *__restrict__ d4, const double *__restrict__ d5,
                                                       const double *__restrict__ radius, int64_t n,
                                                       StepRecs<R> r) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= n) return;
    const double cord[3] = {x[p], y[p], z[p]};
    const double theta[3] = {tx[p], ty[p], tz[p]};
    const double diff[6] = {d0[p], d1[p], d2[p], d3[p], d4[p], d5[p]};
    store_records(r, p, cord, theta, diff, radius[p]);
}


// ---- link break --------------------------------------------------------------
struct BreakArgs {
    const double *x, *y, *z, *tx, *ty, *tz;
    const int32_t *A, *B;
    uint8_t *state;
    int32_t *slotLink, *slotOther;  // broken flag mirrored into both of the link's slots
    const int32_t *sidePos;
    const double *initialLen, *maxLoad;
    int64_t L;
    double bendBreak, scalarMaxLoad;
    uint64_t *counter;
};

// CombinedAccel::iterateLinks (CombinedAccel.h:200-235) applying
// HiDEMDefLinkBreakFunc (HiDEMDefaultFunctions.h:217-241); the strain is FP64
// in every precision form.
__global__ void __launch_bounds__(TPB) k_break(BreakArgs a) {
    const int64_t l = (int64_t)blockIdx.x * TPB + threadIdx.x;
    uint32_t broke = 0;
    if (l < a.L && a.state[l] != 0) {
        const int32_t i = a.A[l], j = a.B[l];
        double xDist = a.x[i] - a.x[j];
        double yDist = a.y[i] - a.y[j];
        double zDist = a.z[i] - a.z[j];
        double stretchLength = sqrt((xDist * xDist) + (yDist * yDist) + (zDist * zDist));
        double bx = fabs(a.tx[i] - a.tx[j]);
        double by = fabs(a.ty[i] - a.ty[j]);
        double bz = fabs(a.tz[i] - a.tz[j]);
        double inner = (by < bz) ? bz : by;
        double maxBend = (bx < inner) ? inner : bx;
        double initialLen = a.initialLen[l];
        double load = (stretchLength - initialLen) / initialLen + a.bendBreak * maxBend;
        double maxLoad = a.maxLoad ? a.maxLoad[l] : a.scalarMaxLoad;
        if (load >= maxLoad) {
            a.state[l] = 0;
            const int32_t s0 = a.sidePos[2 * l], s1 = a.sidePos[2 * l + 1];
            a.slotLink[s0] |= SLOT_BROKEN;
            a.slotLink[s1] |= SLOT_BROKEN;
            a.slotOther[s0] |= SLOT_BROKEN;
            a.slotOther[s1] |= SLOT_BROKEN;
            broke = 1;
        }
    }
    broke = wave_sum_u32(broke);
    if ((threadIdx.x & 63) == 0 && broke && a.counter)
        atomicAdd((unsigned long long *)a.counter, (unsigned long long)broke);
}

// ---- skin reductions -----------------------------------------------------------
// HiDEMLib::getMaxDistOwned (HiDEMLib.h:424-461): max over particles of the
// signed Diff, starting from 0.0.
__global__ void __launch_bounds__(TPB) k_skin_heuristic(const double *__restrict__ dx, const double *__restrict__ dy,
                                                         const double *__restrict__ dz, int64_t n, uint64_t *out) {
    double m[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        double v[3] = {dx[i], dy[i], dz[i]};
#pragma unroll
        for (int c = 0; c < 3; c++) m[c] = v[c] > m[c] ? v[c] : m[c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double w = wave_max(m[c]);
        if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long *)&out[c], (unsigned long long)ordered_key(w));
    }
}

// Rebuild snapshot: the positions in FP64 and rounded to FP32.
__global__ void __launch_bounds__(TPB) k_snapshot(const double *__restrict__ x, const double *__restrict__ y,
                                                   const double *__restrict__ z, int64_t n, double *__restrict__ x0,
                                                   double *__restrict__ y0, double *__restrict__ z0,
                                                   float *__restrict__ x0f, float *__restrict__ y0f,
                                                   float *__restrict__ z0f) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const double a = x[i], b = y[i], c = z[i];
    x0[i] = a;
    y0[i] = b;
    z0[i] = c;
    x0f[i] = (float)a;
    y0f[i] = (float)b;
    z0f[i] = (float)c;
}

// Exact test: max over particles of |x - x0|^2 (FP64).
__global__ void __launch_bounds__(TPB) k_skin_exact(const double *__restrict__ x, const double *__restrict__ y,
                                                     const double *__restrict__ z, const double *__restrict__ x0,
                                                     const double *__restrict__ y0, const double *__restrict__ z0,
                                                     int64_t n, uint64_t *out) {
    double m2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
        double ddx = x[i] - x0[i], ddy = y[i] - y0[i], ddz = z[i] - z0[i];
        double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        m2 = d2 > m2 ? d2 : m2;
    }
    m2 = wave_max(m2);
    if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long *)out, (unsigned long long)ordered_key(m2));
}

inline int blocks_for(int64_t n) { return (int)((n + TPB - 1) / TPB); }
inline int reduce_grid(const Context &c, int64_t n) {
    int64_t nb = blocks_for(n);
    int cap = c.numCU > 0 ? c.numCU * 4 : 1024;
    int g = (int)(nb < cap ? nb : cap);
    return g < 1 ? 1 : g;
}

// State pointers for the integrator: in place (alt == nullptr) or from the
// current buffers into alt[0..11] (x, y, z, xTheta.., Diff0..5).
IntegArgs integ_args(Context &c, const hdm_integ_params &ip, uint64_t *skin, bool withDisplacement, DevBuf *alt) {
    IntegArgs a;
    a.x = c.x.as<double>(); a.y = c.y.as<double>(); a.z = c.z.as<double>();
    a.tx = c.tx.as<double>(); a.ty = c.ty.as<double>(); a.tz = c.tz.as<double>();
    a.d0 = c.diff[0].as<double>(); a.d1 = c.diff[1].as<double>(); a.d2 = c.diff[2].as<double>();
    a.d3 = c.diff[3].as<double>(); a.d4 = c.diff[4].as<double>(); a.d5 = c.diff[5].as<double>();
    double *o[12];
    const double *in[12] = {a.x, a.y, a.z, a.tx, a.ty, a.tz, a.d0, a.d1, a.d2, a.d3, a.d4, a.d5};
    for (int k = 0; k < 12; k++) o[k] = alt ? alt[k].as<double>() : const_cast<double *>(in[k]);
    a.ox = o[0]; a.oy = o[1]; a.oz = o[2]; a.otx = o[3]; a.oty = o[4]; a.otz = o[5];
    a.od0 = o[6]; a.od1 = o[7]; a.od2 = o[8]; a.od3 = o[9]; a.od4 = o[10]; a.od5 = o[11];
    a.radius = c.radius.as<double>();
    a.type = c.type.as<uint8_t>();
    a.n = c.n;
    a.ip = ip;
    a.partials = skin ? c.skinPartials.as<double>() : nullptr;
    a.guard = nullptr;
    a.guardOut = nullptr;
    a.stepsDone = nullptr;
    a.guardThr = 0.0;
    a.x0 = withDisplacement ? c.x0.as<double>() : nullptr;
    a.y0 = withDisplacement ? c.y0.as<double>() : nullptr;
    a.z0 = withDisplacement ? c.z0.as<double>() : nullptr;
    a.x0f = withDisplacement ? c.x0f.as<float>() : nullptr;
    a.y0f = withDisplacement ? c.y0f.as<float>() : nullptr;
    a.z0f = withDisplacement ? c.z0f.as<float>() : nullptr;
    a.blockId = c.nBlocks > 0 ? c.blockId.as<int32_t>() : nullptr;
    a.blockWater = c.nBlocks > 0 ? c.blockWater.as<double>() : nullptr;
    a.mapMask = ip.variant == HDM_INTEG_GLACIER ? c.mapMask : 0;
    a.maps = a.mapMask ? c.mapDesc.as<MapDev>() : nullptr;
    return a;
}

template <class R, class CT, class FO>
ForceArgs<R, CT, FO> force_args(Context &c, const hdm_force_params &fp) {
    ForceArgs<R, CT, FO> a;
    a.x = c.x.as<double>(); a.y = c.y.as<double>(); a.z = c.z.as<double>();
    a.d0 = c.diff[0].as<double>(); a.d1 = c.diff[1].as<double>(); a.d2 = c.diff[2].as<double>();
    a.d3 = c.diff[3].as<double>(); a.d4 = c.diff[4].as<double>(); a.d5 = c.diff[5].as<double>();
    a.radius = c.radius.as<double>();
    a.n = c.n;
    a.nbrCount = c.nbrCount.as<int32_t>(); a.nbr = c.nbr.as<int32_t>(); a.nbrWidth = c.maxNeigh;
    a.linkCount = c.linkCount.as<uint32_t>(); a.slotLink = c.slotLink.as<int32_t>();
    a.slotOther = c.slotOther.as<int32_t>(); a.linkWidth = c.maxLinks;
    a.ct = c.ct.as<CT>();
    a.f0 = c.force[0].as<FO>(); a.f1 = c.force[1].as<FO>(); a.f2 = c.force[2].as<FO>();
    a.f3 = c.force[3].as<FO>(); a.f4 = c.force[4].as<FO>(); a.f5 = c.force[5].as<FO>();
    a.fp = cast_params<R>(fp);
    a.breakCheck = 0;
    a.linkedContacts = c.nbrExcludesLinked ? 1 : 0;
    a.slotLinkW = c.slotLink.as<int32_t>();
    a.slotOtherW = c.slotOther.as<int32_t>();
    a.tx = c.tx.as<double>(); a.ty = c.ty.as<double>(); a.tz = c.tz.as<double>();
    a.initialLen = c.initialLen.as<double>();
    a.maxLoad = c.hasMaxLoad ? c.maxLoad.as<double>() : nullptr;
    a.state = c.state.as<uint8_t>();
    a.bendBreak = 0;
    a.scalarMaxLoad = 0;
    a.brokenCount = c.brokenCount.as<uint64_t>();
    return a;
}

template <class R, class CT, class FO>
hipError_t force_impl(Context &c, const hdm_force_params &fp) {
    hipLaunchKernelGGL((k_force<R, CT, FO>), dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, force_args<R, CT, FO>(c, fp));
    return hipGetLastError();
}

constexpr int FINALIZE_BLOCKS = 32;

// Launch shapes of the fused step the autotuner chooses from.
struct StepShape {
    int block;
    bool xcd, occ;
    const char *name;
};
constexpr StepShape kShapes[STEP_VARIANTS] = {
    {256, true, false, "wg256-xcd"},      {128, true, false, "wg128-xcd"},   {512, true, false, "wg512-xcd"},
    {256, false, false, "wg256-linear"},  {128, true, true, "wg128-xcd-occ"}, {256, true, true, "wg256-xcd-occ"},
    {64, true, false, "wg64-xcd"}};
constexpr int kMinBlock = 64;

int64_t rows_for(int64_t n, int block) { return (n + block - 1) / block; }

// Rows for the rebuild-test partials (plus the finalize stage), and the guard
// words (zeroed once; the folding workgroup re-arms the ticket).
hipError_t ensure_partials(Context &c) {
    hipError_t e = c.skinPartials.ensure(sizeof(double) * 4 * ((size_t)rows_for(c.n, kMinBlock) + FINALIZE_BLOCKS));
    if (e != hipSuccess || c.guardState.ptr) return e;
    e = c.guardState.ensure(sizeof(uint32_t) * GUARD_WORDS);
    if (e != hipSuccess) return e;
    return launch_fill_u32(c.stream, c.guardState.as<uint32_t>(), 0u, GUARD_WORDS);
}

// k_skin_finalize over the rows of a launch of rows_for(n, block) workgroups.
hipError_t skin_finalize(Context &c, uint64_t *skin, int block = TPB) {
    FinalizeArgs f;
    f.rows = rows_for(c.n, block);
    f.partials = c.skinPartials.as<double>();
    f.stage = c.skinPartials.as<double>() + 4 * f.rows;
    f.ticket = c.guardState.as<uint32_t>() + GUARD_TICKET;
    f.skin = skin;
    hipLaunchKernelGGL(k_skin_finalize, dim3(FINALIZE_BLOCKS), dim3(TPB), 0, c.stream, f);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_guard_reset(Context &c) {
    hipError_t e = ensure_partials(c);
    if (e != hipSuccess) return e;
    return launch_fill_u32(c.stream, c.guardState.as<uint32_t>() + GUARD_FLAG, 0u, 3);
}

hipError_t launch_force(Context &c, const hdm_force_params &fp) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    if (c.n == 0) return hipSuccess;
    if (c.prec == HDM_PREC_FP64) return force_impl<double, double, double>(c, fp);
    return force_impl<float, float, float>(c, fp);
}

hipError_t launch_integrate(Context &c, const hdm_integ_params &ip, uint64_t *skin, bool withDisplacement) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    if (c.n == 0) return hipSuccess;
    if (skin) {
        hipError_t e = ensure_partials(c);
        if (e != hipSuccess) return e;
    }
    IntegArgs a = integ_args(c, ip, skin, withDisplacement, nullptr);
    if (c.prec == HDM_PREC_FP64)
        hipLaunchKernelGGL(k_integrate<double>, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, a,
                           c.force[0].as<double>(), c.force[1].as<double>(), c.force[2].as<double>(),
                           c.force[3].as<double>(), c.force[4].as<double>(), c.force[5].as<double>());
    else
        hipLaunchKernelGGL(k_integrate<float>, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, a,
                           c.force[0].as<float>(), c.force[1].as<float>(), c.force[2].as<float>(),
                           c.force[3].as<float>(), c.force[4].as<float>(), c.force[5].as<float>());
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !skin) return e;
    return skin_finalize(c, skin);
}

namespace {

// Bound on every Diff component under which a pair beyond the attraction
// range provably adds +-0 (contact_is_inert): the damping terms then stay
// below 1e300 (FP64) / 1e37 (FP32) in magnitude.  -1: no Diff qualifies.
double tame_bound(const Context &c, const hdm_force_params &fp) {
    if (!inert_skip_allowed(fp)) return -1.0;  // no Diff is tame: no pair is skipped
    double k = 1.0;
    const double a = fabs(fp.damping * fp.invDt), b = fabs(fp.dampingRotation * fp.invDt);
    k = a > k ? a : k;
    k = b > k ? b : k;
    if (!(k < 1e30)) return -1.0;  // also NaN
    return (c.prec == HDM_PREC_FP64 ? 1e299 : 1e36) / (4.0 * k);
}

template <class R>
StepRecs<R> step_recs(Context &c, double tame, const hdm_break_params *brk) {
    StepRecs<R> r;
    using Rec = typename StepRecs<R>::Rec;
    r.rec = c.rec.as<Rec>();
    r.drec = c.drec.as<double>();
    r.tail = c.dtail.as<MixedTail>();
    r.recOut = c.recAlt.as<Rec>();
    r.drecOut = c.drecAlt.as<double>();
    r.tailOut = c.dtailAlt.as<MixedTail>();
    r.tame = tame;
    r.slotThr = c.slotThr.as<uint16_t>();
    r.thrScale = c.thrScale;
    r.l0Max = c.l0Max;
    r.perLinkMaxLoad = c.hasMaxLoad ? 1 : 0;
    r.scalarMaxLoad = brk ? (float)brk->maxLoad : 0.0f;
    r.bendBreak = brk ? (float)brk->bendBreak : 0.0f;
    r.prefilter = (brk && brk->bendBreak >= 0.0 && brk->bendBreak < 1e30) ? 1 : 0;
    r.storeForce = 1;
    return r;
}

hipError_t ensure_records(Context &c) {
    const size_t n = (size_t)(c.n > 0 ? c.n : 1);
    const size_t recBytes = c.prec == HDM_PREC_FP64 ? sizeof(double4) : sizeof(MixedRec);
    for (DevBuf *b : {&c.rec, &c.recAlt}) {
        hipError_t e = b->ensure(recBytes * n);
        if (e != hipSuccess) return e;
    }
    for (DevBuf *b : {&c.drec, &c.drecAlt}) {
        hipError_t e = b->ensure(sizeof(double) * (size_t)drec_stride(c.prec) * n);
        if (e != hipSuccess) return e;
    }
    if (c.prec != HDM_PREC_FP64) {
        for (DevBuf *b : {&c.dtail, &c.dtailAlt}) {
            hipError_t e = b->ensure(sizeof(MixedTail) * n);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

}  // namespace

hipError_t launch_pack_records(Context &c, double tame) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    hipError_t e = ensure_records(c);
    if (e != hipSuccess || c.n == 0) return e;
    // pack writes the "out" side; swap so the packed records are current
    swap_buf(c.rec, c.recAlt);
    swap_buf(c.drec, c.drecAlt);
    swap_buf(c.dtail, c.dtailAlt);
    auto go = [&](auto r) {
        hipLaunchKernelGGL(k_pack_records, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                           c.y.as<double>(), c.z.as<double>(), c.tx.as<double>(), c.ty.as<double>(),
                           c.tz.as<double>(), c.diff[0].as<double>(), c.diff[1].as<double>(), c.diff[2].as<double>(),
                           c.diff[3].as<double>(), c.diff[4].as<double>(), c.diff[5].as<double>(),
                           c.radius.as<double>(), c.n, r);
    };
    if (c.prec == HDM_PREC_FP64) {
        StepRecs<double> r = step_recs<double>(c, tame, nullptr);
        r.recOut = c.rec.as<double4>();
        r.drecOut = c.drec.as<double>();
        go(r);
    } else {
        StepRecs<float> r = step_recs<float>(c, tame, nullptr);
        r.recOut = c.rec.as<MixedRec>();
        r.drecOut = c.drec.as<double>();
        r.tailOut = c.dtail.as<MixedTail>();
        go(r);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    c.recCurrent = true;
    c.recTame = tame;
    return hipSuccess;
}

void swap_step_state(Context &c) {
    c.guardCur ^= 1;
    swap_buf(c.tx, c.alt[3]);
    swap_buf(c.ty, c.alt[4]);
    swap_buf(c.tz, c.alt[5]);
    swap_buf(c.rec, c.recAlt);
    swap_buf(c.drec, c.drecAlt);
    swap_buf(c.dtail, c.dtailAlt);
}

hipError_t sync_state(Context &c) {
    if (!c.soaStale) return hipSuccess;
    c.soaStale = false;
    if (c.n == 0) return hipSuccess;
    auto go = [&](auto *rec) {
        using Rec = typename std::remove_pointer<decltype(rec)>::type;
        hipLaunchKernelGGL((k_unpack_records<Rec>), dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, rec,
                           c.drec.as<double>(), c.dtail.as<MixedTail>(), c.n, c.x.as<double>(), c.y.as<double>(), c.z.as<double>(),
                           c.diff[0].as<double>(), c.diff[1].as<double>(), c.diff[2].as<double>(),
                           c.diff[3].as<double>(), c.diff[4].as<double>(), c.diff[5].as<double>());
    };
    if (c.prec == HDM_PREC_FP64) go(c.rec.as<double4>());
    else go(c.rec.as<MixedRec>());
    return hipGetLastError();
}

hipError_t launch_step(Context &c, const hdm_force_params &fp, const hdm_integ_params &ip, uint64_t *skin,
                       bool withDisplacement, const hdm_break_params *brk, bool storeForce, const StepGuard *guard) {
    if (c.n == 0) return hipSuccess;
    const size_t b = sizeof(double) * (size_t)c.n;
    for (int k = 3; k < 6; k++) {
        hipError_t e = c.alt[k].ensure(b);
        if (e != hipSuccess) return e;
    }
    if (skin) {
        hipError_t e = ensure_partials(c);
        if (e != hipSuccess) return e;
    }
    const double tame = tame_bound(c, fp);
    if (!c.recCurrent || c.recTame != tame) {
        hipError_t e = launch_pack_records(c, tame);
        if (e != hipSuccess) return e;
    }
    IntegArgs g = integ_args(c, ip, skin, withDisplacement, c.alt);
    if (guard) {
        if (!skin || !withDisplacement) return hipErrorInvalidValue;
        uint32_t *words = c.guardState.as<uint32_t>();
        g.guard = words + GUARD_FLAG + c.guardCur;
        g.guardOut = words + GUARD_FLAG + (c.guardCur ^ 1);
        g.stepsDone = words + GUARD_STEPS;
        g.guardThr = guard->threshold;
        g.partials = nullptr;   // the trip words replace the maxima
    }
    const int v = (c.stepVariant >= 0 && c.stepVariant < STEP_VARIANTS) ? c.stepVariant : 0;
    const StepShape shape = kShapes[v];
    auto launch = [&](auto f, auto r) {
        r.storeForce = storeForce ? 1 : 0;
        if (brk && c.L > 0) {
            f.breakCheck = 1;
            f.bendBreak = brk->bendBreak;
            f.scalarMaxLoad = brk->maxLoad;
        }
        using A = decltype(f);
        using R = typename std::conditional<std::is_same<A, ForceArgs<double, double, double>>::value, double,
                                            float>::type;
        const dim3 grid((unsigned)rows_for(c.n, shape.block));
        switch (v) {
        case 0: hipLaunchKernelGGL((k_step<R, R, 256, true, false>), grid, dim3(256), 0, c.stream, f, g, r); break;
        case 1: hipLaunchKernelGGL((k_step<R, R, 128, true, false>), grid, dim3(128), 0, c.stream, f, g, r); break;
        case 2: hipLaunchKernelGGL((k_step<R, R, 512, true, false>), grid, dim3(512), 0, c.stream, f, g, r); break;
        case 3: hipLaunchKernelGGL((k_step<R, R, 256, false, false>), grid, dim3(256), 0, c.stream, f, g, r); break;
        case 4: hipLaunchKernelGGL((k_step<R, R, 128, true, true>), grid, dim3(128), 0, c.stream, f, g, r); break;
        case 5: hipLaunchKernelGGL((k_step<R, R, 256, true, true>), grid, dim3(256), 0, c.stream, f, g, r); break;
        default: hipLaunchKernelGGL((k_step<R, R, 64, true, false>), grid, dim3(64), 0, c.stream, f, g, r); break;
        }
    };
    if (c.prec == HDM_PREC_FP64) launch(force_args<double, double, double>(c, fp), step_recs<double>(c, tame, brk));
    else launch(force_args<float, float, float>(c, fp), step_recs<float>(c, tame, brk));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the new state becomes current (records and rotations); the SoA
    // coordinates and Diff are unpacked on demand
    swap_step_state(c);
    c.soaStale = true;
    if (!skin || guard) return hipSuccess;
    return skin_finalize(c, skin, shape.block);
}

hipError_t launch_break_records(Context &c, const hdm_break_params &bp, uint64_t *counter) {
    if (counter) {
        hipError_t e = launch_fill_u64(c.stream, counter, 0ull, 1);
        if (e != hipSuccess) return e;
    }
    if (c.L == 0 || c.n == 0) return hipSuccess;
    if (!c.recCurrent) return launch_break(c, bp, counter);
    hdm_force_params fp{};
    auto go = [&](auto f, auto r) {
        using A = decltype(f);
        using R = typename std::conditional<std::is_same<A, ForceArgs<double, double, double>>::value, double,
                                            float>::type;
        f.bendBreak = bp.bendBreak;
        f.scalarMaxLoad = bp.maxLoad;
        if constexpr (sizeof(R) == 8) {
            const int64_t nSlots = (c.n + 63) / 64 * (int64_t)f.linkWidth * 64;
            hipLaunchKernelGGL((k_break_slots<R>), dim3((unsigned)((nSlots + TPB - 1) / TPB)), dim3(TPB), 0,
                               c.stream, f, r, nSlots, counter);
        } else {
            hipLaunchKernelGGL((k_break_recs<R>), dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, f, r, counter);
        }
    };
    if (c.prec == HDM_PREC_FP64) go(force_args<double, double, double>(c, fp), step_recs<double>(c, c.recTame, &bp));
    else go(force_args<float, float, float>(c, fp), step_recs<float>(c, c.recTame, &bp));
    return hipGetLastError();
}

hipError_t launch_step_loop(Context &c, const hdm_force_params &fp, const hdm_integ_params &ip, uint64_t *skin,
                            const hdm_break_params &brk, const StepGuard &guard, int32_t steps, bool *barrierFailed) {
    *barrierFailed = false;
    if (c.n == 0 || steps <= 0) return hipSuccess;
    if (!c.coopLaunch) return hipErrorNotSupported;
    const size_t b = sizeof(double) * (size_t)c.n;
    for (int k = 3; k < 6; k++) {
        hipError_t e = c.alt[k].ensure(b);
        if (e != hipSuccess) return e;
    }
    hipError_t e = ensure_partials(c);
    if (e != hipSuccess) return e;
    if (!c.loopBarrier.ptr) {
        e = c.loopBarrier.ensure(sizeof(uint32_t) * (BARRIER_WORDS + 1));
        if (e == hipSuccess) e = launch_fill_u32(c.stream, c.loopBarrier.as<uint32_t>(), 0u, BARRIER_WORDS + 1);
        if (e != hipSuccess) return e;
    }
    const double tame = tame_bound(c, fp);
    if (!c.recCurrent || c.recTame != tame) {
        e = launch_pack_records(c, tame);
        if (e != hipSuccess) return e;
    }
    IntegArgs g = integ_args(c, ip, skin, true, c.alt);
    uint32_t *words = c.guardState.as<uint32_t>();
    g.guard = words + GUARD_FLAG + c.guardCur;
    g.guardOut = words + GUARD_FLAG + (c.guardCur ^ 1);
    g.stepsDone = words + GUARD_STEPS;
    g.guardThr = guard.threshold;
    g.partials = nullptr;
    LoopArgs la;
    la.steps = steps;
    la.barrier = c.loopBarrier.as<uint32_t>();
    la.error = la.barrier + BARRIER_WORDS;
    la.maxSpins = 1u << 22;   // seconds of waiting; a step takes milliseconds
    auto go = [&](auto f, auto r, const void *kernel) -> hipError_t {
        r.storeForce = 0;
        f.breakCheck = c.L > 0 ? 1 : 0;
        f.bendBreak = brk.bendBreak;
        f.scalarMaxLoad = brk.maxLoad;
        int perCU = 0;
        hipError_t e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, TPB, 0);
        if (e2 != hipSuccess) return e2;
        const int64_t tiles = (c.n + TPB - 1) / TPB;
        int64_t grid = (int64_t)perCU * (c.numCU > 0 ? c.numCU : 1);
        if (grid > tiles) grid = tiles;
        grid -= grid % 8;
        if (grid < 8) return hipErrorInvalidConfiguration;
        using Real = typename decltype(r)::Real;
        using CTT = typename decltype(f)::CTType;
        using Set = LoopSet<Real, CTT>;
        e2 = c.loopSets.ensure(2 * sizeof(Set));
        if (e2 != hipSuccess) return e2;
        Set *sets = static_cast<Set *>(c.loopSets.ptr);
        const Set even{f, g, r};
        hipLaunchKernelGGL((k_store_loop_sets<Real, CTT>), dim3(1), dim3(64), 0, c.stream, even, swapped_roles(even),
                           sets);
        e2 = hipGetLastError();
        if (e2 != hipSuccess) return e2;
        void *args[] = {&sets, &la};
        return hipLaunchCooperativeKernel(kernel, dim3((unsigned)grid), dim3(TPB), args, 0, c.stream);
    };
    if (c.prec == HDM_PREC_FP64)
        e = go(force_args<double, double, double>(c, fp), step_recs<double>(c, tame, &brk),
               reinterpret_cast<const void *>(&k_step_loop<double, double, TPB>));
    else
        e = go(force_args<float, float, float>(c, fp), step_recs<float>(c, tame, &brk),
               reinterpret_cast<const void *>(&k_step_loop<float, float, TPB>));
    if (e != hipSuccess) return e;
    // the launch swapped the roles `steps` times (the host undoes the ones a
    // trip skipped, as for skipped per-step launches)
    if (steps & 1) swap_step_state(c);
    c.soaStale = true;
    uint32_t err = 0;
    e = hipMemcpyAsync(&err, la.error, sizeof(err), hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return e;
    if (err) {   // re-arm the barrier words for the next launch
        *barrierFailed = true;
        return launch_fill_u32(c.stream, c.loopBarrier.as<uint32_t>(), 0u, BARRIER_WORDS + 1);
    }
    return hipSuccess;
}

const char *step_variant_name(int v) { return (v >= 0 && v < STEP_VARIANTS) ? kShapes[v].name : "?"; }

// The fastest shape, except that a shape within 1% of it which renumbers
// workgroups per XCD wins the tie: equal time, but each XCD's L2 then serves
// its own particle range, so the step fetches far less from HBM (the linear
// order read 1.8x the bytes at the same speed on the 8M slab).
int pick_step_variant(const double ms[STEP_VARIANTS]) {
    int best = 0;
    for (int v = 1; v < STEP_VARIANTS; v++)
        if (ms[v] < ms[best]) best = v;
    if (kShapes[best].xcd) return best;
    int alt = -1;
    for (int v = 0; v < STEP_VARIANTS; v++)
        if (kShapes[v].xcd && ms[v] <= 1.01 * ms[best] && (alt < 0 || ms[v] < ms[alt])) alt = v;
    return alt >= 0 ? alt : best;
}

hipError_t launch_break(Context &c, const hdm_break_params &bp, uint64_t *counter) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    if (c.L == 0) return hipSuccess;
    BreakArgs a;
    a.x = c.x.as<double>(); a.y = c.y.as<double>(); a.z = c.z.as<double>();
    a.tx = c.tx.as<double>(); a.ty = c.ty.as<double>(); a.tz = c.tz.as<double>();
    a.A = c.la.as<int32_t>(); a.B = c.lb.as<int32_t>();
    a.state = c.state.as<uint8_t>();
    a.initialLen = c.initialLen.as<double>();
    a.maxLoad = c.hasMaxLoad ? c.maxLoad.as<double>() : nullptr;
    a.slotLink = c.slotLink.as<int32_t>();
    a.slotOther = c.slotOther.as<int32_t>();
    a.sidePos = c.sidePos.as<int32_t>();
    a.L = c.L;
    a.bendBreak = bp.bendBreak;
    a.scalarMaxLoad = bp.maxLoad;
    a.counter = counter;
    hipLaunchKernelGGL(k_break, dim3(blocks_for(c.L)), dim3(TPB), 0, c.stream, a);
    return hipGetLastError();
}

hipError_t launch_skin_heuristic(Context &c, uint64_t *out3) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    hipError_t e = launch_fill_u64(c.stream, out3, ordered_key(0.0), 3);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_skin_heuristic, dim3(reduce_grid(c, c.n)), dim3(TPB), 0, c.stream, c.diff[0].as<double>(),
                       c.diff[1].as<double>(), c.diff[2].as<double>(), c.n, out3);
    return hipGetLastError();
}

hipError_t launch_skin_exact(Context &c, uint64_t *out1) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    hipError_t e = launch_fill_u64(c.stream, out1, ordered_key(0.0), 1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_skin_exact, dim3(reduce_grid(c, c.n)), dim3(TPB), 0, c.stream, c.x.as<double>(),
                       c.y.as<double>(), c.z.as<double>(), c.x0.as<double>(), c.y0.as<double>(), c.z0.as<double>(),
                       c.n, out1);
    return hipGetLastError();
}

hipError_t launch_snapshot(Context &c) {
    if (hipError_t se = sync_state(c); se != hipSuccess) return se;
    if (c.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_snapshot, dim3(blocks_for(c.n)), dim3(TPB), 0, c.stream, c.x.as<double>(), c.y.as<double>(),
                       c.z.as<double>(), c.n, c.x0.as<double>(), c.y0.as<double>(), c.z0.as<double>(),
                       c.x0f.as<float>(), c.y0f.as<float>(), c.z0f.as<float>());
    return hipGetLastError();
}

}  // namespace hdm
