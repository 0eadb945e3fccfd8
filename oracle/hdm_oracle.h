/*
 * hdm_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU FP64 restatement of HiDEM's per-timestep hot path, used only as the
 * parity checker (tests/, __graft_entry__.smoke()) and as the timed CPU
 * baseline in bench.py.  Nothing in hidem_amd/ includes, links or calls this.
 *
 * Every function names the reference file:line whose semantics it restates
 * (paths relative to the frroberts/HiDEM tree).  Arithmetic is written in the
 * reference's expression order and compiled with -ffp-contract=off so the
 * rounding sequence matches the reference's FP64 build (SET_FP32=OFF,
 * SET_CT_FP32=OFF, no -ffast-math).
 *
 * Parity pinning (what checks each function independently of the product):
 *   hdmo_grid_build      reference NeighbourGridTester cases + closed form
 *   hdmo_neigh_build     O(N^2) brute force + FCC 12-neighbour + 26-stencil
 *   hdmo_link_lists      invariants (each link once per end, ascending)
 *   hdmo_force           closed forms: 2-particle Hertz contact, axial beam,
 *                        torsion beam, frame orthonormality, action/reaction
 *   hdmo_integrate_*     closed-form recurrences (free fall, terminal drift)
 *   hdmo_break           closed-form single-link threshold crossing
 *   hdmo_skin_*          closed form on constructed displacements
 * None of these functions is "parity unpinned".
 */
#ifndef HDM_ORACLE_H
#define HDM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDMO_TYPE_FIXED 1u

/* SoA particle state, HiDEMParticleData field set (HiDEMParticleData.h:20-39,
 * ParticleData.h:29-37), FP64 build: hpReal_t = lpReal_t = double. */
typedef struct hdmo_particles {
    int64_t n;
    double *x, *y, *z;            /* xCord, yCord, zCord            */
    double *tx, *ty, *tz;         /* xTheta, yTheta, zTheta         */
    double *dx, *dy, *dz;         /* xDiff, yDiff, zDiff            */
    double *du, *dv, *dw;         /* uDiff, vDiff, wDiff            */
    double *radius;
    double *fx, *fy, *fz, *fu, *fv, *fw; /* xForce .. wForce        */
    uint8_t *type;                /* bit 0: clamped (build-defined) */
} hdmo_particles;

/* SoA link state, HiDEMLinkData/LinkData (HiDEMLinkData.h:24-31,
 * LinkData.h:28-43).  Side data CT0..CT5 is kept in the layout that
 * LinkDataHandler::appendInitLinks (LinkDataHandler.h:229-236) creates:
 * side 2*l is end A, side 2*l+1 is end B.  The reference later permutes the
 * side arrays (LinkAccelSide::sortSideData); that is a storage permutation
 * only and does not change any value. */
typedef struct hdmo_links {
    int64_t n;
    int32_t *a, *b;               /* localPartIndA, localPartIndB   */
    uint8_t *state;
    double *initialLen;
    double *maxLoad;              /* per-link threshold (build-defined extension) */
    double *ct[6];                /* CT0_flat .. CT5_flat, 2*n entries each */
} hdmo_links;

typedef struct hdmo_grid {
    double spacing;
    double xMin, yMin, zMin, xMax, yMax, zMax;
    int32_t X, Y, Z;
    int64_t totalCells;
    int64_t dirOffsets[27];
    int64_t *colours;      /* n */
    int32_t *order;        /* iterationOrder, n */
    int64_t nOcc;          /* occupied cells (bucket count) */
    int64_t *occColour;    /* nOcc, ascending */
    int32_t *bucketStart;  /* nOcc */
    int32_t *bucketSize;   /* nOcc */
} hdmo_grid;

typedef struct hdmo_neigh {
    int64_t n;
    int32_t *count;        /* numNeighs, n */
    int64_t *start;        /* CSR start, n+1 */
    int32_t *list;         /* CSR entries, in the reference's visit order */
} hdmo_neigh;

typedef struct hdmo_linklists {
    int64_t n;
    int32_t *count;        /* linksPerParticle */
    int64_t *start;        /* n+1 */
    int32_t *linkId;       /* ascending per particle */
    int32_t *other;        /* otherPartLink */
    int32_t *side;         /* linkSideLind: 2*l (end A) or 2*l+1 (end B) */
} hdmo_linklists;

/* HiDEMDefNeighForceFunc + HiDEMDefLinkForceFunc parameters
 * (HiDEMDefaultFunctions.h:21-35, 90-103). */
typedef struct hdmo_force_params {
    double invDt, attractDistance, attractScaleForce, damping, dampingRotation, scaleE;
    double E, W2, W4, G;
} hdmo_force_params;

/* A 2-D field on a regular xPoints x yPoints grid over [xMin, xMax] x
 * [yMin, yMax] (Map2d, Map2d.h:22-61): data[ix + iy * xPoints]; xRes, yRes,
 * invXMaxMin, invYMaxMin as its constructor computes them (in float, see
 * hdmo_map_init); xOffset, yOffset as set by Map2d::setOffset. */
typedef struct hdmo_map2d {
    int32_t xPoints, yPoints;
    float xMin, yMin, xMax, yMax;
    float xRes, yRes, invXMaxMin, invYMaxMin;
    float xOffset, yOffset;
    const double *data;
} hdmo_map2d;

/* Map2d constructor arithmetic (Map2d.h:43-61) for the derived fields. */
void hdmo_map_init(hdmo_map2d *m);
/* Map2d::getIndexes (Map2d.h:92-108): xLow, yLow, xHigh, yHigh. */
void hdmo_map_indexes(const hdmo_map2d *m, float x, float y, int32_t out[4]);
/* Map2d::interp (Map2d.h:110-145), FP32 bilinear interpolation; the value
 * itself (not rounded to float) where both axes clamp to one point. */
double hdmo_map_interp(const hdmo_map2d *m, float x, float y);

/* GlacierDem map inputs, indexes into hdmo_integ_params.maps. */
enum { HDMO_MAP_BED = 0, HDMO_MAP_FRICTION = 1, HDMO_MAP_EXT = 2, HDMO_MAP_COUNT = 8 };

/* Time integration parameters: IceBlock applyForce (MainIceBlock.cpp:299-400)
 * and GlacierDem applyForce (MainDem.cpp:145-259). */
typedef struct hdmo_integ_params {
    int32_t variant;       /* 0 = IceBlock (+clamp), 1 = GlacierDem (analytic bed) */
    double rhoIce, g, lift, waterLevel, airDrag, waterDrag;
    double invDt, invDt2, dtSqr, invDtSqr, dt;
    /* GlacierDem only */
    double bedDrag, bedInteract, bedStiffness, frictionCap;
    double bedZ, bedXEnd, deepZ, frictionXEnd; /* analytic bed: z=bedZ for x<bedXEnd else deepZ; fric=1 for x<frictionXEnd */
    double ext[6];         /* constant external force x,y,z,u,v,w */
    double maxVel, maxRot;
    int32_t velocityLimit;
    int32_t bound[6];      /* xMin, xMax, yMin, yMax, zMin, zMax enabled */
    double lower[3], upper[3];
    /* batched form (build-defined): waterline of particle i is
     * blockWater[blockId[i]] when blockWater != NULL */
    const int32_t *blockId;
    const double *blockWater;
    /* GlacierDem map inputs (NULL: the analytic bed / friction above, ext[k]):
     * bed height, friction, external force x, y, z, u, v, w */
    const hdmo_map2d *maps[HDMO_MAP_COUNT];
} hdmo_integ_params;

/* HiDEMDefLinkBreakFunc (HiDEMDefaultFunctions.h:207-242). */
typedef struct hdmo_break_params {
    double bendBreak;
    double maxLoad;        /* used when links->maxLoad == NULL */
} hdmo_break_params;

/* ---- neighbour grid (NeighbourGrid.cpp:24-184, NeighbourGrid.h:85-105) ---- */
int  hdmo_grid_build(const hdmo_particles *p, double searchDistance, hdmo_grid *g);
void hdmo_grid_free(hdmo_grid *g);

/* ---- redistribution order (single rank) ----
 * HiDEMDistribution's sort (HiDEMDistribution.h:150-181): particles ordered by
 * their cell colours on four grids of spacing s, 2s, 4s, 8s (bounding box and
 * colours as NeighbourGrid::generateNeighbourGrid / colourParticles,
 * NeighbourGrid.cpp:72-161), coarsest first (ParallelSort::sortIndirectNLevel,
 * ParallelSort.h:84-89), ties by index (the reference's sort leaves them
 * unspecified); ParticleDataHandler::reorder (ParticleDataHandler.h:36-69).
 * Links ordered by their old (A, B) (LinkData::getNewOrder, LinkData.h:166-
 * 181), ties by link index, then renamed to the new particle indices
 * (reorderLinkTargets, :116-137); each link's two CT sides move with it.
 * newOrder[i] = old index of the new particle i, linkOrder[j] = old index of
 * the new link j (either may be NULL).  Returns 0, -1 for fewer than two
 * particles, -2 when out of memory. */
int  hdmo_reorder(hdmo_particles *p, hdmo_links *l, double searchDistance, int64_t *newOrder, int64_t *linkOrder);

/* ---- neighbour list (NeighAccel.h:25-137) ---- */
int  hdmo_neigh_build(const hdmo_particles *p, const hdmo_grid *g, double interactDist, hdmo_neigh *nl);
void hdmo_neigh_free(hdmo_neigh *nl);

/* ---- per-particle link lists (LinkAccelSide.h:43-157) ---- */
int  hdmo_link_lists(int64_t nParticles, const hdmo_links *l, hdmo_linklists *ll);
void hdmo_linklists_free(hdmo_linklists *ll);

/* ---- force gather (CombinedAccel.h:121-197 with the default functors) ---- */
void hdmo_force(hdmo_particles *p, hdmo_links *l, const hdmo_neigh *nl,
                const hdmo_linklists *ll, const hdmo_force_params *fp);

/* Single-pair functors, exposed for closed-form tests. */
void hdmo_contact_pair(double mask, double xDist, double yDist, double zDist,
                       const double self[7], const double other[7],
                       const hdmo_force_params *fp, double force[6]);
void hdmo_beam_side(double mask, double xDist, double yDist, double zDist,
                    const double self[7], const double other[7],
                    const hdmo_force_params *fp, double ct[6], double force[6]);

/* ---- time integration ---- */
void hdmo_integrate(hdmo_particles *p, const hdmo_integ_params *ip);

/* ---- link break (CombinedAccel.h:200-235 + HiDEMDefLinkBreakFunc) ---- */
int64_t hdmo_break(const hdmo_particles *p, hdmo_links *l, const hdmo_break_params *bp);

/* ---- rebuild policy ---- */
/* Reference heuristic (HiDEMLib.h:424-474): returns the stepSize-scaled
 * displacement estimate stepSize*sqrt(mx^2+my^2+mz^2), mx = max(0, max xDiff). */
double hdmo_skin_heuristic(const hdmo_particles *p, int stepSize);
/* Exact test (build-defined): max_i |x_i - x0_i| over all particles. */
double hdmo_skin_exact(const hdmo_particles *p, const double *x0, const double *y0, const double *z0);

/* ---- whole-step driver (MainIceBlock.cpp:409-447, HiDEMLib.h:533-628) ---- */
typedef struct hdmo_sim {
    hdmo_particles *p;
    hdmo_links *l;
    hdmo_force_params fp;
    hdmo_integ_params ip;
    hdmo_break_params bp;
    double searchDistance, interactDist, dt, breakTime;
    int32_t policy;          /* 0 = reference heuristic, 1 = exact displacement, 2 = fixed interval */
    int32_t regenInterval;   /* policy 2 */
    int64_t ts;
    double accumulatedDist;
    int64_t rebuilds;
    int64_t broken;
    /* internal */
    hdmo_grid grid;
    hdmo_neigh nl;
    hdmo_linklists ll;
    double *x0, *y0, *z0;
    int32_t built;
} hdmo_sim;

int  hdmo_sim_init(hdmo_sim *s);
/* Runs nSteps full steps; records, for each step, whether a rebuild ran
 * (rebuildLog may be NULL) and the links broken in that step as
 * (step, link) pairs appended to brokenLog (capacity brokenCap, may be NULL).
 * Returns the number of broken-link records written. */
int64_t hdmo_sim_run(hdmo_sim *s, int64_t nSteps, uint8_t *rebuildLog,
                     int64_t *brokenLog, int64_t brokenCap);
void hdmo_sim_free(hdmo_sim *s);

int hdmo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
