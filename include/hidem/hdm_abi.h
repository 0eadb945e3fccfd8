/*
 * hdm_abi.h -- C ABI of the MI355X (gfx950) HiDEM timestep hot path.
 *
 * Plain C: POD structs, raw pointers and sizes, int status codes.  A context
 * (hdm_ctx) is the accelerator object that HiDEMLib selects at
 * HiDEMLib.h:54-60 (CombinedAccel / CombinedAccelSmem); it owns device
 * mirrors of the HiDEMParticleData / HiDEMLinkData fields the step touches,
 * the neighbour grid, the neighbour list and the per-particle link lists.
 * All work is ordered on one HIP stream owned by the context; no call
 * performs an implicit device-wide synchronisation unless it returns data to
 * the host.  A context is not re-entrant: one host thread at a time.
 *
 * Every function returns HDM_OK (0) or a negative HDM_E_* code; the message
 * of the last failure is available from hdm_last_error().  The C++ wrapper
 * (CombinedAccelHip.h) maps failures to the reference's convention (message,
 * then exit(1); DevCall.h:69-91).
 *
 * Precision forms (hdm_ctx_create):
 *   HDM_PREC_FP64   every quantity in FP64 (reference build with
 *                   SET_FP32=OFF, SET_CT_FP32=OFF).
 *   HDM_PREC_MIXED  contact/beam arithmetic, CT side state and forces in
 *                   FP32; coordinates, rotations, per-step displacements
 *                   (Diff), the integrator and the break strain in FP64.
 * Host-side structs always carry FP64 values; the context converts.
 */
#ifndef HIDEM_HDM_ABI_H
#define HIDEM_HDM_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDM_ABI_VERSION 1

enum {
    HDM_OK = 0,
    HDM_E_INVALID = -1,     /* bad argument or call order */
    HDM_E_HIP = -2,         /* HIP runtime error */
    HDM_E_NOMEM = -3,       /* device or host allocation failed */
    HDM_E_NODEVICE = -4,    /* no usable gfx950 device */
    HDM_E_CAPACITY = -5,    /* a size exceeds what the structures allow */
    HDM_E_STATE = -6        /* data not uploaded / structures not built */
};

enum { HDM_PREC_FP64 = 0, HDM_PREC_MIXED = 1 };

/* Particle type bits (ParticleData::type is uint8_t, ParticleData.h:31). */
#define HDM_TYPE_FIXED 1u   /* clamped particle: never moves, Diff = 0 */

typedef struct hdm_ctx hdm_ctx;

/* Host view of the HiDEMParticleData fields on the hot path
 * (HiDEMParticleData.h:20-39, ParticleData.h:29-37).  Any pointer may be NULL
 * on download (field skipped); on upload only the force fields may be NULL
 * (zero-filled).  Arrays hold n entries. */
/* The arrays of hdm_particles_host / hdm_links_host may be host memory or
 * device (or managed) memory of the context's GPU; uploads and downloads copy
 * on the device for the latter. */
typedef struct hdm_particles_host {
    int64_t n;
    double *xCord, *yCord, *zCord;
    double *xTheta, *yTheta, *zTheta;
    double *xDiff, *yDiff, *zDiff, *uDiff, *vDiff, *wDiff;
    double *radius;
    double *xForce, *yForce, *zForce, *uForce, *vForce, *wForce;
    uint8_t *type;
} hdm_particles_host;

/* Host view of HiDEMLinkData (HiDEMLinkData.h:24-31, LinkData.h:28-43).
 * CT[c] holds 2n side entries in checkpoint order: side 2l is the
 * localPartIndA end of link l, side 2l+1 the localPartIndB end (the order
 * LinkDataHandler::appendInitLinks creates, LinkDataHandler.h:229-236).
 * maxLoad is the per-link fracture threshold; NULL means "use the scalar
 * maxLoad of hdm_break_params". */
typedef struct hdm_links_host {
    int64_t n;
    int32_t *localPartIndA, *localPartIndB;
    uint8_t *state;
    double *initialLen;
    double *maxLoad;
    double *CT[6];
} hdm_links_host;

/* HiDEMDefNeighForceFunc (HiDEMDefaultFunctions.h:21-35) and
 * HiDEMDefLinkForceFunc (:90-103) parameters. */
typedef struct hdm_force_params {
    double invDt, attractDistance, attractScaleForce, damping, dampingRotation, scaleE;
    double E, W2, W4, G;
} hdm_force_params;

enum { HDM_INTEG_ICEBLOCK = 0, HDM_INTEG_GLACIER = 1 };

/* applyForce integrator parameters.  HDM_INTEG_ICEBLOCK: MainIceBlock.cpp
 * applyForce (:299-400) plus the clamp (HDM_TYPE_FIXED).  HDM_INTEG_GLACIER:
 * MainDem.cpp applyForce (:145-259), HiDEMDefVelocityLimit and
 * HiDEMDefBoundaryCondition, with the bed, friction and external force from
 * the maps of hdm_set_maps where bound, else analytic: height bedZ for
 * x < bedXEnd else deepZ, friction 1 for x < frictionXEnd else 0, a constant
 * external force ext[6]. */
typedef struct hdm_integ_params {
    int32_t variant;
    double rhoIce, g, lift, waterLevel, airDrag, waterDrag;
    double invDt, invDt2, dtSqr, invDtSqr, dt;
    double bedDrag, bedInteract, bedStiffness, frictionCap;
    double bedZ, bedXEnd, deepZ, frictionXEnd;
    double ext[6];
    double maxVel, maxRot;
    int32_t velocityLimit;
    int32_t bound[6];              /* xMin, xMax, yMin, yMax, zMin, zMax */
    double lower[3], upper[3];
} hdm_integ_params;

/* HiDEMDefLinkBreakFunc (HiDEMDefaultFunctions.h:207-215). */
typedef struct hdm_break_params {
    double bendBreak;
    double maxLoad;
} hdm_break_params;

typedef struct hdm_grid_info {
    double spacing, xMin, yMin, zMin, xMax, yMax, zMax;
    int32_t X, Y, Z;
    int64_t totalCells;
    int64_t occupiedCells;
    int32_t maxPerCell;
} hdm_grid_info;

/* Rebuild policies for hdm_run (gridService, HiDEMLib.h:533-628). */
enum {
    HDM_REBUILD_HEURISTIC = 0,  /* reference accumulated-Diff estimate (HiDEMLib.h:548-559) */
    HDM_REBUILD_EXACT = 1,      /* max |x - x_at_last_build| vs (interactDist - attractDistance)/2 */
    HDM_REBUILD_INTERVAL = 2,   /* every regenInterval steps ([grid] gridRegenInterval) */
    HDM_REBUILD_LEARNED = 3     /* exact test evaluated on the device every step (a step that finds the
                                   list stale does nothing); the host enqueues batches of steps sized by
                                   the learned scheduler (hdm_sched_*) without a per-step check.  Same
                                   rebuild steps and results as HDM_REBUILD_EXACT.  No reference
                                   counterpart: replaces the gridService check (HiDEMLib.h:544-559). */
};

typedef struct hdm_run_params {
    hdm_force_params force;
    hdm_integ_params integ;
    hdm_break_params brk;
    double searchDistance;      /* grid cell size */
    double interactDist;        /* neighbour-list skin */
    double dt, breakTime;
    int32_t policy;
    int32_t regenInterval;
} hdm_run_params;

typedef struct hdm_run_stats {
    int64_t timestep;           /* current timestep after the call */
    int64_t steps;              /* steps run by this call */
    int64_t rebuilds;           /* grid + neighbour-list rebuilds in this call */
    int64_t brokenTotal;        /* links with state == 0 after the call */
    double seconds;             /* wall time of the call, device-synchronised */
} hdm_run_stats;

/* Per-kernel-class device time accumulated with HIP events on the context's
 * stream since the last reset (milliseconds, and launch counts). */
enum { HDM_T_FORCE = 0, HDM_T_INTEGRATE, HDM_T_BREAK, HDM_T_SKIN, HDM_T_GRID, HDM_T_NEIGH, HDM_T_LINKS,
       HDM_T_STEP, HDM_T_COUNT };

/* ---- context ------------------------------------------------------------ */
int  hdm_ctx_create(int32_t precision, int32_t device, hdm_ctx **out);
int  hdm_ctx_destroy(hdm_ctx *ctx);
const char *hdm_last_error(const hdm_ctx *ctx);
int  hdm_abi_version(void);
int  hdm_device_count(int32_t *count);
int  hdm_sync(hdm_ctx *ctx);
int  hdm_set_timing(hdm_ctx *ctx, int32_t enabled);
int  hdm_kernel_times(hdm_ctx *ctx, double ms[HDM_T_COUNT], int64_t launches[HDM_T_COUNT]);
int  hdm_reset_kernel_times(hdm_ctx *ctx);
int  hdm_device_bytes(hdm_ctx *ctx, int64_t *bytes);

/* ---- data (ParticleDataHandler / LinkDataHandler ownership) --------------- */
int  hdm_particles_upload(hdm_ctx *ctx, const hdm_particles_host *p);
int  hdm_particles_download(hdm_ctx *ctx, hdm_particles_host *p);
int  hdm_links_upload(hdm_ctx *ctx, const hdm_links_host *l);
int  hdm_links_download(hdm_ctx *ctx, hdm_links_host *l);

/* ---- neighbour grid: NeighbourGrid::generate (NeighbourGrid.cpp:48-70) ---- */
int  hdm_grid_build(hdm_ctx *ctx, double searchDistance);
int  hdm_grid_info_get(hdm_ctx *ctx, hdm_grid_info *info);
/* colours[n] (cell id per particle), iterationOrder[n] (pInd sorted by
 * (colour, pInd)), either may be NULL. */
int  hdm_grid_download(hdm_ctx *ctx, int64_t *colours, int32_t *iterationOrder);

/* ---- neighbour list: NeighAccel::generateNeighbours (NeighAccel.h:25-57) -- */
int  hdm_neighbours_build(hdm_ctx *ctx, double interactDist);
/* CSR export: counts[n], starts[n+1], list[starts[n]] (capacity listCap). */
int  hdm_neighbours_download(hdm_ctx *ctx, int32_t *counts, int64_t *starts, int32_t *list, int64_t listCap);
int  hdm_neighbours_total(hdm_ctx *ctx, int64_t *total, int32_t *maxPerParticle);

/* ---- per-particle link lists + side layout: LinkAccelSide::generateLinks
 *      (LinkAccelSide.h:92-143) and sortSideData (:159-417) ---------------- */
int  hdm_links_build(hdm_ctx *ctx);
/* CSR export per particle in ascending link id: linkId, other endpoint, end
 * (0 = this particle is localPartIndA, 1 = localPartIndB). */
int  hdm_linklists_download(hdm_ctx *ctx, int32_t *counts, int64_t *starts, int32_t *linkId,
                            int32_t *other, int32_t *end, int64_t cap);

/* ---- per-step kernels ------------------------------------------------------ */
/* computePartForce with HiDEMDefNeighForceFunc + HiDEMDefLinkForceFunc:
 * CombinedAccel::iterateCombined (CombinedAccel.h:121-197). */
int  hdm_force(hdm_ctx *ctx, const hdm_force_params *fp);
/* applyForce (HiDEMLib.h:303-314) with the integrator of ip->variant. */
int  hdm_integrate(hdm_ctx *ctx, const hdm_integ_params *ip);
/* linkUpdate with HiDEMDefLinkBreakFunc: CombinedAccel::iterateLinks
 * (CombinedAccel.h:200-235).  nBroken (may be NULL) receives the number of
 * links broken by this call (reading it synchronises). */
int  hdm_break(hdm_ctx *ctx, const hdm_break_params *bp, int64_t *nBroken);
/* Reference rebuild estimate (HiDEMLib.h:424-474): stepSize*|max+ Diff|. */
int  hdm_skin_heuristic(hdm_ctx *ctx, int32_t stepSize, double *dist);
/* Exact test: max_i |x_i - x_i(at last snapshot)|. */
int  hdm_skin_snapshot(hdm_ctx *ctx);
int  hdm_skin_exact(hdm_ctx *ctx, double *maxDisp);

/* Allocation-free variant of hdm_skin_heuristic.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It takes no context: xDiff, yDiff, zDiff are n doubles in device memory of
 * the current GPU (e.g. HiDEMParticleData's Diff buffers), the work is
 * ordered on `stream` (a hipStream_t; NULL = the null stream) and *dist (host)
 * receives stepSize * |max+ Diff| after the call synchronises that stream.
 * Its partial maxima live in a fixed device array of 393,240 B that each GPU
 * holds until the process ends; the call allocates nothing.  n beyond
 * 134,217,728 gives HDM_E_CAPACITY.  Calls that overlap on one GPU share
 * that array and race.  hdm_last_error(NULL) returns the message of this
 * call's last failure on the calling thread. */
int  hdm_skin_heuristic_noalloc(const double *xDiff, const double *yDiff, const double *zDiff, int64_t n,
                                int32_t stepSize, void *stream, double *dist);

/* ---- the timestep loop: nSteps of gridService + computePartForce +
 *      applyForce + linkUpdate (when ts*dt > breakTime) + finalizeTs
 *      (MainIceBlock.cpp:409-447).  rebuildLog (nSteps bytes, may be NULL)
 *      receives 1 for steps that rebuilt the grid and neighbour list. ------- */
int  hdm_run(hdm_ctx *ctx, const hdm_run_params *rp, int64_t nSteps, uint8_t *rebuildLog, hdm_run_stats *stats);
int  hdm_set_timestep(hdm_ctx *ctx, int64_t ts);
int  hdm_count_broken(hdm_ctx *ctx, int64_t *broken);
/* finalizeTs statistics: computeActiveLinks and computeIteratedNeighbours
 * (HiDEMLib.h:197-232) and their sums over the owned particles
 * (HiDEMLib.h:518-523).  activeLinks[n]: the particle's links with
 * state != 0; iteratedNeighbours[n]: the entries of its neighbour list as the
 * reference builds it (NeighAccel, link partners within the skin included
 * also where the FP32 form's list leaves them out).  Host arrays; any
 * argument may be NULL.  Needs a built neighbour list and link lists. */
int  hdm_particle_stats(hdm_ctx *ctx, int32_t *activeLinks, int32_t *iteratedNeighbours, int64_t *activeLinksSum,
                        int64_t *iteratedNeighboursSum);

/* ---- learned rebuild scheduler (no reference counterpart; it schedules the
 *      gridService rebuild check, HiDEMLib.h:544-559) -------------------------
 * Blocks of blockCells^3 neighbour-grid cells get 16 features each; a BF16
 * MLP 16-64-64-1 (MFMA, FP32 accumulation) predicts log2 of the steps until
 * the block would trip the exact test.  HDM_REBUILD_LEARNED enqueues
 * clamp(safety * 2^min_prediction, minBatch, maxBatch) steps per host check
 * (maxBatch when no model is set). */
#define HDM_SCHED_FEATURES 16
#define HDM_SCHED_HIDDEN 64
typedef struct hdm_mlp_weights {        /* FP32 master weights */
    float W1[HDM_SCHED_FEATURES * HDM_SCHED_HIDDEN];   /* [in][out] */
    float b1[HDM_SCHED_HIDDEN];
    float W2[HDM_SCHED_HIDDEN * HDM_SCHED_HIDDEN];     /* [in][out] */
    float b2[HDM_SCHED_HIDDEN];
    float w3[HDM_SCHED_HIDDEN];
    float b3;
    float featMean[HDM_SCHED_FEATURES];                /* input standardisation */
    float featStd[HDM_SCHED_FEATURES];
} hdm_mlp_weights;

typedef struct hdm_sched_params {
    int32_t blockCells;         /* block edge in grid cells (4) */
    int32_t minBatch, maxBatch; /* steps per host check */
    float safety;               /* fraction of the predicted horizon enqueued */
} hdm_sched_params;

typedef struct hdm_mlp_train_params {
    int32_t epochs, batch;      /* minibatches are consecutive slices of the (pre-shuffled) rows */
    float lr, beta1, beta2, eps;
} hdm_mlp_train_params;

int  hdm_sched_configure(hdm_ctx *ctx, const hdm_sched_params *sp);
int  hdm_sched_set_model(hdm_ctx *ctx, const hdm_mlp_weights *w);   /* NULL: no model */
/* Features of the current state (blocks laid over the last grid build).
 * Host arrays may be NULL, else hold `capacity` blocks (HDM_E_CAPACITY when
 * nBlocks exceeds it): features [nBlocks][16], blockDisp [nBlocks] (max
 * |x - x0| of the block, metres), valid [nBlocks]. */
int  hdm_sched_features(hdm_ctx *ctx, const hdm_run_params *rp, int64_t stepsSinceRebuild, int32_t capacity,
                        int32_t *nBlocks, float *features, float *blockDisp, uint8_t *valid);
/* Predictions (log2 steps) for the features of the last feature evaluation;
 * pred (may be NULL) holds `capacity` floats; minPred: minimum over valid
 * blocks. */
int  hdm_sched_predict(hdm_ctx *ctx, int32_t capacity, float *pred, double *minPred);
/* The network on host arrays (copied to the device): X [n][16] raw features. */
int  hdm_mlp_forward(hdm_ctx *ctx, const hdm_mlp_weights *w, const float *X, int64_t n, float *y);
/* Sum over rows of the gradient of the loss (squared error in log2 steps;
 * censored rows only below their target) with respect to W1, b1, W2, b2, w3,
 * b3 (the first 5313 floats of hdm_mlp_weights' layout), and the loss sum. */
int  hdm_mlp_gradient(hdm_ctx *ctx, const hdm_mlp_weights *w, const float *X, const float *target,
                      const uint8_t *censored, int64_t n, float *grad, double *loss);
/* Minibatch Adam on the device; w is updated in place.  loss: mean loss of
 * the last epoch. */
int  hdm_mlp_train(hdm_ctx *ctx, hdm_mlp_weights *w, const float *X, const float *target, const uint8_t *censored,
                   int64_t n, const hdm_mlp_train_params *tp, double *loss);
/* Training labels (up to 32 concurrent references).  _begin takes the blocks
 * and block displacements of the last hdm_sched_features call and the
 * current positions as a reference; after every later step, _track records
 * per block the first step count at which one of its particles has moved the
 * block's remaining skin budget ((interactDist - attractDistance)/2 minus the
 * block displacement at _begin) from its reference position -- a lower bound
 * of the steps until the block trips the exact test.  _end downloads the
 * counts (-1: not reached) and frees the slot. */
int  hdm_sched_label_begin(hdm_ctx *ctx, int32_t slot, const hdm_run_params *rp);
int  hdm_sched_label_track(hdm_ctx *ctx);
int  hdm_sched_label_end(hdm_ctx *ctx, int32_t slot, int32_t capacity, int32_t *firstCross, int32_t *nBlocks);
/* Host checks and skipped launches of HDM_REBUILD_LEARNED runs since the
 * context was created. */
int  hdm_sched_stats(hdm_ctx *ctx, int64_t *hostChecks, int64_t *stepsLaunched, int64_t *stepsExecuted);

/* ---- execution options of HDM_REBUILD_LEARNED runs (no reference
 *      counterpart; the reference tunes launch constants statically in
 *      Tune.h) ---------------------------------------------------------------
 * autotune: the next run's first steps launch every shape of the fused step
 * (workgroup size, XCD-contiguous workgroup numbering) in turn -- all shapes
 * compute identical results -- and keep the fastest.  graphs: plain middle
 * steps run as replays of a captured two-step hipGraph. */
typedef struct hdm_exec_params {
    int32_t autotune;
    int32_t graphs;
    int32_t variant;            /* shape to use (0..variants-1); -1 keeps the current one */
    int32_t listWithPartners;   /* HDM_PREC_MIXED: 1 keeps link partners in the neighbour list
                                   (the reference's list); 0 (default) leaves them out and the
                                   link loop applies their contact law (fewer gathers).  FP64
                                   always uses the reference's list. */
    int32_t persistent;         /* 1: the steps of a batch run in one co-resident launch that
                                   synchronises the grid between steps (256-thread workgroups,
                                   no autotuned shape, no graphs); same results */
} hdm_exec_params;

typedef struct hdm_exec_info {
    int32_t variant, variants, tuned;
    char name[32];
    double tuneMs[8];           /* best step time per shape seen by the autotuner (ms) */
    int64_t graphCaptures, graphReplays;
    int32_t mergedContacts;     /* the current neighbour list leaves out link partners */
    int64_t persistentLaunches; /* co-resident multi-step launches (hdm_exec_params.persistent) */
} hdm_exec_info;

int  hdm_set_exec(hdm_ctx *ctx, const hdm_exec_params *ep);
int  hdm_exec_info_get(hdm_ctx *ctx, hdm_exec_info *info);

/* ---- HBM bandwidth probes (the idea of Benchmark/BenchMain.cpp:14-83):
 *      kind 0 = streaming copy (read `bytes`, write `bytes`),
 *      kind 1 = six-array read-modify-write, the access pattern of the CT
 *      side update (reads and writes `bytes` in total each way).
 *      Reports achieved GB/s (bytes read + written / time, HIP events). ------ */
enum { HDM_PROBE_COPY = 0, HDM_PROBE_CT_RMW = 1 };
int  hdm_bw_probe(hdm_ctx *ctx, int32_t kind, int64_t bytes, int32_t iters, double *gbps);

/* ---- seeded FCC block and link generator (setup; build-defined, replaces the
 *      non-deterministic FCCParticleGenerator/LinkGenerator) ---------------- */
typedef struct hdm_fcc_spec {
    int32_t nx, ny, nz;         /* unit cells */
    double a;                   /* lattice constant */
    double origin[3];
    double radius;
    double jitter;              /* uniform +-jitter per coordinate */
    uint64_t seed;
    double clampXBelow;         /* particles with x < clampXBelow get HDM_TYPE_FIXED (use -inf to disable) */
    int32_t order;              /* HDM_ORDER_LATTICE or HDM_ORDER_MORTON (unit cells along a Z-order curve) */
} hdm_fcc_spec;

enum { HDM_ORDER_LATTICE = 0, HDM_ORDER_MORTON = 1 };

int  hdm_gen_fcc_count(const hdm_fcc_spec *spec, int64_t *nParticles);
int  hdm_gen_fcc_fill(const hdm_fcc_spec *spec, hdm_particles_host *out);
/* Links join every pair with dist <= (r_i + r_j) * linkGenScale
 * (LinkGenerator.h:83-89); end A = the smaller global key (index), links in
 * ascending (A, B); initialLen = dist; maxLoad = threshold*(1 + spread*U(-1,1))
 * from a counter-based hash of (seed, A, B). */
int  hdm_gen_links_count(const hdm_particles_host *p, double linkGenScale, int64_t *nLinks);
int  hdm_gen_links_fill(const hdm_particles_host *p, double linkGenScale, double threshold, double spread,
                        uint64_t seed, hdm_links_host *out);

/* ---- batched form: many independent ice blocks in one state --------------
 * Block b gets N_b particles, log-uniform in [minParticles, maxParticles]
 * (FCC, nx = 2*ny = 2*nz unit cells), its own fracture-threshold mean
 * U[thrLo, thrHi] (per link: mean*(1 + spread*U(-1,1))), its own waterline
 * U[waterLo, waterHi] x block height, and a clamp on its first clampX metres.
 * Blocks are packed on a 16 x ceil(B/16) shelf grid in x-y with `gap` metres
 * between them (gap > search distance: no cross-block neighbours or links).
 * Particles of a block are contiguous; blockId[i] names the block. */
typedef struct hdm_batch_spec {
    int32_t nBlocks;
    int64_t minParticles, maxParticles;
    double a, radius, jitter, gap;
    double thrLo, thrHi, spread;
    double waterLo, waterHi;
    double clampX;
    uint64_t seed;
} hdm_batch_spec;

typedef struct hdm_batch_block {
    int32_t nx, ny, nz;
    double origin[3];
    int64_t particleStart, particleCount;
    double threshold;           /* mean fracture threshold of the block */
    double waterLevel;          /* absolute z of the block's waterline */
} hdm_batch_block;

int  hdm_gen_batch_layout(const hdm_batch_spec *spec, hdm_batch_block *blocks, int64_t *nParticles);
int  hdm_gen_batch_fill(const hdm_batch_spec *spec, const hdm_batch_block *blocks, hdm_particles_host *out,
                        int32_t *blockId);
/* Scales per-link thresholds (generated with threshold 1) by the mean of the
 * block that owns end A. */
int  hdm_gen_batch_thresholds(const hdm_batch_spec *spec, const hdm_batch_block *blocks, const int32_t *blockId,
                              hdm_links_host *links);
/* Per-particle block ids and per-block waterlines for the integrator
 * (replaces hdm_integ_params.waterLevel per particle); nBlocks = 0 clears. */
int  hdm_set_blocks(hdm_ctx *ctx, const int32_t *blockId, int32_t nBlocks, const double *waterLevel);

/* Redistribution order on one rank (HiDEMDistribution.h:150-181): particles
 * sorted by their cell colours on grids of spacing s, 2s, 4s, 8s (s =
 * searchDistance; NeighbourGrid's bounding box and colours), coarsest first
 * (ParallelSort::sortIndirectNLevel), ties by old index; links sorted by
 * their old (A, B) (LinkData::getNewOrder, LinkData.h:166-181), ties by link
 * index, endpoints renamed (reorderLinkTargets, :116-137), CT sides moved
 * with their link.  Every particle and link array is permuted in place, so
 * later downloads return the new order.  order[i] / linkOrder[j] (host, n /
 * L entries, may be NULL) receive the old index of new particle i / link j.
 * The grid and neighbour list must be built again (hdm_run does so on its
 * first step); link lists built before are rebuilt by the call. */
int  hdm_reorder(hdm_ctx *ctx, double searchDistance, int64_t *order, int64_t *linkOrder);

/* ---- GlacierDem map inputs ----------------------------------------------------
 * A 2-D field on a regular xPoints x yPoints grid over [xMin, xMax] x
 * [yMin, yMax] (Map2d, Map2d.h:22-61), data[ix + iy * xPoints] (host
 * memory, lpReal_t values), read at (x, y) by Map2d::interp (Map2d.h:110-145):
 * bilinear interpolation in FP32 of the four grid values around the point,
 * indexes clamped to the grid, the value itself where both axes clamp to one
 * point (so a 1 x 1 map is a constant, as ExternalForceBuilder makes for a
 * single value, ExternalForces.h:111-121).  xOffset, yOffset: Map2d::setOffset. */
typedef struct hdm_map2d {
    int32_t xPoints, yPoints;
    float xMin, yMin, xMax, yMax;
    float xOffset, yOffset;
    const double *data;
} hdm_map2d;

/* Slots of hdm_set_maps: bed height and friction (InputBed, InputGeometry.h
 * :53-67), external force x, y, z, u, v, w (ExternalForces, ExternalForces.h
 * :24-70). */
enum { HDM_MAP_BED = 0, HDM_MAP_FRICTION = 1, HDM_MAP_EXT = 2, HDM_MAP_COUNT = 8 };

/* Map-driven inputs of the HDM_INTEG_GLACIER integrator (MainDem.cpp:176-217:
 * bed->getHeight, bed->getFric, extForcePtr->xForce->interp ...).  maps[k]
 * NULL keeps the analytic input of hdm_integ_params for that slot (bedZ /
 * bedXEnd / deepZ, frictionXEnd, ext[k]).  The data is copied to the device;
 * a later call replaces every slot (all NULL, or maps NULL, clears).
 * HDM_E_INVALID for a map with fewer than 1 point per axis or no data. */
int  hdm_set_maps(hdm_ctx *ctx, const hdm_map2d *const *maps);

/* Map2dLoader::readInput + get (Map2dLoader.cpp:8-86, Map2dLoader.h:29-40):
 * reads a whitespace-separated table whose first two columns are x and y
 * (rows with the same y form one grid row, every row the same length) and
 * fills *map with column fieldID on the grid the x and y columns span (each
 * row placed at its rounded grid index, Map2d::setPointCord).  map->data is
 * owned by the library until hdm_map2d_free.  HDM_E_INVALID for a missing or
 * malformed file (a row with another column count, rows of unequal length, a
 * field index beyond the columns). */
int  hdm_map2d_load(const char *path, int32_t fieldID, hdm_map2d *map);
void hdm_map2d_free(hdm_map2d *map);

/* ---- TOML-driven simulation (HiDEMLib::setup + the sample main loop) --------
 * argv holds the command line after the program name, as the reference's
 * ArgParser expects it (ArgParser.h:170-196): optional --section.key=value
 * overrides, then the config path last.  The [general], [iteration], [grid]
 * and [lattice] keys are the reference's; [hdm] holds the build-defined
 * extras (block size, jitter, clamp, threshold spread, integrator, bed,
 * rebuild policy).  hdm_sim_steps runs gridService + computePartForce +
 * applyForce + linkUpdate + finalizeTs per step (MainIceBlock.cpp:409-447). */
typedef struct hdm_sim hdm_sim;

/* Everything setup derives from a config, without touching a GPU: the step
 * parameters, the lattice spec and the link generation parameters. */
typedef struct hdm_setup_info {
    hdm_run_params run;
    hdm_fcc_spec fcc;
    double linkGenScale, threshold, spread;
    uint64_t linkSeed;
    int64_t endTimestep;
    int32_t precision;
} hdm_setup_info;
int  hdm_config_derive(int32_t argc, const char *const *argv, hdm_setup_info *out);

int  hdm_sim_create(int32_t precision, int32_t argc, const char *const *argv, hdm_sim **out);
int  hdm_sim_destroy(hdm_sim *sim);
const char *hdm_sim_last_error(void);
int  hdm_sim_ctx(hdm_sim *sim, hdm_ctx **ctx);
int  hdm_sim_run_params(hdm_sim *sim, hdm_run_params *rp);
int  hdm_sim_info(hdm_sim *sim, int64_t *nParticles, int64_t *nLinks, int64_t *timestep, int64_t *endTimestep);
int  hdm_sim_steps(hdm_sim *sim, int64_t nSteps, uint8_t *rebuildLog, hdm_run_stats *stats);
const char *hdm_sim_config_dump(hdm_sim *sim);
/* HiDEMH5Checkpoint::writeCheckpoint (HiDEMH5Checkpoint.h:23-49) of the
 * sim's current state and step, in the reference's layout (hdm_io.h);
 * loads libhidem_io.so from this library's directory. */
int  hdm_sim_checkpoint(hdm_sim *sim, const char *path);

#ifdef __cplusplus
}
#endif
#endif
