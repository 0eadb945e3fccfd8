// Internal state of an hdm_ctx: device mirrors of the hot-path fields, the
// neighbour grid, the neighbour list, the per-particle link lists and the side
// (CT) store, all owned by the context and ordered on its HIP stream.
//
// Device layout (every array is one hipMalloc, structure of arrays):
//   particles  x/y/zCord, x/y/zTheta, x..wDiff, radius: f64;  x..wForce: f64
//              (FP64 form) or f32 (mixed form); type: u8.  Index = pInd.
//              Coordinates, rotations and Diff have a second buffer each that
//              the fused step kernel integrates into before the two swap.
//   links      localPartIndA/B: i32, state: u8, initialLen, maxLoad: f64.
//   slots      "wave tiles": particle p owns slot k of tile p/64 at
//              tile*64*width + k*64 + p%64, so the 64 lanes of a wavefront
//              that process 64 consecutive particles read slot k of all of
//              them from one contiguous 256 B (i32) / 512 B (f64) run.
//              Neighbour list (width maxNeigh) and link slots (width
//              maxLinks: link id, other endpoint, CT0..CT5 side state) use it.
//   grid       colour (i32 per particle), iterationOrder (i32), dense
//              cellStart (i32, totalCells+1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hidem/hdm_abi.h"

namespace hdm {

// Device descriptor of an hdm_map2d (Map2d's fields, with xRes, yRes and
// the inverse extents computed in float as its constructor does).
struct MapDev {
    int32_t xPoints, yPoints;
    float xMin, yMin, xRes, yRes, invXMaxMin, invYMaxMin, xOffset, yOffset;
    const double *data;
};

// Raw device allocation that grows on demand (contents are not preserved).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t want) {
        if (want <= bytes && ptr) return hipSuccess;
        release();
        if (want == 0) want = 256;
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) { ptr = nullptr; return e; }
        bytes = want;
        return hipSuccess;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

inline void swap_buf(DevBuf &a, DevBuf &b) {
    void *p = a.ptr;
    size_t s = a.bytes;
    a.ptr = b.ptr;
    a.bytes = b.bytes;
    b.ptr = p;
    b.bytes = s;
}

struct TimedLaunch {
    int cls;
    hipEvent_t start, stop;
};

struct GridState {
    double spacing = 0, xMin = 0, yMin = 0, zMin = 0, xMax = 0, yMax = 0, zMax = 0;
    int32_t X = 0, Y = 0, Z = 0;
    int64_t totalCells = 0;
    int64_t occupied = 0;
    int32_t maxPerCell = 0;
    bool built = false;
    // the cell records hold the positions of this state (Context::stateEpoch, ts)
    uint64_t recEpoch = 0;
    int64_t recTs = -1;
};

// One training-label reference of the learned scheduler (hdm_sched_label_*).
struct LabelSlot {
    DevBuf ref, blockOf, remaining, first;
    int32_t nb = 0;
    int64_t start = 0;
    bool active = false;
};
constexpr int LABEL_SLOTS = 32;

struct Context {
    int32_t prec = HDM_PREC_FP64;
    int32_t device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int32_t numCU = 0;
    bool coopLaunch = false;     // the device takes cooperative (co-resident) launches

    // particles
    int64_t n = 0;
    DevBuf x, y, z, tx, ty, tz, diff[6], radius, force[6], type;
    // second buffers of the fused step: alt[3..5] hold the integrated
    // rotations (each hdm_run step writes them and then swaps them with
    // tx, ty, tz); the other entries are unused
    DevBuf alt[12];
    // gather records of the fused step, ping-pong like the state:
    //   rec   FP64 form: double4 {x, y, z, radius}; FP32 form: MixedRec
    //         {x, y, z, Diff0..5 in FP32, rotation bound, radius, low parts
    //         of Diff0 and Diff1} (64 B); the
    //         radius reads NaN when a Diff component exceeds recTame (then
    //         the contact skip test fails)
    //   drec  FP64 form: FP64 Diff0..5 of every particle, the rotation
    //         bound |xTheta| + |yTheta| + |zTheta| (as a float) and a 0 (8
    //         doubles, gathered by partners).  FP32 form: 3 doubles, the
    //         FP64 translation Diff of the particles whose Diff does not
    //         split exactly into the record's FP32 part and low parts
    //         (split_diff3); others leave their entry untouched
    //   dtail FP32 form: MixedTail, the FP64 rotation Diff and the low part
    //         of Diff2 (NaN when the translation Diff is in drec), 32 B.
    //         Only the particle reads its own; partners gather its MixedRec
    //         (coordinates, FP32 Diff, rotation bound, radius flag)
    // The fused step reads and writes only the records for coordinates and
    // Diff: after it the SoA x, y, z and diff[] are stale (soaStale) until
    // sync_state unpacks them.
    DevBuf rec, recAlt, drec, drecAlt, dtail, dtailAlt;
    bool recCurrent = false;     // rec/drec/dtail hold the current state
    int32_t guardCur = 0;        // trip word of rec (0/1); swaps with the records
    // Whether the state the last learned run left trips the exact test
    // (valid while stateEpoch and ts are those it recorded).
    bool tripKnown = false, tripPending = false;
    uint64_t tripEpoch = 0;
    int64_t tripTs = -1;
    // Broken-link count kept by every kernel that breaks links (brokenCount,
    // u64 on the device); a full count refreshes it when not valid.
    DevBuf brokenCount;
    bool brokenCountValid = false;
    bool soaStale = false;       // x, y, z, diff[] lag the records
    bool altHoldsPrev = false;   // recAlt/drecAlt/dtailAlt hold the state before the last fused step
    // FP32 force: the neighbour list leaves out each particle's link partners
    // and the link loop applies their contact law (merge_linked_contacts)
    bool partnersUnique = false;     // no particle has two sides to one partner or a side to itself
    bool mergeContacts = true;       // allowed (hdm_set_exec)
    bool nbrExcludesLinked = false;  // the current list was built that way
    double recTame = 0;          // Diff bound the radius flags were made with
    // batched form: block id per particle, waterline per block (nBlocks = 0: unused)
    int32_t nBlocks = 0;
    DevBuf blockId, blockWater;
    // GlacierDem map inputs (hdm_set_maps): descriptors (MapDev[HDM_MAP_COUNT])
    // and their values; bit k of mapMask: slot k holds a map
    int32_t mapMask = 0;
    DevBuf mapDesc, mapData[HDM_MAP_COUNT];
    // links
    int64_t L = 0;
    bool hasMaxLoad = false;
    DevBuf la, lb, state, initialLen, maxLoad;
    DevBuf ctStage[6];     // CT in checkpoint order (2l + end) during upload/download
    // link slots (wave tiles) and side store
    int32_t maxLinks = 0;
    bool linksBuilt = false;
    DevBuf linkCount, linkCursor, slotLink, slotOther;
    DevBuf ct;             // CT0..5 of every slot, 6 x 64 values per (wave tile, side) block (ct_offset)
    DevBuf sidePos;        // 2L: tiled slot of side (l, end)
    DevBuf slotThr;        // u16 per slot: break prefilter threshold code of the side's link (k_side_codes)
    float thrScale = 0.0f; // threshold = code * thrScale (at most the link's L0 (1 + maxLoad), or L0)
    float l0Max = 0.0f;    // largest initialLen (rounded up to float)
    // neighbour list (wave tiles)
    int32_t maxNeigh = 0;      // ELL width of the list (at least the largest count; kept across builds)
    int32_t maxNeighCount = 0; // largest count of the current list
    bool neighBuilt = false;
    DevBuf nbrCount, nbr;
    DevBuf statsBuf;       // i32 [2][n]: hdm_particle_stats output staging
    DevBuf nbrLinked;      // i32 per particle: link partners within the skin that the list leaves out
                           // (merge_linked_contacts), so the reference list holds nbrCount + nbrLinked
    // grid
    GridState grid;
    DevBuf colour, order, cellStart, cellCursor, scanPartials;
    DevBuf cellRec;        // double4 [n]: {x, y, z, radius} of order[i] (the list build's candidate records)
    DevBuf cellRecF;       // float4 [n]: the same rounded to FP32 (the list build's first test)
    DevBuf cellPos;        // i32 [n]: position of each particle in order
    // reductions / scalars
    DevBuf red;            // u64 scratch for bbox / max reductions and counters
    DevBuf skinPartials;   // f64 [workgroups][4]: rebuild-test maxima per integrating workgroup
    DevBuf guardState;     // u32 [GUARD_WORDS]: ticket, guard flag, executed-step counter
    DevBuf x0, y0, z0;     // positions at the last neighbour-list build
    DevBuf x0f, y0f, z0f;  // the same rounded to FP32 (the device guard's first test)
    DevBuf probe;          // bandwidth-probe buffer (hdm_bw_probe)
    bool snapshotValid = false;
    bool skinFused = false;      // red[32..34] hold max+ Diff from the last integrate of hdm_run
    bool skinFusedDisp = false;  // red[36] holds max |x - x0|^2 from that integrate
    // execution: launch shape of the fused step (autotuned) and graph replay
    int32_t stepVariant = 0;     // index into the step shapes (kernels_step.hip)
    bool autotune = false, tuned = false, useGraphs = false, capturing = false;
    double tuneMs[8] = {0};
    int64_t tuneCount[8] = {0};
    bool usePersistent = false;  // hdm_exec_params.persistent
    int64_t loopLaunches = 0;    // k_step_loop launches
    DevBuf loopBarrier;          // grid-barrier words of k_step_loop (+1: timeout flag)
    DevBuf loopSets;             // k_step_loop argument sets (this step's, roles swapped)
    hipGraphExec_t graphExec = nullptr;      // replays GRAPH_STEPS_LONG steps
    hipGraphExec_t graphExecPair = nullptr;  // replays 2 steps
    std::vector<unsigned char> graphKey, graphKeyPair;
    int64_t graphCaptures = 0, graphReplays = 0;
    // learned rebuild scheduler
    hdm_sched_params sched = {4, 8, 256, 0.5f};
    bool haveModel = false;
    bool guardArmed = false;     // guard flag may be set (HDM_REBUILD_LEARNED left it)
    int32_t schedBX = 0, schedBY = 0, schedBZ = 0;
    double schedOx = 0, schedOy = 0, schedOz = 0, schedEdge = 0;
    bool schedAssigned = false;  // block ids match the last grid build
    int64_t stepsSinceRebuild = 0, schedLastEval = 0;
    int64_t schedChecks = 0, schedLaunched = 0, schedExecuted = 0;
    // Horizon left over when a learned run ended inside a scheduled batch:
    // the next run continues that batch instead of evaluating the network
    // again, as one longer run would.  Valid while the state is the one the
    // run left (stateEpoch unchanged, same timestep).
    int64_t schedCarry = 0, schedCarryTs = -1;
    uint64_t stateEpoch = 0, schedCarryEpoch = 0;
    DevBuf schedBlockOf, schedAcc, schedFeat, schedDisp, schedValid, schedPred, schedBrokenPrev;
    int32_t schedNb = 0;         // blocks of the last feature evaluation
    hdm_mlp_weights schedModel;  // the scheduler's network (host copy)
    bool mlpHoldsModel = false;  // mlpW holds schedModel (standalone MLP calls replace it)
    DevBuf mlpW, mlpW1b, mlpW2b, mlpGpart;
    DevBuf mlpX, mlpT, mlpC, mlpY, mlpGrad, mlpM, mlpV, mlpLoss;
    LabelSlot labels[LABEL_SLOTS];
    // run state
    int64_t ts = 0;
    double accumulatedDist = 0;
    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> eventPool;
    double ms[HDM_T_COUNT] = {0};
    int64_t launches[HDM_T_COUNT] = {0};

    int64_t tiles() const { return (n + 63) / 64; }
    size_t ctBytes() const { return prec == HDM_PREC_FP64 ? 8 : 4; }
    size_t forceBytes() const { return prec == HDM_PREC_FP64 ? 8 : 4; }
};

// Whether a neighbour list built now leaves out link partners: FP32 force
// only (its summation order is its own; the FP64 force keeps the reference's
// list order), with every particle's partners distinct and not itself.  The
// contact law of a linked pair then runs in the link loop, so each step skips
// the ~12 record gathers of a lattice particle's linked neighbours, which are
// inert contacts while the link holds.
inline bool merge_linked_contacts(const Context &c) {
    return c.prec == HDM_PREC_MIXED && c.linksBuilt && c.partnersUnique && c.mergeContacts && c.maxLinks <= 64;
}

// Link slots hold the link id; the sign bit marks a broken link (state == 0).
constexpr int32_t SLOT_BROKEN = (int32_t)0x80000000u;
constexpr int32_t SLOT_LINK_MASK = 0x7fffffff;

// Launch helpers shared by the kernel files.
int xcd_grid(int64_t nBlocks);

// Message of the last failed call on this thread that has no context (the
// allocation-free entry point); hdm_last_error(NULL) returns it.
void set_free_error(const std::string &msg);

// Block grid of the learned scheduler.
struct SchedGeometry {
    double ox, oy, oz, edge;
    int32_t BX, BY, BZ;
};
struct SchedFeatureInputs {
    double budget;          // (interactDist - attractDistance) / 2
    double waterLevel;
    int64_t stepsSince, stepsSincePrev;
};

// Kernel launchers (defined in kernels_*.hip).  All enqueue on ctx.stream.
hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, int64_t n);
hipError_t launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t v, int64_t n);
hipError_t launch_bbox(Context &c);
hipError_t launch_colour_count(Context &c);
hipError_t launch_exclusive_scan_i32(Context &c, const int32_t *in, int32_t *out, int64_t count);
hipError_t launch_scatter_sort(Context &c);   // ends with launch_cell_records
// cellRec / cellPos from the current positions in the grid's order.
hipError_t launch_cell_records(Context &c);
hipError_t launch_neigh_count(Context &c, double interactDist);
hipError_t launch_neigh_fill(Context &c, double interactDist);
hipError_t launch_max_i32(Context &c, const int32_t *v, int64_t n, uint64_t *out);
hipError_t launch_link_count(Context &c);
hipError_t launch_link_fill(Context &c);
hipError_t launch_ct_import(Context &c);   // ctStage (checkpoint order) -> ct (tiled)
hipError_t launch_ct_export(Context &c);   // ct (tiled) -> ctStage (checkpoint order)
// initialLen / maxLoad -> slotThr codes (after link_fill); maxBits: 2 device
// words of scratch; returns the largest initialLen (rounded up) and the code
// scale
hipError_t launch_side_thresholds(Context &c, uint64_t *maxBits, float *l0Max, float *scale);
// *flag != 0 after completion when a particle has two sides to one partner or
// a side to itself (after link_fill).
hipError_t launch_partner_check(Context &c, uint64_t *flag);
hipError_t launch_pack_records(Context &c, double tame);  // SoA state -> rec / drec
// Redistribution order (hdm_reorder): permutes the SoA particle state, the
// links and the CT stage (checkpoint order) on the device; order / linkOrder
// (host, may be null) receive the old index of each new particle / link.
hipError_t launch_reorder(Context &c, double searchDistance, int64_t *order, int64_t *linkOrder);
// Brings x, y, z and diff[] up to date with the records after fused steps
// (no-op otherwise).  Every launcher that reads them calls it first.
hipError_t sync_state(Context &c);
// The buffer swap of one fused step (undone by a second call).
void swap_step_state(Context &c);
hipError_t launch_force(Context &c, const hdm_force_params &fp);
// skin != nullptr: the rebuild-test maxima of the new state land in
// skin[0..2] (max+ Diff per axis) and skin[4] (max |x - x0|^2, when
// withDisplacement), as ordered keys, when the launch completes.
hipError_t launch_integrate(Context &c, const hdm_integ_params &ip, uint64_t *skin = nullptr,
                            bool withDisplacement = false);
// Launch shapes of the fused step (autotuner candidates).
constexpr int STEP_VARIANTS = 7;
const char *step_variant_name(int v);
// Autotuner choice from the per-shape step times (ms).
int pick_step_variant(const double ms[STEP_VARIANTS]);

// Guard words in Context::guardState.
// GUARD_FLAG + guardCur is the trip word of the current record buffer, the
// other one that of the alternate buffer (they swap with the records).
enum { GUARD_TICKET = 0, GUARD_FLAG = 1, GUARD_STEPS = 3, GUARD_WORDS = 4 };
// u64 slot of Context::red where the skin keys of the last integrating
// launch start (max+ Diff per axis at SKIN_KEYS + 0..2, max |x - x0|^2 at
// SKIN_KEYS + 4).
enum { SKIN_KEYS = 32 };

// Exact rebuild test evaluated on the device by launch_step: every particle
// whose new position lies 2 |x - x0| >= threshold from its snapshot sets the
// trip word of the record buffer the launch writes (2 |x - x0| >= threshold
// for some particle exactly when it holds for the largest |x - x0|); a launch
// whose input buffer's word is set does nothing but set its output word, so
// steps launched after a trip do nothing until the host clears the words.
// Workgroup 0 of each executed launch counts it (GUARD_STEPS).
struct StepGuard {
    double threshold;   // interactDist - attractDistance
};

// Force + integration of one timestep in one launch (state ping-pong; the
// integrated state is current afterwards).  skin as for launch_integrate.
// brk != nullptr: first apply the link break that the previous timestep's
// launch_break would have applied to the current state.  storeForce: write
// the force arrays (else they keep their previous contents).  guard: see
// StepGuard (requires skin and withDisplacement).
hipError_t launch_step(Context &c, const hdm_force_params &fp, const hdm_integ_params &ip, uint64_t *skin,
                       bool withDisplacement, const hdm_break_params *brk, bool storeForce,
                       const StepGuard *guard = nullptr);
// Clears the guard flag and the executed-step counter (GUARD_FLAG, GUARD_STEPS).
hipError_t launch_guard_reset(Context &c);
// `steps` guarded fused steps in one co-resident launch (k_step_loop), for
// HDM_REBUILD_LEARNED with hdm_exec_params.persistent; every step applies
// the previous step's break and none stores the force.  Ends early after a
// step that trips the exact test, like the per-step launches.
// *barrierFailed: a grid-barrier wait timed out (the steps are not valid).
hipError_t launch_step_loop(Context &c, const hdm_force_params &fp, const hdm_integ_params &ip, uint64_t *skin,
                            const hdm_break_params &brk, const StepGuard &guard, int32_t steps, bool *barrierFailed);
hipError_t launch_break(Context &c, const hdm_break_params &bp, uint64_t *counter);
// The same break per particle from the current records (hdm_run's last step;
// launch_break when the records are not current).  counter: links broken.
hipError_t launch_break_records(Context &c, const hdm_break_params &bp, uint64_t *counter);
hipError_t launch_skin_heuristic(Context &c, uint64_t *out3);
hipError_t launch_skin_exact(Context &c, uint64_t *out1);
hipError_t launch_snapshot(Context &c);
hipError_t launch_count_broken(Context &c, uint64_t *counter);
// computeActiveLinks / computeIteratedNeighbours (HiDEMLib.h:197-232):
// per particle its intact links and its reference neighbour-list entries
// (activeLinks / iteratedNeighbours, device arrays of n, may be nullptr) and
// their sums into sums[0], sums[1].
hipError_t launch_particle_stats(Context &c, int32_t *activeLinks, int32_t *iteratedNeighbours, uint64_t *sums);
// learned scheduler (kernels_sched.hip)
int64_t mlp_param_count();
hipError_t sched_assign_blocks(Context &c, const SchedGeometry &g);
hipError_t sched_features(Context &c, const SchedGeometry &g, const SchedFeatureInputs &in);
hipError_t sched_min_prediction(Context &c, int64_t nBlocks, uint64_t *out);
hipError_t sched_label_begin(Context &c, int slot, double budget);
hipError_t sched_label_track(Context &c);
hipError_t mlp_upload(Context &c, const hdm_mlp_weights &w);
hipError_t mlp_forward(Context &c, const float *X, int64_t n, float *y);
hipError_t mlp_gradient(Context &c, const float *X, const float *target, const uint8_t *censored, int64_t n,
                        float *grad, float *lossPart);
hipError_t mlp_adam(Context &c, const float *grad, float *m, float *v, float scale, float lr, float beta1,
                    float beta2, float eps, int64_t t);
hipError_t launch_probe(Context &c, int kind, void *buf, int64_t bytes);
hipError_t launch_probe_zero(Context &c, void *buf, int64_t bytes);

// Monotone map between doubles and unsigned 64-bit keys (used for
// order-independent, hence deterministic, atomic min/max of doubles).
__host__ __device__ inline uint64_t ordered_key(double d) {
    uint64_t u = __builtin_bit_cast(uint64_t, d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double ordered_value(uint64_t k) {
    uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, u);
}

}  // namespace hdm
