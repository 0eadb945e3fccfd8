// HiDEMLibHip: the timestep API of HiDEMLib (HiDEMLib/HiDEMLib.h:45-960) on
// the gfx950 accelerator, for one GPU (one rank, World of one Domain).
//
// Same call sequence as the reference mains (MainIceBlock.cpp:409-447):
//   while (hiDem.run()) {
//       hiDem.gridService();
//       hiDem.computePartForce(HiDEMDefNeighForceFunc(...), HiDEMDefLinkForceFunc(...));
//       hiDem.applyForce(integrator);
//       if (hiDem.getCurrentTimestep() * dt > breakTime) hiDem.linkUpdate(HiDEMDefLinkBreakFunc(...));
//       hiDem.finalizeTs();
//   }
// setup() reads the reference's [iteration], [grid] and [lattice] keys
// (HiDEMLib.h:707-822) plus the build-defined [hdm] table (block size,
// jitter, clamp, threshold spread, rebuild policy) and generates the seeded
// FCC block and its links on the host.
#pragma once

#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hidem/CombinedAccelHip.h"
#include "toml_lite.h"

namespace hdm {

// [general] of MainIceBlock.cpp:25-51 / MainDem.cpp with the reference
// defaults, scaled as MainIceBlock.cpp:99-121.
struct GeneralParams {
    double breakTime, dt, scale, beamWidth, E, waterLevel, attractDistance, attractScaleForce, maxVel, maxRot;
    double rhoWater, rhoIce, bedStiffness, bedInteract, waterDrag, airDrag, bedDrag, g, damping, dampingRotation;
    double maxLoad, bendBreak;
    bool bound[6];
    double ext[6];
    // derived
    double lift, invDt, invDt2, dtSqr, invDtSqr, G, W2, W4, scaleE;

    static GeneralParams read(ArgParser &parser);
    hidem::HiDEMDefNeighForceFunc neighForce() const;
    hidem::HiDEMDefLinkForceFunc linkForce() const;
    hidem::HiDEMDefLinkBreakFunc linkBreak() const;
    // IceBlock (variant HDM_INTEG_ICEBLOCK) or GlacierDem (HDM_INTEG_GLACIER)
    // applyForce parameters; bed geometry from the [hdm] table.
    hdm_integ_params integrator(ArgParser &parser) const;
};

// GlacierDem map inputs named by a config: the bed's base and friction fields
// ([[collisionGeometry]] type = "bed") and external forces given as
// {path, fieldID} tables ([externalForce]); slot = HDM_MAP_*.
struct MapFile {
    int slot;
    std::string path;
    int32_t fieldID;
};
std::vector<MapFile> map_inputs(ArgParser &parser);
// Loads the files (hdm_map2d_load) and binds them to the context
// (hdm_set_maps); throws TomlError naming what failed.
void set_map_inputs(hdm_ctx *ctx, const std::vector<MapFile> &files);

// The pure part of HiDEMLib::setup: everything derived from the config.
struct SetupInfo {
    hdm_setup_info info;
    GeneralParams gp;
    // [lattice] type = "File": particle input file and its dataset names
    // (HiDEMGeometry.cpp:111-121); [restart] fromRestart: a checkpoint to
    // start from.  Both read through libhidem_io.so.
    std::string particleFile;
    std::string fileNames[4];
    std::string restartFile;
};
SetupInfo configure(ArgParser &parser);

class HiDEMLibHip {
public:
    explicit HiDEMLibHip(int32_t precision = HDM_PREC_FP64, int32_t device = 0);

    hidem::CombinedAccelHip accel;
    hidem::World world;

    // HiDEMLib::setup (HiDEMLib.h:707-822)
    void setup(ArgParser &parser, double scale);
    void setup(const hdm_setup_info &info);
    // the generator-free part of setup: uploads a given state (host arrays
    // are copied into the h_* members first) at the given step
    void setupState(const hdm_setup_info &info, const hdm_particles_host &p, const hdm_links_host &l, int timestep);
    // setup from a SetupInfo that may name a particle file or a checkpoint
    void setup(const SetupInfo &s);
    // checkpoint of the current state and step (hdm_sim_checkpoint)
    void writeCheckpoint(const std::string &path);
    // [restart] keys (HiDEMLib.h:729-733) and the writes they ask for:
    // periodic from finalizeTs (HiDEMLib.h:536-541), at the end from
    // finalize (:476-484), to outWorkingPath/checkpointDirectory/
    // checkpoint[_<step>].h5 (HiDEMH5Checkpoint.h:23-49)
    void readCheckpointConfig(ArgParser &parser);
    void periodicCheckpoint();
    void finalize();
    bool writeEndCheckpoint = false, timestepCheckpoint = false, keepOldCheckpoint = false, finalized = false;
    // hdm.redistribute with the [distribution] keys (HiDEMLib.h:560-587,
    // 725-727) on one rank: the redistribution order (hdm_reorder) on the
    // first step, on firstBalanceTimestep and, unless dynamicRedistribution,
    // every redistributionInterval steps (the dynamic bounds check concerns
    // domain boundaries, which one rank does not have)
    void readDistributionConfig(ArgParser &parser);
    bool redistributionDue() const;
    int nextRedistribution() const;   // first due step after the current one (INT_MAX: none)
    void redistributeNow();
    bool redistribute = false, dynamicRedistribution = true;
    int redistributionInterval = 100, firstBalanceTimestep = 2000;
    int checkpointInterval = 10000;
    std::string checkpointDirectory = "checkpoints", outWorkingPath;

    bool run() const;                 // HiDEMLib.h:236-240
    void gridService();               // HiDEMLib.h:533-628 (rebuild part)
    void computePartForce(const hidem::HiDEMDefNeighForceFunc &n, const hidem::HiDEMDefLinkForceFunc &l);
    void applyForce(const hdm_integ_params &ip);
    void linkUpdate(const hidem::HiDEMDefLinkBreakFunc &b);
    void finalizeTs();                // HiDEMLib.h:495-530
    // finalizeTs statistics (HiDEMLib.h:507-529): after the steps named in
    // [output] statIntervals and every multiple of the last one, the sums of
    // computeActiveLinks and computeIteratedNeighbours (hdm_particle_stats),
    // printed as the reference logs them and kept in the members below
    std::vector<int> statIntervals{100, 200, 1000, 5000, 10000, 20000, 50000};
    bool printStats = true;
    bool statsDue() const;
    int nextStats() const;            // first due step after the current one (INT_MAX: none)
    void writeStats();
    int statsTimestep = -1;
    int64_t linksIterated = 0, neighboursIterated = 0;
    int getCurrentTimestep() const { return currentTimestep; }

    int64_t numParticles() const { return nParticles; }
    int64_t numLinks() const { return nLinks; }
    int64_t rebuilds() const { return nRebuilds; }
    bool lastStepRebuilt() const { return lastRebuilt; }
    // after steps run through hdm_run on this context (fused mode)
    void syncFromRun(const hdm_run_stats &st) {
        currentTimestep = (int)st.timestep;
        nRebuilds += st.rebuilds;
        built = true;
    }
    bool fusedSteps = true;           // hdm.fusedSteps: hdm_sim_steps runs hdm_run

    double searchDistance = 0, interactDist = 0, attractDistance = 0;
    int32_t policy = HDM_REBUILD_HEURISTIC;
    int gridRegenInterval = -1;
    int iterationBlockSize = 10;
    int currentTimestep = 0, startTimestep = 0, endTimestep = 10000;
    unsigned int maxMinutes = (unsigned int)-1;
    double accumulatedDist = 0;

    // host copies of the generated state (checkpoint order), kept for callers
    // that want to inspect the initial condition
    std::vector<double> h_cords[3], h_radius;
    std::vector<uint8_t> h_type;
    std::vector<int32_t> h_a, h_b;
    std::vector<double> h_len, h_maxLoad;

private:
    int64_t nParticles = 0, nLinks = 0, nRebuilds = 0;
    bool lastRebuilt = false;
    bool built = false;
    std::chrono::steady_clock::time_point runStart;
    void rebuild();
};

}  // namespace hdm
