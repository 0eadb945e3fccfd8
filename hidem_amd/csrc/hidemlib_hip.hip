// HiDEMLibHip and the hdm_sim_* C entry points (TOML-driven setup + the
// reference-shaped timestep loop).
#include "hidemlib_hip.h"

#include <dlfcn.h>

#include <cmath>
#include <filesystem>
#include <cstdio>
#include <limits>
#include <new>

namespace hdm {

namespace {
template <class T> T arg_or(ArgParser &p, const std::string &name, T dflt) { return p.getArg<T>(name).value_or(dflt); }

// The TOML table at "section.key" when it is a {path, fieldID} map table
// (ExternalForceBuilder::createForce, ExternalForces.h:87-123), else null.
const TomlValue *map_spec(ArgParser &a, const std::string &name) {
    const size_t dot = name.find('.');
    const std::string sec = name.substr(0, dot), key = name.substr(dot + 1);
    if (!a.table.contains(sec) || !a.table.at(sec).contains(key)) return nullptr;
    const TomlValue &v = a.table.at(sec).at(key);
    return v.is_table() ? &v : nullptr;
}
}  // namespace

std::vector<MapFile> map_inputs(ArgParser &a) {
    std::vector<MapFile> out;
    // [[collisionGeometry]] type = "bed": base and friction fields of one table
    // (GeometryCollection::parseCollisionTable, GeometryCollection.cpp:281-289)
    if (a.table.contains("collisionGeometry")) {
        const TomlValue &arr = a.table.at("collisionGeometry");
        for (const TomlValue &g : arr.arr) {
            if (!g.contains("type") || g.at("type").s != "bed") continue;
            if (!g.contains("path") || !g.contains("baseFieldID") || !g.contains("frictionFieldID"))
                throw TomlError("collisionGeometry bed needs path, baseFieldID and frictionFieldID");
            out.push_back({HDM_MAP_BED, g.at("path").s, (int32_t)toml_as<int64_t>(g.at("baseFieldID"))});
            out.push_back({HDM_MAP_FRICTION, g.at("path").s, (int32_t)toml_as<int64_t>(g.at("frictionFieldID"))});
        }
    }
    const char *en[6] = {"externalForce.xForce", "externalForce.yForce", "externalForce.zForce",
                         "externalForce.uForce", "externalForce.vForce", "externalForce.wForce"};
    for (int k = 0; k < 6; k++) {
        const TomlValue *m = map_spec(a, en[k]);
        if (!m) continue;
        if (!m->contains("path") || !m->contains("fieldID")) throw TomlError(std::string(en[k]) + " needs path and fieldID");
        out.push_back({HDM_MAP_EXT + k, m->at("path").s, (int32_t)toml_as<int64_t>(m->at("fieldID"))});
    }
    return out;
}

void set_map_inputs(hdm_ctx *ctx, const std::vector<MapFile> &files) {
    hdm_map2d maps[HDM_MAP_COUNT] = {};
    const hdm_map2d *slots[HDM_MAP_COUNT] = {};
    int rc = HDM_OK;
    std::string failed;
    for (const MapFile &f : files) {
        hdm_map2d_free(&maps[f.slot]);
        rc = hdm_map2d_load(f.path.c_str(), f.fieldID, &maps[f.slot]);
        if (rc != HDM_OK) {
            failed = f.path;
            break;
        }
        slots[f.slot] = &maps[f.slot];
    }
    if (rc == HDM_OK) rc = hdm_set_maps(ctx, slots);
    for (auto &m : maps) hdm_map2d_free(&m);
    if (!failed.empty()) throw TomlError("cannot read map " + failed);
    if (rc != HDM_OK) throw TomlError(std::string("hdm_set_maps: ") + hdm_last_error(ctx));
}

// MainIceBlock.cpp:25-51 (defaults) and :99-121 (scaling / derived values);
// MainDem.cpp:116-138 is identical.  Evaluated in double, in the reference's
// expression order.
GeneralParams GeneralParams::read(ArgParser &a) {
    GeneralParams g;
    g.breakTime = arg_or<double>(a, "general.breakTime", 10.0);
    g.dt = arg_or<double>(a, "general.dt", 0.0001);
    g.scale = arg_or<double>(a, "general.scale", 1.0);
    g.beamWidth = arg_or<double>(a, "general.beamWidth", 0.6);
    g.E = arg_or<double>(a, "general.E", 1e+09);
    g.waterLevel = arg_or<double>(a, "general.waterLevel", 0.0);
    g.attractDistance = arg_or<double>(a, "general.attractDistance", 0.0);
    g.attractScaleForce = arg_or<double>(a, "general.attractScaleForce", 0.0);
    g.maxVel = arg_or<double>(a, "general.maxVel", 20);
    g.maxRot = arg_or<double>(a, "general.maxRot", 5);
    g.rhoWater = arg_or<double>(a, "general.rhoWater", 1030.0);
    g.rhoIce = arg_or<double>(a, "general.rhoIce", 910.0);
    g.bedStiffness = arg_or<double>(a, "general.bedStiffness", 2e+08);
    g.bedInteract = arg_or<double>(a, "general.bedInteract", 2.0);
    g.waterDrag = arg_or<double>(a, "general.waterDrag", 10);
    g.airDrag = arg_or<double>(a, "general.airDrag", 1);
    g.bedDrag = arg_or<double>(a, "general.bedDrag", 1000.0);
    g.g = arg_or<double>(a, "general.g", 9.81);
    g.damping = arg_or<double>(a, "general.damping", 1000.0);
    g.dampingRotation = arg_or<double>(a, "general.dampingRotation", 1000.0);
    g.maxLoad = arg_or<double>(a, "general.maxLoad", 0.35);
    g.bendBreak = arg_or<double>(a, "general.bendBreak", 0.35);
    const char *bn[6] = {"general.xBoundaryMin", "general.xBoundaryMax", "general.yBoundaryMin",
                         "general.yBoundaryMax", "general.zBoundaryMin", "general.zBoundaryMax"};
    for (int k = 0; k < 6; k++) g.bound[k] = arg_or<bool>(a, bn[k], false);
    const char *en[6] = {"externalForce.xForce", "externalForce.yForce", "externalForce.zForce",
                         "externalForce.uForce", "externalForce.vForce", "externalForce.wForce"};
    // a single value each; a {path, fieldID} table is a map (load_maps)
    for (int k = 0; k < 6; k++) g.ext[k] = map_spec(a, en[k]) ? 0.0 : arg_or<double>(a, en[k], 0.0);

    const double scale = g.scale;
    g.beamWidth *= scale;
    g.bedInteract *= scale;
    g.bedStiffness *= scale * scale;
    g.waterDrag *= scale * scale;
    g.airDrag *= scale * scale;
    g.bedDrag *= scale * scale;
    g.damping *= scale * scale;
    g.dampingRotation *= scale * scale;
    g.attractDistance *= scale;
    g.lift = (g.rhoWater - g.rhoIce) / g.rhoIce * g.g;
    g.invDt = 1.0 / g.dt;
    g.invDt2 = 1.0 / (2.0 * g.dt);
    g.dtSqr = std::pow(g.dt, 2.0);
    g.invDtSqr = 1.0 / std::pow(g.dt, 2.0);
    g.G = g.E / 2.0;
    g.W2 = g.beamWidth * g.beamWidth;
    g.W4 = g.beamWidth * g.beamWidth * g.beamWidth * g.beamWidth;
    g.scaleE = g.E * scale;
    return g;
}

hidem::HiDEMDefNeighForceFunc GeneralParams::neighForce() const {
    return hidem::HiDEMDefNeighForceFunc(invDt, attractDistance, attractScaleForce, damping, dampingRotation, scaleE);
}
hidem::HiDEMDefLinkForceFunc GeneralParams::linkForce() const {
    return hidem::HiDEMDefLinkForceFunc(invDt, damping, dampingRotation, E, W2, W4, G);
}
hidem::HiDEMDefLinkBreakFunc GeneralParams::linkBreak() const { return hidem::HiDEMDefLinkBreakFunc(bendBreak, maxLoad); }

hdm_integ_params GeneralParams::integrator(ArgParser &a) const {
    hdm_integ_params ip{};
    std::string kind = arg_or<std::string>(a, "hdm.integrator", "iceblock");
    if (kind != "iceblock" && kind != "glacier") throw TomlError("hdm.integrator must be \"iceblock\" or \"glacier\"");
    ip.variant = kind == "glacier" ? HDM_INTEG_GLACIER : HDM_INTEG_ICEBLOCK;
    ip.rhoIce = rhoIce;
    ip.g = this->g;
    ip.lift = lift;
    ip.waterLevel = waterLevel;
    ip.airDrag = airDrag;
    ip.waterDrag = waterDrag;
    ip.invDt = invDt;
    ip.invDt2 = invDt2;
    ip.dtSqr = dtSqr;
    ip.invDtSqr = invDtSqr;
    ip.dt = dt;
    ip.bedDrag = bedDrag;
    ip.bedInteract = bedInteract;
    ip.bedStiffness = bedStiffness;
    ip.frictionCap = 1e+5 * scale * scale;  // MainDem.cpp:181
    ip.bedZ = arg_or<double>(a, "hdm.bedZ", 0.0);
    ip.bedXEnd = arg_or<double>(a, "hdm.bedXEnd", 1e300);
    ip.deepZ = arg_or<double>(a, "hdm.deepZ", -1e300);
    ip.frictionXEnd = arg_or<double>(a, "hdm.frictionXEnd", 1e300);
    for (int k = 0; k < 6; k++) ip.ext[k] = ext[k];
    ip.maxVel = maxVel;
    ip.maxRot = maxRot;
    ip.velocityLimit = ip.variant == HDM_INTEG_GLACIER ? 1 : 0;
    for (int k = 0; k < 6; k++) ip.bound[k] = bound[k] ? 1 : 0;
    std::vector<double> lo = arg_or<std::vector<double>>(a, "hdm.lowerLimit", {0.0, 0.0, 0.0});
    std::vector<double> hi = arg_or<std::vector<double>>(a, "hdm.upperLimit", {0.0, 0.0, 0.0});
    if (lo.size() != 3 || hi.size() != 3) throw TomlError("hdm.lowerLimit/upperLimit need 3 values");
    for (int k = 0; k < 3; k++) {
        ip.lower[k] = lo[(size_t)k];
        ip.upper[k] = hi[(size_t)k];
    }
    return ip;
}

HiDEMLibHip::HiDEMLibHip(int32_t precision, int32_t device) : accel(precision, device) {}

// The configuration half of HiDEMLib::setup (HiDEMLib.h:707-761 and
// HiDEMGeometry::setupLattice, HiDEMGeometry.cpp:86-95): no device work.
SetupInfo configure(ArgParser &parser) {
    SetupInfo s;
    s.gp = GeneralParams::read(parser);
    const double scale = s.gp.scale;
    hdm_setup_info &o = s.info;
    o = hdm_setup_info{};
    o.endTimestep = arg_or<int>(parser, "iteration.timeSteps", 10000);
    double searchDistanceScale = arg_or<double>(parser, "grid.searchDistanceScale", 1.4);
    int regen = arg_or<int>(parser, "grid.gridRegenInterval", -1);

    std::string type = arg_or<std::string>(parser, "lattice.type", "");
    if (type != "FCC" && type != "File")
        throw TomlError("lattice.type: \"FCC\" (generated) or \"File\" (particle file) on this path");
    auto need = [&](const char *k) {
        auto v = parser.getArg<double>(k);
        if (!v) throw TomlError(std::string("missing ") + k);
        return *v;
    };
    // the largest radius sets the search distance (HiDEMGeometry.cpp:86-121)
    double radius = need("lattice.particleRadius") * scale;
    o.linkGenScale = need("lattice.linkGenScale");
    o.fcc.radius = radius;
    if (type == "File") {
        s.particleFile = arg_or<std::string>(parser, "lattice.path", "");
        if (s.particleFile.empty()) throw TomlError("lattice.path: the particle file of type \"File\"");
        const char *keys[4] = {"lattice.xCords", "lattice.yCords", "lattice.zCords", "lattice.radius"};
        const char *dflt[4] = {"xCord", "yCord", "zCord", "radius"};
        for (int k = 0; k < 4; k++) s.fileNames[k] = arg_or<std::string>(parser, keys[k], dflt[k]);
    } else {
    double spacing = need("lattice.fccSpacing") * scale;
    auto cells = parser.getArg<std::vector<int>>("hdm.cells");
    if (!cells || cells->size() != 3) throw TomlError("hdm.cells needs 3 values");
    std::vector<double> origin = arg_or<std::vector<double>>(parser, "hdm.origin", {0.0, 0.0, 0.0});
    if (origin.size() != 3) throw TomlError("hdm.origin needs 3 values");
    o.fcc.nx = (*cells)[0];
    o.fcc.ny = (*cells)[1];
    o.fcc.nz = (*cells)[2];
    o.fcc.a = spacing;
    for (int k = 0; k < 3; k++) o.fcc.origin[k] = origin[(size_t)k];
    }
    s.restartFile = arg_or<std::string>(parser, "restart.fromRestart", "");
    o.fcc.jitter = arg_or<double>(parser, "hdm.jitter", 0.0);
    o.fcc.seed = (uint64_t)arg_or<int64_t>(parser, "lattice.seed", 1);
    o.fcc.clampXBelow = arg_or<double>(parser, "hdm.clampXBelow", -std::numeric_limits<double>::infinity());
    std::string order = arg_or<std::string>(parser, "hdm.order", "lattice");
    if (order != "lattice" && order != "morton") throw TomlError("hdm.order must be \"lattice\" or \"morton\"");
    o.fcc.order = order == "morton" ? HDM_ORDER_MORTON : HDM_ORDER_LATTICE;
    o.spread = arg_or<double>(parser, "hdm.maxLoadSpread", 0.0);
    o.linkSeed = (uint64_t)arg_or<int64_t>(parser, "hdm.linkSeed", 7);
    o.threshold = s.gp.maxLoad;
    std::string prec = arg_or<std::string>(parser, "hdm.precision", "fp64");
    if (prec != "fp64" && prec != "mixed") throw TomlError("hdm.precision must be \"fp64\" or \"mixed\"");
    o.precision = prec == "mixed" ? HDM_PREC_MIXED : HDM_PREC_FP64;

    std::string pol = arg_or<std::string>(parser, "hdm.rebuildPolicy", "heuristic");
    int32_t policy;
    if (pol == "heuristic") policy = HDM_REBUILD_HEURISTIC;
    else if (pol == "exact") policy = HDM_REBUILD_EXACT;
    else if (pol == "interval") policy = HDM_REBUILD_INTERVAL;
    else if (pol == "learned") policy = HDM_REBUILD_LEARNED;
    else throw TomlError("hdm.rebuildPolicy must be heuristic, exact, interval or learned");
    if (regen != -1) policy = HDM_REBUILD_INTERVAL;

    // HiDEMLib.h:759-761
    double maxParticleDist = radius * 2.0;
    o.run.searchDistance = searchDistanceScale * maxParticleDist;
    o.run.interactDist = o.run.searchDistance - maxParticleDist;
    o.run.force = hidem::make_force_params(s.gp.neighForce(), s.gp.linkForce());
    o.run.integ = s.gp.integrator(parser);
    o.run.brk = hdm_break_params{s.gp.bendBreak, s.gp.maxLoad};
    o.run.dt = s.gp.dt;
    o.run.breakTime = s.gp.breakTime;
    o.run.policy = policy;
    o.run.regenInterval = regen;
    return s;
}

void HiDEMLibHip::setup(ArgParser &parser, double scale) {
    SetupInfo s = configure(parser);
    if (scale != s.gp.scale) throw TomlError("setup: scale differs from general.scale");
    maxMinutes = arg_or<unsigned int>(parser, "iteration.maxMinutes", (unsigned int)-1);
    readCheckpointConfig(parser);
    readDistributionConfig(parser);
    fusedSteps = arg_or<bool>(parser, "hdm.fusedSteps", true);
    setup(s);
}

namespace {

// libhidem_io.so, loaded on first use from the directory of this library
// (the compute library itself does not link HDF5).
struct IoLib {
    void *h = nullptr;
    const char *(*lastError)(void) = nullptr;
    int (*ckptInfo)(const char *, int64_t *, int64_t *, int32_t *) = nullptr;
    int (*ckptWrite)(const char *, int32_t, const hdm_particles_host *, const hdm_links_host *) = nullptr;
    int (*ckptRead)(const char *, hdm_particles_host *, hdm_links_host *) = nullptr;
    int (*fileCount)(const char *, const char *, int64_t *) = nullptr;
    int (*fileRead)(const char *, const char *, const char *, const char *, const char *, hdm_particles_host *) = nullptr;

    static IoLib &get() {
        static IoLib io;
        if (!io.h) {
            Dl_info info{};
            std::string dir = ".";
            if (dladdr(reinterpret_cast<void *>(&hdm_sim_create), &info) && info.dli_fname) {
                dir = info.dli_fname;
                const size_t slash = dir.rfind('/');
                dir = slash == std::string::npos ? "." : dir.substr(0, slash);
            }
            const std::string path = dir + "/libhidem_io.so";
            io.h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (!io.h) throw TomlError("cannot load " + path + " (HDF5 IO): " + dlerror());
            io.lastError = reinterpret_cast<const char *(*)(void)>(dlsym(io.h, "hdm_io_last_error"));
            io.ckptInfo = reinterpret_cast<decltype(io.ckptInfo)>(dlsym(io.h, "hdm_checkpoint_info"));
            io.ckptRead = reinterpret_cast<decltype(io.ckptRead)>(dlsym(io.h, "hdm_checkpoint_read"));
            io.ckptWrite = reinterpret_cast<decltype(io.ckptWrite)>(dlsym(io.h, "hdm_checkpoint_write"));
            io.fileCount = reinterpret_cast<decltype(io.fileCount)>(dlsym(io.h, "hdm_particle_file_count"));
            io.fileRead = reinterpret_cast<decltype(io.fileRead)>(dlsym(io.h, "hdm_particle_file_read"));
            if (!io.lastError || !io.ckptInfo || !io.ckptRead || !io.ckptWrite || !io.fileCount || !io.fileRead)
                throw TomlError(path + " lacks the HDF5 IO entry points");
        }
        return io;
    }
    void check(int rc, const std::string &what) {
        if (rc != HDM_OK) throw TomlError(what + ": " + lastError());
    }
};

// Host storage behind hdm_particles_host / hdm_links_host views.
struct HostState {
    std::vector<double> pf[19];
    std::vector<uint8_t> type;
    std::vector<int32_t> a, b;
    std::vector<uint8_t> state;
    std::vector<double> len, maxLoad, ct[6];
    hdm_particles_host p{};
    hdm_links_host l{};
    void particles(int64_t n) {
        for (auto &v : pf) v.assign((size_t)n, 0.0);
        type.assign((size_t)n, 0);
        double **f[19] = {&p.xCord, &p.yCord, &p.zCord, &p.xTheta, &p.yTheta, &p.zTheta, &p.xDiff,
                          &p.yDiff, &p.zDiff, &p.uDiff, &p.vDiff, &p.wDiff, &p.radius, &p.xForce,
                          &p.yForce, &p.zForce, &p.uForce, &p.vForce, &p.wForce};
        for (int k = 0; k < 19; k++) *f[k] = pf[k].data();
        p.type = type.data();
        p.n = n;
    }
    void links(int64_t L, bool withCT) {
        a.assign((size_t)L, 0);
        b.assign((size_t)L, 0);
        state.assign((size_t)L, 1);
        len.assign((size_t)L, 0.0);
        maxLoad.assign((size_t)L, 0.0);
        for (auto &c : ct) c.assign(withCT ? (size_t)(2 * L) : 0, 0.0);
        l.n = L;
        l.localPartIndA = a.data();
        l.localPartIndB = b.data();
        l.state = state.data();
        l.initialLen = len.data();
        l.maxLoad = maxLoad.data();
        for (int c = 0; c < 6; c++) l.CT[c] = withCT ? ct[c].data() : nullptr;
    }
};

}  // namespace

void HiDEMLibHip::writeCheckpoint(const std::string &path) {
    IoLib &io = IoLib::get();
    HostState st;
    st.particles(nParticles);
    st.links(nLinks, true);
    accel.downloadParticles(st.p);
    hdm_links_host l = st.l;
    if (h_maxLoad.empty()) l.maxLoad = nullptr;  // no per-link thresholds bound
    accel.downloadLinks(l);
    io.check(io.ckptWrite(path.c_str(), (int32_t)currentTimestep, &st.p, &l), "hdm_checkpoint_write");
}

void HiDEMLibHip::setup(const SetupInfo &s) {
    if (s.restartFile.empty() && s.particleFile.empty()) {
        setup(s.info);
        return;
    }
    IoLib &io = IoLib::get();
    HostState st;
    int timestep = 0;
    if (!s.restartFile.empty()) {
        // HiDEMLib::setup from a checkpoint (ParticleDataHandler::fromCheckpoint,
        // LinkDataHandler::fromCheckpoint); per-link thresholds from the file
        // when it has them, else the scalar maxLoad
        int64_t n = 0, L = 0;
        int32_t ts = 0;
        io.check(io.ckptInfo(s.restartFile.c_str(), &n, &L, &ts), "hdm_checkpoint_info");
        st.particles(n);
        st.links(L, true);
        if (io.ckptRead(s.restartFile.c_str(), &st.p, &st.l) != HDM_OK) {
            st.l.maxLoad = nullptr;
            io.check(io.ckptRead(s.restartFile.c_str(), &st.p, &st.l), "hdm_checkpoint_read");
        }
        timestep = ts;
    } else {
        int64_t n = 0;
        io.check(io.fileCount(s.particleFile.c_str(), s.fileNames[0].c_str(), &n), "hdm_particle_file_count");
        st.particles(n);
        io.check(io.fileRead(s.particleFile.c_str(), s.fileNames[0].c_str(), s.fileNames[1].c_str(),
                             s.fileNames[2].c_str(), s.fileNames[3].c_str(), &st.p),
                 "hdm_particle_file_read");
        int64_t L = 0;
        hidem::hdm_check(nullptr, hdm_gen_links_count(&st.p, s.info.linkGenScale, &L), "hdm_gen_links_count");
        st.links(L, false);
        hidem::hdm_check(nullptr, hdm_gen_links_fill(&st.p, s.info.linkGenScale, s.info.threshold, s.info.spread,
                                                     s.info.linkSeed, &st.l),
                         "hdm_gen_links_fill");
    }
    setupState(s.info, st.p, st.l, timestep);
}

void HiDEMLibHip::setup(const hdm_setup_info &info) {
    runStart = std::chrono::steady_clock::now();
    endTimestep = (int)info.endTimestep;
    searchDistance = info.run.searchDistance;
    interactDist = info.run.interactDist;
    attractDistance = info.run.force.attractDistance;
    policy = info.run.policy;
    gridRegenInterval = info.run.regenInterval;
    hdm_fcc_spec spec = info.fcc;

    hidem::hdm_check(nullptr, hdm_gen_fcc_count(&spec, &nParticles), "hdm_gen_fcc_count");
    size_t n = (size_t)nParticles;
    for (auto &v : h_cords) v.assign(n, 0.0);
    h_radius.assign(n, 0.0);
    h_type.assign(n, 0);
    std::vector<double> th[3], df[6];
    for (auto &v : th) v.assign(n, 0.0);
    for (auto &v : df) v.assign(n, 0.0);
    hdm_particles_host p{};
    p.n = nParticles;
    p.xCord = h_cords[0].data();
    p.yCord = h_cords[1].data();
    p.zCord = h_cords[2].data();
    p.xTheta = th[0].data();
    p.yTheta = th[1].data();
    p.zTheta = th[2].data();
    p.xDiff = df[0].data();
    p.yDiff = df[1].data();
    p.zDiff = df[2].data();
    p.uDiff = df[3].data();
    p.vDiff = df[4].data();
    p.wDiff = df[5].data();
    p.radius = h_radius.data();
    p.type = h_type.data();
    hidem::hdm_check(nullptr, hdm_gen_fcc_fill(&spec, &p), "hdm_gen_fcc_fill");
    hidem::hdm_check(nullptr, hdm_gen_links_count(&p, info.linkGenScale, &nLinks), "hdm_gen_links_count");
    size_t L = (size_t)nLinks;
    h_a.assign(L, 0);
    h_b.assign(L, 0);
    h_len.assign(L, 0.0);
    h_maxLoad.assign(L, 0.0);
    std::vector<uint8_t> state(L, 1);
    hdm_links_host l{};
    l.n = nLinks;
    l.localPartIndA = h_a.data();
    l.localPartIndB = h_b.data();
    l.state = state.data();
    l.initialLen = h_len.data();
    l.maxLoad = h_maxLoad.data();
    hidem::hdm_check(nullptr, hdm_gen_links_fill(&p, info.linkGenScale, info.threshold, info.spread, info.linkSeed, &l),
                     "hdm_gen_links_fill");
    setupState(info, p, l, 0);
}

void HiDEMLibHip::setupState(const hdm_setup_info &info, const hdm_particles_host &p, const hdm_links_host &l,
                             int timestep) {
    runStart = std::chrono::steady_clock::now();
    endTimestep = (int)info.endTimestep;
    searchDistance = info.run.searchDistance;
    interactDist = info.run.interactDist;
    attractDistance = info.run.force.attractDistance;
    policy = info.run.policy;
    gridRegenInterval = info.run.regenInterval;
    const double radius = info.fcc.radius;
    nParticles = p.n;
    nLinks = l.n;
    const size_t n = (size_t)p.n;
    if (h_cords[0].data() != p.xCord) {
        const double *src[3] = {p.xCord, p.yCord, p.zCord};
        for (int k = 0; k < 3; k++) h_cords[k].assign(src[k], src[k] + n);
        h_radius.assign(p.radius, p.radius + n);
        h_type.assign(n, 0);
        if (p.type) h_type.assign(p.type, p.type + n);
        h_a.assign(l.localPartIndA, l.localPartIndA + l.n);
        h_b.assign(l.localPartIndB, l.localPartIndB + l.n);
        h_len.assign(l.initialLen, l.initialLen + l.n);
        h_maxLoad.clear();
        if (l.maxLoad) h_maxLoad.assign(l.maxLoad, l.maxLoad + l.n);
    }

    accel.uploadParticles(p);
    accel.uploadLinks(l);
    accel.generateLinks();

    hidem::Domain d;
    d.owner = 0;
    d.lower = hidem::Real3(INFINITY);
    d.upper = hidem::Real3(-INFINITY);
    for (size_t i = 0; i < n; i++) {
        d.lower = {std::fmin(d.lower.x, h_cords[0][i]), std::fmin(d.lower.y, h_cords[1][i]), std::fmin(d.lower.z, h_cords[2][i])};
        d.upper = {std::fmax(d.upper.x, h_cords[0][i]), std::fmax(d.upper.y, h_cords[1][i]), std::fmax(d.upper.z, h_cords[2][i])};
    }
    d.upper = d.upper + radius;
    d.lower = d.lower - radius;
    d.setHalo(searchDistance);
    world.domains = {d};
    currentTimestep = startTimestep = timestep;
    built = false;
    if (timestep != 0) hidem::hdm_check(accel.ctx, hdm_set_timestep(accel.ctx, timestep), "hdm_set_timestep");
}

bool HiDEMLibHip::run() const {
    auto mins = std::chrono::duration_cast<std::chrono::minutes>(std::chrono::steady_clock::now() - runStart).count();
    return currentTimestep < endTimestep && (unsigned long long)mins < (unsigned long long)maxMinutes;
}

void HiDEMLibHip::rebuild() {
    accumulatedDist = 0;
    accel.generateNeighbours(searchDistance, interactDist);
    hidem::hdm_check(accel.ctx, hdm_skin_snapshot(accel.ctx), "hdm_skin_snapshot");
    built = true;
    nRebuilds++;
}

// HiDEMLib::gridService (HiDEMLib.h:544-559, 589-627) with the build's exact
// displacement test as an alternative policy.
void HiDEMLibHip::gridService() {
    if (redistributionDue()) redistributeNow();
    bool needsGridRegen = currentTimestep == startTimestep || !built;
    if (policy == HDM_REBUILD_INTERVAL) {
        if (gridRegenInterval > 0 && (currentTimestep % gridRegenInterval) == 0) needsGridRegen = true;
    } else if (policy == HDM_REBUILD_HEURISTIC) {
        if (currentTimestep % iterationBlockSize == 0) {
            double d = 0;
            hidem::hdm_check(accel.ctx, hdm_skin_heuristic(accel.ctx, 10, &d), "hdm_skin_heuristic");
            accumulatedDist += d;
            if (accumulatedDist >= interactDist / 2.0) needsGridRegen = true;
        }
    } else if (built) {  // exact (and learned, whose device-side test is the exact one)
        double d = 0;
        hidem::hdm_check(accel.ctx, hdm_skin_exact(accel.ctx, &d), "hdm_skin_exact");
        if (2.0 * d >= interactDist - attractDistance) needsGridRegen = true;
    }
    lastRebuilt = needsGridRegen;
    if (needsGridRegen) rebuild();
}

void HiDEMLibHip::computePartForce(const hidem::HiDEMDefNeighForceFunc &n, const hidem::HiDEMDefLinkForceFunc &l) {
    accel.iterateCombinedPre(n, l);
    accel.iterateCombinedPost(n, l);
}

void HiDEMLibHip::applyForce(const hdm_integ_params &ip) { accel.applyForce(ip); }

void HiDEMLibHip::linkUpdate(const hidem::HiDEMDefLinkBreakFunc &b) { accel.iterateLinks(b); }

void HiDEMLibHip::finalizeTs() {
    currentTimestep++;
    if (statsDue()) writeStats();
    periodicCheckpoint();
}

bool HiDEMLibHip::statsDue() const {
    if (statIntervals.empty()) return false;
    for (int i : statIntervals)
        if (i == currentTimestep) return true;
    return statIntervals.back() > 0 && currentTimestep % statIntervals.back() == 0;
}

int HiDEMLibHip::nextStats() const {
    int next = std::numeric_limits<int>::max();
    if (statIntervals.empty()) return next;
    for (int i : statIntervals)
        if (i > currentTimestep) next = std::min(next, i);
    const int last = statIntervals.back();
    if (last > 0) next = std::min(next, (currentTimestep / last + 1) * last);
    return next;
}

void HiDEMLibHip::writeStats() {
    hidem::hdm_check(accel.ctx, hdm_particle_stats(accel.ctx, nullptr, nullptr, &linksIterated, &neighboursIterated),
                     "hdm_particle_stats");
    statsTimestep = currentTimestep;
    if (printStats)
        std::printf("timestep done %d\niteratedLinks %lld links\niteratedNegibours %lld neighbours\n", currentTimestep,
                    (long long)linksIterated, (long long)neighboursIterated);
}

void HiDEMLibHip::readCheckpointConfig(ArgParser &parser) {
    writeEndCheckpoint = arg_or<bool>(parser, "restart.writeEndCheckpoint", false);
    timestepCheckpoint = arg_or<bool>(parser, "restart.timestepCheckpoint", false);
    keepOldCheckpoint = arg_or<bool>(parser, "restart.keepOld", false);
    checkpointInterval = arg_or<int>(parser, "restart.checkpointInterval", 10000);
    checkpointDirectory = arg_or<std::string>(parser, "restart.checkpointDirectory", "checkpoints");
    outWorkingPath = arg_or<std::string>(parser, "output.outWorkingPath", "");
    statIntervals = arg_or<std::vector<int>>(parser, "output.statIntervals", statIntervals);
    printStats = arg_or<bool>(parser, "hdm.printStats", true);
}

static void write_named_checkpoint(HiDEMLibHip &h) {
    // an empty working path means the current directory (the reference
    // would write under the filesystem root)
    const std::string dir = (h.outWorkingPath.empty() ? std::string(".") : h.outWorkingPath) + "/" +
                            h.checkpointDirectory + "/";
    const std::string suffix = h.timestepCheckpoint ? "_" + std::to_string(h.getCurrentTimestep()) : "";
    const std::string target = dir + "checkpoint" + suffix + ".h5";
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (h.keepOldCheckpoint && std::filesystem::exists(target))
        std::filesystem::rename(target, dir + "checkpoint" + suffix + "_old.h5", ec);
    h.writeCheckpoint(target);
}

void HiDEMLibHip::readDistributionConfig(ArgParser &parser) {
    redistribute = arg_or<bool>(parser, "hdm.redistribute", false);
    redistributionInterval = arg_or<int>(parser, "distribution.redistributionInterval", 100);
    firstBalanceTimestep = arg_or<int>(parser, "distribution.firstBalanceTimestep", 2000);
    dynamicRedistribution = arg_or<bool>(parser, "distribution.dynamicRedistribution", true);
}

bool HiDEMLibHip::redistributionDue() const {
    if (!redistribute) return false;
    if (currentTimestep == startTimestep || currentTimestep == firstBalanceTimestep) return true;
    return !dynamicRedistribution && redistributionInterval > 0 && currentTimestep % redistributionInterval == 0;
}

int HiDEMLibHip::nextRedistribution() const {
    int next = std::numeric_limits<int>::max();
    if (!redistribute) return next;
    if (firstBalanceTimestep > currentTimestep) next = firstBalanceTimestep;
    if (!dynamicRedistribution && redistributionInterval > 0)
        next = std::min(next, (currentTimestep / redistributionInterval + 1) * redistributionInterval);
    return next;
}

void HiDEMLibHip::redistributeNow() {
    hidem::hdm_check(accel.ctx, hdm_reorder(accel.ctx, searchDistance, nullptr, nullptr), "hdm_reorder");
    built = false;
}

void HiDEMLibHip::periodicCheckpoint() {
    if (checkpointInterval != -1 && checkpointInterval > 0 && currentTimestep != 0 &&
        currentTimestep % checkpointInterval == 0)
        write_named_checkpoint(*this);
}

void HiDEMLibHip::finalize() {
    if (writeEndCheckpoint && !finalized) write_named_checkpoint(*this);
    finalized = true;
}

}  // namespace hdm

// ---------------------------------------------------------------------------
// C entry points
struct hdm_sim {
    hdm::HiDEMLibHip lib;
    hdm::ArgParser args;
    hdm::GeneralParams gp;
    hdm_integ_params integ;
    std::string dump;
    explicit hdm_sim(int32_t prec) : lib(prec) {}
};

static thread_local std::string g_sim_error;

extern "C" {

const char *hdm_sim_last_error(void) { return g_sim_error.c_str(); }

int hdm_config_derive(int32_t argc, const char *const *argv, hdm_setup_info *out) {
    if (!out || argc < 1 || !argv) return HDM_E_INVALID;
    try {
        std::vector<const char *> av{"hdm_config"};
        for (int i = 0; i < argc; i++) av.push_back(argv[i]);
        hdm::ArgParser args((int)av.size(), av.data());
        *out = hdm::configure(args).info;
        return HDM_OK;
    } catch (const std::exception &e) {
        g_sim_error = e.what();
        return HDM_E_INVALID;
    }
}

int hdm_sim_create(int32_t precision, int32_t argc, const char *const *argv, hdm_sim **out) {
    if (!out || argc < 1 || !argv) return HDM_E_INVALID;
    *out = nullptr;
    hdm_ctx *probe = nullptr;
    int rc = hdm_ctx_create(precision, 0, &probe);
    if (rc != HDM_OK) {
        g_sim_error = "no usable gfx950 device";
        return rc;
    }
    hdm_ctx_destroy(probe);
    try {
        std::vector<const char *> av{"hdm_sim"};
        for (int i = 0; i < argc; i++) av.push_back(argv[i]);
        hdm::ArgParser args((int)av.size(), av.data());
        hdm::SetupInfo si = hdm::configure(args);
        hdm_sim *s = new hdm_sim(precision);
        s->args = std::move(args);
        s->gp = si.gp;
        s->integ = si.info.run.integ;
        s->lib.maxMinutes = s->args.getArg<unsigned int>("iteration.maxMinutes").value_or((unsigned int)-1);
        s->lib.setup(si);
        s->lib.readCheckpointConfig(s->args);
        // hdm.redistribute (off by default: the generator already emits a
        // spatial order): the redistribution sort of the reference's
        // gridService (HiDEMLib.h:560-587), from the first step on
        s->lib.readDistributionConfig(s->args);
        if (s->integ.variant == HDM_INTEG_GLACIER) {
            std::vector<hdm::MapFile> files = hdm::map_inputs(s->args);
            if (!files.empty()) {
                try {
                    hdm::set_map_inputs(s->lib.accel.ctx, files);
                } catch (...) {
                    delete s;
                    throw;
                }
            }
        }
        *out = s;
        return HDM_OK;
    } catch (const std::exception &e) {
        g_sim_error = e.what();
        return HDM_E_INVALID;
    }
}

int hdm_sim_destroy(hdm_sim *s) {
    delete s;
    return HDM_OK;
}

int hdm_sim_ctx(hdm_sim *s, hdm_ctx **ctx) {
    if (!s || !ctx) return HDM_E_INVALID;
    *ctx = s->lib.accel.ctx;
    return HDM_OK;
}

int hdm_sim_run_params(hdm_sim *s, hdm_run_params *rp) {
    if (!s || !rp) return HDM_E_INVALID;
    rp->force = hidem::make_force_params(s->gp.neighForce(), s->gp.linkForce());
    rp->integ = s->integ;
    rp->brk = hdm_break_params{s->gp.bendBreak, s->gp.maxLoad};
    rp->searchDistance = s->lib.searchDistance;
    rp->interactDist = s->lib.interactDist;
    rp->dt = s->gp.dt;
    rp->breakTime = s->gp.breakTime;
    rp->policy = s->lib.policy;
    rp->regenInterval = s->lib.gridRegenInterval;
    return HDM_OK;
}

int hdm_sim_checkpoint(hdm_sim *s, const char *path) {
    if (!s || !path) return HDM_E_INVALID;
    try {
        s->lib.writeCheckpoint(path);
        return HDM_OK;
    } catch (const std::exception &e) {
        g_sim_error = e.what();
        return HDM_E_INVALID;
    }
}

int hdm_sim_info(hdm_sim *s, int64_t *nParticles, int64_t *nLinks, int64_t *timestep, int64_t *endTimestep) {
    if (!s) return HDM_E_INVALID;
    if (nParticles) *nParticles = s->lib.numParticles();
    if (nLinks) *nLinks = s->lib.numLinks();
    if (timestep) *timestep = s->lib.getCurrentTimestep();
    if (endTimestep) *endTimestep = s->lib.endTimestep;
    return HDM_OK;
}

// The loop of MainIceBlock.cpp:409-447, nSteps iterations (or until run()
// is false).  rebuildLog may be NULL.
int hdm_sim_steps(hdm_sim *s, int64_t nSteps, uint8_t *rebuildLog, hdm_run_stats *stats) {
    if (!s || nSteps < 0) return HDM_E_INVALID;
    auto &h = s->lib;
    auto nF = s->gp.neighForce();
    auto lF = s->gp.linkForce();
    auto bF = s->gp.linkBreak();
    auto t0 = std::chrono::steady_clock::now();
    int64_t r0 = h.rebuilds(), done = 0;
    if (h.fusedSteps) {
        // the same loop as one hdm_run call: one fused launch per step, same results
        int64_t n = nSteps;
        if (n > (int64_t)h.endTimestep - h.getCurrentTimestep()) n = (int64_t)h.endTimestep - h.getCurrentTimestep();
        if (n < 0) n = 0;
        hdm_run_params rp;
        int rc = hdm_sim_run_params(s, &rp);
        if (rc != HDM_OK) return rc;
        hdm_run_stats total{};
        total.timestep = h.getCurrentTimestep();
        try {
            // runs end at the periodic checkpoint steps, where the state is written
            while (done < n) {
                int64_t k = n - done;
                const int64_t iv = h.checkpointInterval;
                if (iv > 0) {
                    const int64_t next = ((int64_t)h.getCurrentTimestep() / iv + 1) * iv;
                    k = std::min<int64_t>(k, next - h.getCurrentTimestep());
                }
                // runs also end where a redistribution is due and after steps that print statistics
                if (h.redistributionDue()) h.redistributeNow();
                k = std::min<int64_t>(k, (int64_t)h.nextRedistribution() - h.getCurrentTimestep());
                k = std::min<int64_t>(k, (int64_t)h.nextStats() - h.getCurrentTimestep());
                hdm_run_stats st{};
                rc = hdm_run(h.accel.ctx, &rp, k, rebuildLog ? rebuildLog + done : nullptr, &st);
                if (rc != HDM_OK) {
                    g_sim_error = hdm_last_error(h.accel.ctx);
                    return rc;
                }
                h.syncFromRun(st);
                if (h.statsDue()) h.writeStats();
                h.periodicCheckpoint();
                done += st.steps;
                total.steps += st.steps;
                total.rebuilds += st.rebuilds;
                total.seconds += st.seconds;
                total.timestep = st.timestep;
                total.brokenTotal = st.brokenTotal;
                if (st.steps < k) break;
            }
            if (!h.run()) h.finalize();
        } catch (const std::exception &e) {
            g_sim_error = e.what();
            return HDM_E_INVALID;
        }
        if (stats) *stats = total;
        return HDM_OK;
    }
    try {
        for (; done < nSteps && h.run(); done++) {
            h.gridService();
            if (rebuildLog) rebuildLog[done] = h.lastStepRebuilt() ? 1 : 0;
            h.computePartForce(nF, lF);
            h.applyForce(s->integ);
            if (h.getCurrentTimestep() * s->gp.dt > s->gp.breakTime) h.linkUpdate(bF);
            h.finalizeTs();
        }
        h.accel.sync();
        if (!h.run()) h.finalize();
    } catch (const std::exception &e) {
        g_sim_error = e.what();
        return HDM_E_INVALID;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->timestep = h.getCurrentTimestep();
        stats->steps = done;
        stats->rebuilds = h.rebuilds() - r0;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        int64_t b = 0;
        int rc = hdm_count_broken(h.accel.ctx, &b);
        if (rc != HDM_OK) return rc;
        stats->brokenTotal = b;
    }
    return HDM_OK;
}

const char *hdm_sim_config_dump(hdm_sim *s) {
    if (!s) return "";
    s->dump = s->args.table.dump();
    return s->dump.c_str();
}

}  // extern "C"
