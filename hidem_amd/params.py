"""Derives the per-step kernel parameters from a TOML config.

Follows the parameter handling of the reference mains: defaults of
MainIceBlock.cpp:25-51 / MainDem.cpp (same keys), scaling of
MainIceBlock.cpp:99-121 (MainDem.cpp:116-138), searchDistance / interactDist
of HiDEMLib::setup (HiDEMLib.h:744-761) and the lattice scaling of
HiDEMGeometry::setupLattice (HiDEMGeometry.cpp:86-95).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

from . import _abi as A

try:
    import tomllib as _toml  # Python >= 3.11
except ModuleNotFoundError:  # pragma: no cover
    import tomli as _toml

CONFIG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs")

GENERAL_DEFAULTS = {
    "breakTime": 10.0, "dt": 0.0001, "scale": 1.0, "beamWidth": 0.6, "E": 1e9, "waterLevel": 0.0,
    "attractDistance": 0.0, "attractScaleForce": 0.0, "maxVel": 20.0, "maxRot": 5.0,
    "rhoWater": 1030.0, "rhoIce": 910.0, "bedStiffness": 2e8, "bedInteract": 2.0, "waterDrag": 10.0,
    "airDrag": 1.0, "bedDrag": 1000.0, "g": 9.81, "damping": 1000.0, "dampingRotation": 1000.0,
    "maxLoad": 0.35, "bendBreak": 0.35,
}

POLICIES = {"heuristic": A.HDM_REBUILD_HEURISTIC, "exact": A.HDM_REBUILD_EXACT, "interval": A.HDM_REBUILD_INTERVAL}


EXT_FORCES = ("xForce", "yForce", "zForce", "uForce", "vForce", "wForce")


def map_inputs(cfg: dict) -> list:
    """GlacierDem map slots a config names (hdm_set_maps order: bed, friction,
    external force x..w; None where the analytic input applies): the bed's
    base and friction fields of a [[collisionGeometry]] table with type
    "bed", external forces given as {path, fieldID} tables.  Paths as written
    (relative to the working directory, like the reference)."""
    slots = [None] * A.HDM_MAP_COUNT
    for g in cfg.get("collisionGeometry", []):
        if g.get("type") == "bed":
            slots[A.HDM_MAP_BED] = A.load_map(g["path"], int(g["baseFieldID"]))
            slots[A.HDM_MAP_FRICTION] = A.load_map(g["path"], int(g["frictionFieldID"]))
    ext = cfg.get("externalForce", {})
    for k, name in enumerate(EXT_FORCES):
        v = ext.get(name)
        if isinstance(v, dict):
            slots[A.HDM_MAP_EXT + k] = A.load_map(v["path"], int(v["fieldID"]))
    return slots


def load_config(name_or_path: str) -> dict:
    path = name_or_path
    if not os.path.exists(path):
        path = os.path.join(CONFIG_DIR, name_or_path)
        if not path.endswith(".toml"):
            path += ".toml"
    with open(path, "rb") as f:
        return _toml.load(f)


@dataclass
class Setup:
    cfg: dict
    force: A.HdmForceParams
    integ: A.HdmIntegParams
    brk: A.HdmBreakParams
    run: A.HdmRunParams
    radius: float
    a: float
    search_distance: float
    interact_dist: float
    link_gen_scale: float
    dt: float
    break_time: float
    time_steps: int


def derive(cfg: dict) -> Setup:
    g = dict(GENERAL_DEFAULTS)
    g.update(cfg.get("general", {}))
    hdm = cfg.get("hdm", {})
    scale = g["scale"]
    beamWidth = g["beamWidth"] * scale
    bedInteract = g["bedInteract"] * scale
    bedStiffness = g["bedStiffness"] * (scale * scale)  # "x *= scale * scale"
    waterDrag = g["waterDrag"] * (scale * scale)  # "x *= scale * scale"
    airDrag = g["airDrag"] * (scale * scale)  # "x *= scale * scale"
    bedDrag = g["bedDrag"] * (scale * scale)  # "x *= scale * scale"
    damping = g["damping"] * (scale * scale)  # "x *= scale * scale"
    dampingRotation = g["dampingRotation"] * (scale * scale)  # "x *= scale * scale"
    attractDistance = g["attractDistance"] * scale
    lift = (g["rhoWater"] - g["rhoIce"]) / g["rhoIce"] * g["g"]
    dt = g["dt"]
    invDt = 1.0 / dt
    invDt2 = 1.0 / (2.0 * dt)
    dtSqr = math.pow(dt, 2.0)
    invDtSqr = 1.0 / math.pow(dt, 2.0)
    E = g["E"]
    G = E / 2.0
    W2 = beamWidth * beamWidth
    W4 = beamWidth * beamWidth * beamWidth * beamWidth
    scaleE = E * scale

    fp = A.HdmForceParams(invDt, attractDistance, g["attractScaleForce"], damping, dampingRotation, scaleE,
                          E, W2, W4, G)
    ip = A.HdmIntegParams()
    ip.variant = A.HDM_INTEG_GLACIER if hdm.get("integrator", "iceblock") == "glacier" else A.HDM_INTEG_ICEBLOCK
    ip.rhoIce, ip.g, ip.lift, ip.waterLevel = g["rhoIce"], g["g"], lift, g["waterLevel"]
    ip.airDrag, ip.waterDrag = airDrag, waterDrag
    ip.invDt, ip.invDt2, ip.dtSqr, ip.invDtSqr, ip.dt = invDt, invDt2, dtSqr, invDtSqr, dt
    ip.bedDrag, ip.bedInteract, ip.bedStiffness = bedDrag, bedInteract, bedStiffness
    ip.frictionCap = 1e5 * scale * scale
    ip.bedZ = hdm.get("bedZ", 0.0)
    ip.bedXEnd = hdm.get("bedXEnd", 1e300)
    ip.deepZ = hdm.get("deepZ", -1e300)
    ip.frictionXEnd = hdm.get("frictionXEnd", 1e300)
    ext = cfg.get("externalForce", {})
    for k, name in enumerate(EXT_FORCES):
        v = ext.get(name, 0.0)
        ip.ext[k] = 0.0 if isinstance(v, dict) else float(v)  # a {path, fieldID} table is a map (map_inputs)
    ip.maxVel, ip.maxRot = g["maxVel"], g["maxRot"]
    ip.velocityLimit = 1 if ip.variant == A.HDM_INTEG_GLACIER else 0
    bounds = ("xBoundaryMin", "xBoundaryMax", "yBoundaryMin", "yBoundaryMax", "zBoundaryMin", "zBoundaryMax")
    for k, b in enumerate(bounds):
        ip.bound[k] = 1 if g.get(b, False) else 0
    lower = hdm.get("lowerLimit", [0.0, 0.0, 0.0])
    upper = hdm.get("upperLimit", [0.0, 0.0, 0.0])
    for k in range(3):
        ip.lower[k], ip.upper[k] = lower[k], upper[k]

    bp = A.HdmBreakParams(g["bendBreak"], g["maxLoad"])

    lat = cfg["lattice"]
    radius = lat["particleRadius"] * scale
    a = lat["fccSpacing"] * scale
    grid = cfg.get("grid", {})
    maxParticleDist = radius * 2.0
    search = grid.get("searchDistanceScale", 1.4) * maxParticleDist
    interact = search - maxParticleDist
    regen = int(grid.get("gridRegenInterval", -1))
    policy = POLICIES[hdm.get("rebuildPolicy", "heuristic")]
    if regen != -1:
        policy = A.HDM_REBUILD_INTERVAL

    rp = A.HdmRunParams()
    rp.force, rp.integ, rp.brk = fp, ip, bp
    rp.searchDistance, rp.interactDist = search, interact
    rp.dt, rp.breakTime = dt, g["breakTime"]
    rp.policy, rp.regenInterval = policy, regen

    return Setup(cfg, fp, ip, bp, rp, radius, a, search, interact, lat["linkGenScale"], dt, g["breakTime"],
                 int(cfg.get("iteration", {}).get("timeSteps", 10000)))


def build_inputs(cfg: dict, cells=None):
    """Generates the seeded particles and links of a config (optionally with
    fewer unit cells, for parity cases the CPU oracle finishes quickly)."""
    s = derive(cfg)
    hdm = cfg.get("hdm", {})
    nx, ny, nz = cells if cells is not None else hdm["cells"]
    origin = hdm.get("origin", [0.0, 0.0, 0.0])
    order = {"lattice": A.HDM_ORDER_LATTICE, "morton": A.HDM_ORDER_MORTON}[hdm.get("order", "lattice")]
    p = A.gen_fcc(nx, ny, nz, s.a, s.radius, hdm.get("jitter", 0.0), int(cfg["lattice"].get("seed", 1)),
                  origin=tuple(origin), clamp_x_below=hdm.get("clampXBelow", -np.inf), order=order)
    l = A.gen_links(p, s.link_gen_scale, cfg["general"].get("maxLoad", 0.35), hdm.get("maxLoadSpread", 0.0),
                    int(hdm.get("linkSeed", 7)))
    return s, p, l
