"""ctypes wrapper of the CPU oracle (oracle/hdm_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhdm_oracle.so")
LIB_FAST = os.path.join(HERE, "libhdm_oracle_fast.so")

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(f, _dp) for f in (
        "x", "y", "z", "tx", "ty", "tz", "dx", "dy", "dz", "du", "dv", "dw", "radius",
        "fx", "fy", "fz", "fu", "fv", "fw")] + [("type", _u8p)]


class Links(C.Structure):
    _fields_ = [("n", C.c_int64), ("a", _i32p), ("b", _i32p), ("state", _u8p), ("initialLen", _dp),
                ("maxLoad", _dp), ("ct", _dp * 6)]


class Grid(C.Structure):
    _fields_ = ([(f, C.c_double) for f in ("spacing", "xMin", "yMin", "zMin", "xMax", "yMax", "zMax")] +
                [("X", C.c_int32), ("Y", C.c_int32), ("Z", C.c_int32), ("totalCells", C.c_int64),
                 ("dirOffsets", C.c_int64 * 27), ("colours", _i64p), ("order", _i32p), ("nOcc", C.c_int64),
                 ("occColour", _i64p), ("bucketStart", _i32p), ("bucketSize", _i32p)])


class Neigh(C.Structure):
    _fields_ = [("n", C.c_int64), ("count", _i32p), ("start", _i64p), ("list", _i32p)]


class LinkLists(C.Structure):
    _fields_ = [("n", C.c_int64), ("count", _i32p), ("start", _i64p), ("linkId", _i32p), ("other", _i32p),
                ("side", _i32p)]


class ForceParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("invDt", "attractDistance", "attractScaleForce", "damping",
                                          "dampingRotation", "scaleE", "E", "W2", "W4", "G")]


class Map2d(C.Structure):
    _fields_ = [("xPoints", C.c_int32), ("yPoints", C.c_int32)] + \
               [(f, C.c_float) for f in ("xMin", "yMin", "xMax", "yMax", "xRes", "yRes", "invXMaxMin", "invYMaxMin",
                                         "xOffset", "yOffset")] + [("data", _dp)]


MAP_COUNT = 8


class IntegParams(C.Structure):
    _fields_ = ([("variant", C.c_int32)] +
                [(f, C.c_double) for f in ("rhoIce", "g", "lift", "waterLevel", "airDrag", "waterDrag",
                                           "invDt", "invDt2", "dtSqr", "invDtSqr", "dt",
                                           "bedDrag", "bedInteract", "bedStiffness", "frictionCap",
                                           "bedZ", "bedXEnd", "deepZ", "frictionXEnd")] +
                [("ext", C.c_double * 6), ("maxVel", C.c_double), ("maxRot", C.c_double),
                 ("velocityLimit", C.c_int32), ("bound", C.c_int32 * 6),
                 ("lower", C.c_double * 3), ("upper", C.c_double * 3),
                 ("blockId", C.POINTER(C.c_int32)), ("blockWater", C.POINTER(C.c_double)),
                 ("maps", C.POINTER(Map2d) * MAP_COUNT)])


class BreakParams(C.Structure):
    _fields_ = [("bendBreak", C.c_double), ("maxLoad", C.c_double)]


class Sim(C.Structure):
    _fields_ = [("p", C.POINTER(Particles)), ("l", C.POINTER(Links)), ("fp", ForceParams), ("ip", IntegParams),
                ("bp", BreakParams), ("searchDistance", C.c_double), ("interactDist", C.c_double),
                ("dt", C.c_double), ("breakTime", C.c_double), ("policy", C.c_int32),
                ("regenInterval", C.c_int32), ("ts", C.c_int64), ("accumulatedDist", C.c_double),
                ("rebuilds", C.c_int64), ("broken", C.c_int64), ("grid", Grid), ("nl", Neigh),
                ("ll", LinkLists), ("x0", _dp), ("y0", _dp), ("z0", _dp), ("built", C.c_int32)]


_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(fast=False):
    path = LIB_FAST if fast else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.hdmo_grid_build.argtypes = [C.POINTER(Particles), C.c_double, C.POINTER(Grid)]
        L.hdmo_grid_free.argtypes = [C.POINTER(Grid)]
        L.hdmo_neigh_build.argtypes = [C.POINTER(Particles), C.POINTER(Grid), C.c_double, C.POINTER(Neigh)]
        L.hdmo_neigh_free.argtypes = [C.POINTER(Neigh)]
        L.hdmo_link_lists.argtypes = [C.c_int64, C.POINTER(Links), C.POINTER(LinkLists)]
        L.hdmo_linklists_free.argtypes = [C.POINTER(LinkLists)]
        L.hdmo_force.argtypes = [C.POINTER(Particles), C.POINTER(Links), C.POINTER(Neigh), C.POINTER(LinkLists),
                                 C.POINTER(ForceParams)]
        L.hdmo_force.restype = None
        L.hdmo_contact_pair.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                        C.POINTER(ForceParams), _dp]
        L.hdmo_contact_pair.restype = None
        L.hdmo_beam_side.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                     C.POINTER(ForceParams), _dp, _dp]
        L.hdmo_beam_side.restype = None
        L.hdmo_integrate.argtypes = [C.POINTER(Particles), C.POINTER(IntegParams)]
        L.hdmo_integrate.restype = None
        L.hdmo_break.argtypes = [C.POINTER(Particles), C.POINTER(Links), C.POINTER(BreakParams)]
        L.hdmo_break.restype = C.c_int64
        L.hdmo_skin_heuristic.argtypes = [C.POINTER(Particles), C.c_int]
        L.hdmo_skin_heuristic.restype = C.c_double
        L.hdmo_skin_exact.argtypes = [C.POINTER(Particles), _dp, _dp, _dp]
        L.hdmo_skin_exact.restype = C.c_double
        L.hdmo_reorder.argtypes = [C.POINTER(Particles), C.POINTER(Links), C.c_double, _i64p, _i64p]
        L.hdmo_reorder.restype = C.c_int
        L.hdmo_map_init.argtypes = [C.POINTER(Map2d)]
        L.hdmo_map_init.restype = None
        L.hdmo_map_indexes.argtypes = [C.POINTER(Map2d), C.c_float, C.c_float, C.POINTER(C.c_int32)]
        L.hdmo_map_indexes.restype = None
        L.hdmo_map_interp.argtypes = [C.POINTER(Map2d), C.c_float, C.c_float]
        L.hdmo_map_interp.restype = C.c_double
        L.hdmo_sim_init.argtypes = [C.POINTER(Sim)]
        L.hdmo_sim_run.argtypes = [C.POINTER(Sim), C.c_int64, _u8p, _i64p, C.c_int64]
        L.hdmo_sim_run.restype = C.c_int64
        L.hdmo_sim_free.argtypes = [C.POINTER(Sim)]
        L.hdmo_sim_free.restype = None
        L.hdmo_num_threads.restype = C.c_int
        _libs[path] = L
    return _libs[path]


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


_PF = {"x": "xCord", "y": "yCord", "z": "zCord", "tx": "xTheta", "ty": "yTheta", "tz": "zTheta",
       "dx": "xDiff", "dy": "yDiff", "dz": "zDiff", "du": "uDiff", "dv": "vDiff", "dw": "wDiff",
       "radius": "radius", "fx": "xForce", "fy": "yForce", "fz": "zForce", "fu": "uForce", "fv": "vForce",
       "fw": "wForce"}


def particles_struct(p):
    """p: hidem_amd._abi.Particles-like (dict of float64 arrays + type)."""
    s = Particles()
    s.n = p.n
    for k, name in _PF.items():
        setattr(s, k, _p(p.arrays[name], C.c_double))
    s.type = _p(p.type, C.c_uint8)
    return s


def links_struct(l):
    s = Links()
    s.n = l.n
    s.a = _p(l.a, C.c_int32)
    s.b = _p(l.b, C.c_int32)
    s.state = _p(l.state, C.c_uint8)
    s.initialLen = _p(l.initialLen, C.c_double)
    s.maxLoad = _p(l.maxLoad, C.c_double)
    for k in range(6):
        s.ct[k] = _p(l.ct[k], C.c_double)
    return s


def reorder(p, l, search_distance):
    """hdmo_reorder on copies: (particles, links, new_order, link_order) with
    new_order[i] = old index of new particle i, link_order[j] = old index of
    new link j."""
    q, m = p.copy(), l.copy()
    ps, ls = particles_struct(q), links_struct(m)
    order = np.zeros(p.n, np.int64)
    lorder = np.zeros(max(l.n, 1), np.int64)
    rc = lib().hdmo_reorder(C.byref(ps), C.byref(ls), search_distance, _p(order, C.c_int64), _p(lorder, C.c_int64))
    if rc != 0:
        raise RuntimeError(f"hdmo_reorder failed: {rc}")
    return q, m, order, lorder[:l.n]


def force_params(fp):
    return ForceParams(*[getattr(fp, f) for f, _ in ForceParams._fields_])


def map_struct(values, x_min, y_min, x_max, y_max, x_offset=0.0, y_offset=0.0):
    """Map2d over values[iy, ix] (float64; kept alive by the returned pair),
    derived fields from hdmo_map_init."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    m = Map2d(v.shape[1], v.shape[0], x_min, y_min, x_max, y_max, 0, 0, 0, 0, x_offset, y_offset,
              v.ctypes.data_as(_dp))
    lib().hdmo_map_init(C.byref(m))
    return m, v


def map_interp(m, x, y):
    """Map2d::interp of one point (x, y rounded to float as the callers do)."""
    return lib().hdmo_map_interp(C.byref(m), C.c_float(x), C.c_float(y))


def map_indexes(m, x, y):
    out = (C.c_int32 * 4)()
    lib().hdmo_map_indexes(C.byref(m), C.c_float(x), C.c_float(y), out)
    return tuple(out)


def integ_params(ip, block_id=None, block_water=None, maps=None):
    """Copies an hdm_integ_params; optional per-particle block ids and
    per-block waterlines (numpy arrays kept alive by the caller) and the
    GlacierDem map slots (a list of 8 Map2d structs or None, kept alive by
    the caller)."""
    o = IntegParams()
    if block_id is not None:
        o.blockId = block_id.ctypes.data_as(C.POINTER(C.c_int32))
        o.blockWater = block_water.ctypes.data_as(C.POINTER(C.c_double))
    for k, m in enumerate(maps or []):
        if m is not None:
            o.maps[k] = C.pointer(m)
    for f, _ in IntegParams._fields_:
        if f in ("blockId", "blockWater", "maps"):
            continue
        v = getattr(ip, f)
        if f in ("ext", "bound", "lower", "upper"):
            arr = getattr(o, f)
            for k in range(len(arr)):
                arr[k] = v[k]
        else:
            setattr(o, f, v)
    return o


def break_params(bp):
    return BreakParams(bp.bendBreak, bp.maxLoad)


def grid_build(p, search_distance, fast=False):
    L = lib(fast)
    ps = particles_struct(p)
    g = Grid()
    rc = L.hdmo_grid_build(C.byref(ps), search_distance, C.byref(g))
    if rc != 0:
        raise RuntimeError(f"hdmo_grid_build failed: {rc}")
    n = p.n
    out = {
        "dims": (g.X, g.Y, g.Z), "mins": (g.xMin, g.yMin, g.zMin), "maxs": (g.xMax, g.yMax, g.zMax),
        "totalCells": g.totalCells, "dirOffsets": np.array(list(g.dirOffsets)),
        "colours": np.ctypeslib.as_array(g.colours, (n,)).copy(),
        "order": np.ctypeslib.as_array(g.order, (n,)).copy(),
        "nOcc": g.nOcc,
        "occColour": np.ctypeslib.as_array(g.occColour, (max(g.nOcc, 1),))[:g.nOcc].copy(),
        "bucketStart": np.ctypeslib.as_array(g.bucketStart, (max(g.nOcc, 1),))[:g.nOcc].copy(),
        "bucketSize": np.ctypeslib.as_array(g.bucketSize, (max(g.nOcc, 1),))[:g.nOcc].copy(),
    }
    return g, out


def neigh_build(p, search_distance, interact_dist):
    L = lib()
    ps = particles_struct(p)
    g, _ = grid_build(p, search_distance)
    nl = Neigh()
    rc = L.hdmo_neigh_build(C.byref(ps), C.byref(g), interact_dist, C.byref(nl))
    if rc != 0:
        raise RuntimeError(f"hdmo_neigh_build failed: {rc}")
    n = p.n
    counts = np.ctypeslib.as_array(nl.count, (n,)).copy()
    starts = np.ctypeslib.as_array(nl.start, (n + 1,)).copy()
    tot = int(starts[-1])
    lst = np.ctypeslib.as_array(nl.list, (max(tot, 1),))[:tot].copy()
    L.hdmo_neigh_free(C.byref(nl))
    L.hdmo_grid_free(C.byref(g))
    return counts, starts, lst


def link_lists(n_particles, l):
    L = lib()
    ls = links_struct(l)
    ll = LinkLists()
    rc = L.hdmo_link_lists(n_particles, C.byref(ls), C.byref(ll))
    if rc != 0:
        raise RuntimeError(f"hdmo_link_lists failed: {rc}")
    counts = np.ctypeslib.as_array(ll.count, (n_particles,)).copy()
    starts = np.ctypeslib.as_array(ll.start, (n_particles + 1,)).copy()
    tot = int(starts[-1])
    lid = np.ctypeslib.as_array(ll.linkId, (max(tot, 1),))[:tot].copy()
    oth = np.ctypeslib.as_array(ll.other, (max(tot, 1),))[:tot].copy()
    side = np.ctypeslib.as_array(ll.side, (max(tot, 1),))[:tot].copy()
    L.hdmo_linklists_free(C.byref(ll))
    return counts, starts, lid, oth, side


class OracleSim:
    """Owns copies of the particle/link arrays and steps them on the CPU."""

    def __init__(self, p, l, setup, fast=False, block_id=None, block_water=None, maps=None):
        """maps: GlacierDem map slots, 8 entries of (values, x_min, y_min,
        x_max, y_max) or None (bed, friction, external force x..w)."""
        self.L = lib(fast)
        self._blocks = (block_id, block_water)
        self._maps = [map_struct(*m) if m is not None else None for m in (maps or [])]
        self.p = p.copy()
        self.l = l.copy()
        self._ps = particles_struct(self.p)
        self._ls = links_struct(self.l)
        s = Sim()
        s.p = C.pointer(self._ps)
        s.l = C.pointer(self._ls)
        s.fp = force_params(setup.force)
        s.ip = integ_params(setup.integ, block_id, block_water, [m[0] if m else None for m in self._maps])
        s.bp = break_params(setup.brk)
        s.searchDistance = setup.search_distance
        s.interactDist = setup.interact_dist
        s.dt = setup.dt
        s.breakTime = setup.break_time
        s.policy = setup.run.policy
        s.regenInterval = setup.run.regenInterval
        self.s = s
        rc = self.L.hdmo_sim_init(C.byref(s))
        if rc != 0:
            raise RuntimeError(f"hdmo_sim_init failed: {rc}")

    def run(self, n_steps, log=False, broken_cap=1 << 20):
        rlog = np.zeros(max(n_steps, 1), np.uint8) if log else None
        blog = np.zeros(2 * broken_cap, np.int64) if log else None
        nb = self.L.hdmo_sim_run(C.byref(self.s), n_steps, _p(rlog, C.c_uint8), _p(blog, C.c_int64),
                                 broken_cap if log else 0)
        if log:
            nb = min(nb, broken_cap)
            return rlog[:n_steps], blog[:2 * nb].reshape(-1, 2)
        return None

    # single phases, on the sim's own structures
    def rebuild(self):
        s = self.s
        self.L.hdmo_grid_free(C.byref(s.grid))
        self.L.hdmo_neigh_free(C.byref(s.nl))
        if self.L.hdmo_grid_build(s.p, s.searchDistance, C.byref(s.grid)) != 0:
            raise RuntimeError("hdmo_grid_build failed")
        if self.L.hdmo_neigh_build(s.p, C.byref(s.grid), s.interactDist, C.byref(s.nl)) != 0:
            raise RuntimeError("hdmo_neigh_build failed")

    def force(self):
        self.L.hdmo_force(self.s.p, self.s.l, C.byref(self.s.nl), C.byref(self.s.ll), C.byref(self.s.fp))

    def integrate(self):
        self.L.hdmo_integrate(self.s.p, C.byref(self.s.ip))

    def brk(self):
        return self.L.hdmo_break(self.s.p, self.s.l, C.byref(self.s.bp))

    @property
    def ts(self):
        return self.s.ts

    @property
    def rebuilds(self):
        return self.s.rebuilds

    def close(self):
        if self.s is not None:
            self.L.hdmo_sim_free(C.byref(self.s))
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def num_threads(fast=False):
    return lib(fast).hdmo_num_threads()
