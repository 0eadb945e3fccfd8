"""ctypes bindings to libhidem_amd.so (the C ABI in include/hidem/hdm_abi.h).

Plumbing for tests and bench.py: numpy arrays on the host side, the library
owns every device buffer.  The product's compute path is the HIP library; this
module only marshals arguments and raises on non-zero status codes.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhidem_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "hidem", "hdm_abi.h")

HDM_OK = 0
HDM_PREC_FP64 = 0
HDM_PREC_MIXED = 1
HDM_TYPE_FIXED = 1
HDM_INTEG_ICEBLOCK = 0
HDM_INTEG_GLACIER = 1
HDM_REBUILD_HEURISTIC = 0
HDM_REBUILD_EXACT = 1
HDM_REBUILD_INTERVAL = 2
HDM_REBUILD_LEARNED = 3
SCHED_FEATURES = 16
SCHED_HIDDEN = 64
HDM_ORDER_LATTICE = 0
HDM_ORDER_MORTON = 1
HDM_PROBE_COPY = 0
HDM_PROBE_CT_RMW = 1
TIMER_NAMES = ("force", "integrate", "break", "skin", "grid", "neigh", "links", "step")

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)


class HdmParticlesHost(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(f, _dp) for f in (
        "xCord", "yCord", "zCord", "xTheta", "yTheta", "zTheta",
        "xDiff", "yDiff", "zDiff", "uDiff", "vDiff", "wDiff", "radius",
        "xForce", "yForce", "zForce", "uForce", "vForce", "wForce")] + [("type", _u8p)]


class HdmLinksHost(C.Structure):
    _fields_ = [("n", C.c_int64), ("localPartIndA", _i32p), ("localPartIndB", _i32p),
                ("state", _u8p), ("initialLen", _dp), ("maxLoad", _dp), ("CT", _dp * 6)]


class HdmForceParams(C.Structure):
    _fields_ = [(f, C.c_double) for f in (
        "invDt", "attractDistance", "attractScaleForce", "damping", "dampingRotation", "scaleE",
        "E", "W2", "W4", "G")]


class HdmIntegParams(C.Structure):
    _fields_ = ([("variant", C.c_int32)] +
                [(f, C.c_double) for f in ("rhoIce", "g", "lift", "waterLevel", "airDrag", "waterDrag",
                                           "invDt", "invDt2", "dtSqr", "invDtSqr", "dt",
                                           "bedDrag", "bedInteract", "bedStiffness", "frictionCap",
                                           "bedZ", "bedXEnd", "deepZ", "frictionXEnd")] +
                [("ext", C.c_double * 6), ("maxVel", C.c_double), ("maxRot", C.c_double),
                 ("velocityLimit", C.c_int32), ("bound", C.c_int32 * 6),
                 ("lower", C.c_double * 3), ("upper", C.c_double * 3)])


class HdmBreakParams(C.Structure):
    _fields_ = [("bendBreak", C.c_double), ("maxLoad", C.c_double)]


class HdmGridInfo(C.Structure):
    _fields_ = ([(f, C.c_double) for f in ("spacing", "xMin", "yMin", "zMin", "xMax", "yMax", "zMax")] +
                [("X", C.c_int32), ("Y", C.c_int32), ("Z", C.c_int32), ("totalCells", C.c_int64),
                 ("occupiedCells", C.c_int64), ("maxPerCell", C.c_int32)])


class HdmRunParams(C.Structure):
    _fields_ = [("force", HdmForceParams), ("integ", HdmIntegParams), ("brk", HdmBreakParams),
                ("searchDistance", C.c_double), ("interactDist", C.c_double),
                ("dt", C.c_double), ("breakTime", C.c_double),
                ("policy", C.c_int32), ("regenInterval", C.c_int32)]


class HdmRunStats(C.Structure):
    _fields_ = [("timestep", C.c_int64), ("steps", C.c_int64), ("rebuilds", C.c_int64),
                ("brokenTotal", C.c_int64), ("seconds", C.c_double)]


class HdmFccSpec(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("a", C.c_double),
                ("origin", C.c_double * 3), ("radius", C.c_double), ("jitter", C.c_double),
                ("seed", C.c_uint64), ("clampXBelow", C.c_double), ("order", C.c_int32)]


class HdmSetupInfo(C.Structure):
    _fields_ = [("run", HdmRunParams), ("fcc", HdmFccSpec), ("linkGenScale", C.c_double),
                ("threshold", C.c_double), ("spread", C.c_double), ("linkSeed", C.c_uint64),
                ("endTimestep", C.c_int64), ("precision", C.c_int32)]


class HdmMlpWeights(C.Structure):
    _fields_ = [("W1", C.c_float * (16 * 64)), ("b1", C.c_float * 64), ("W2", C.c_float * (64 * 64)),
                ("b2", C.c_float * 64), ("w3", C.c_float * 64), ("b3", C.c_float),
                ("featMean", C.c_float * 16), ("featStd", C.c_float * 16)]

    N_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1

    def to_numpy(self):
        """Flat float32 copy in struct order (parameters first, then standardisation)."""
        return np.frombuffer(bytes(self), dtype=np.float32).copy()

    @classmethod
    def from_numpy(cls, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size * 4 == C.sizeof(cls)
        return cls.from_buffer_copy(flat.tobytes())


class HdmExecParams(C.Structure):
    _fields_ = [("autotune", C.c_int32), ("graphs", C.c_int32), ("variant", C.c_int32),
                ("listWithPartners", C.c_int32), ("persistent", C.c_int32)]


class HdmExecInfo(C.Structure):
    _fields_ = [("variant", C.c_int32), ("variants", C.c_int32), ("tuned", C.c_int32), ("name", C.c_char * 32),
                ("tuneMs", C.c_double * 8), ("graphCaptures", C.c_int64), ("graphReplays", C.c_int64),
                ("mergedContacts", C.c_int32), ("persistentLaunches", C.c_int64)]


class HdmSchedParams(C.Structure):
    _fields_ = [("blockCells", C.c_int32), ("minBatch", C.c_int32), ("maxBatch", C.c_int32), ("safety", C.c_float)]


class HdmMap2d(C.Structure):
    _fields_ = [("xPoints", C.c_int32), ("yPoints", C.c_int32)] + \
               [(f, C.c_float) for f in ("xMin", "yMin", "xMax", "yMax", "xOffset", "yOffset")] + \
               [("data", C.POINTER(C.c_double))]


HDM_MAP_BED, HDM_MAP_FRICTION, HDM_MAP_EXT, HDM_MAP_COUNT = 0, 1, 2, 8


@dataclass
class Map2d:
    """A 2-D field on a regular grid (hdm_map2d): values[iy, ix] over
    [x_min, x_max] x [y_min, y_max]; offsets as Map2d::setOffset."""
    values: np.ndarray
    x_min: float
    y_min: float
    x_max: float
    y_max: float
    x_offset: float = 0.0
    y_offset: float = 0.0

    @classmethod
    def constant(cls, v):
        """One point: the single-value form of the reference's maps."""
        return cls(np.array([[float(v)]]), 0.0, 0.0, 1.0, 1.0)

    def struct(self):
        """hdm_map2d over this map's values (kept alive by the returned pair)."""
        v = np.ascontiguousarray(self.values, dtype=np.float64)
        m = HdmMap2d(v.shape[1], v.shape[0], self.x_min, self.y_min, self.x_max, self.y_max,
                     self.x_offset, self.y_offset, v.ctypes.data_as(C.POINTER(C.c_double)))
        return m, v


def load_map(path: str, field: int) -> Map2d:
    """Map2dLoader::readInput + get through hdm_map2d_load."""
    L = lib()
    m = HdmMap2d()
    check(L.hdm_map2d_load(os.fsencode(path), field, C.byref(m)), what=f"hdm_map2d_load({path}, {field})")
    try:
        v = np.ctypeslib.as_array(m.data, (m.yPoints, m.xPoints)).copy()
    finally:
        L.hdm_map2d_free(C.byref(m))
    return Map2d(v, m.xMin, m.yMin, m.xMax, m.yMax)


class HdmMlpTrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float)]


_lib = None


def lib():
    """Loads libhidem_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build() (make -C hidem_amd/csrc)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = C.c_void_p
    sig = {
        "hdm_abi_version": ([], C.c_int),
        "hdm_device_count": ([_i32p], C.c_int),
        "hdm_ctx_create": ([C.c_int32, C.c_int32, C.POINTER(P)], C.c_int),
        "hdm_ctx_destroy": ([P], C.c_int),
        "hdm_last_error": ([P], C.c_char_p),
        "hdm_sync": ([P], C.c_int),
        "hdm_set_timing": ([P, C.c_int32], C.c_int),
        "hdm_kernel_times": ([P, _dp, _i64p], C.c_int),
        "hdm_reset_kernel_times": ([P], C.c_int),
        "hdm_device_bytes": ([P, _i64p], C.c_int),
        "hdm_particles_upload": ([P, C.POINTER(HdmParticlesHost)], C.c_int),
        "hdm_particles_download": ([P, C.POINTER(HdmParticlesHost)], C.c_int),
        "hdm_links_upload": ([P, C.POINTER(HdmLinksHost)], C.c_int),
        "hdm_links_download": ([P, C.POINTER(HdmLinksHost)], C.c_int),
        "hdm_grid_build": ([P, C.c_double], C.c_int),
        "hdm_grid_info_get": ([P, C.POINTER(HdmGridInfo)], C.c_int),
        "hdm_grid_download": ([P, _i64p, _i32p], C.c_int),
        "hdm_neighbours_build": ([P, C.c_double], C.c_int),
        "hdm_neighbours_download": ([P, _i32p, _i64p, _i32p, C.c_int64], C.c_int),
        "hdm_neighbours_total": ([P, _i64p, _i32p], C.c_int),
        "hdm_links_build": ([P], C.c_int),
        "hdm_linklists_download": ([P, _i32p, _i64p, _i32p, _i32p, _i32p, C.c_int64], C.c_int),
        "hdm_force": ([P, C.POINTER(HdmForceParams)], C.c_int),
        "hdm_integrate": ([P, C.POINTER(HdmIntegParams)], C.c_int),
        "hdm_break": ([P, C.POINTER(HdmBreakParams), _i64p], C.c_int),
        "hdm_skin_heuristic": ([P, C.c_int32, _dp], C.c_int),
        "hdm_skin_snapshot": ([P], C.c_int),
        "hdm_skin_exact": ([P, _dp], C.c_int),
        "hdm_skin_heuristic_noalloc": ([P, P, P, C.c_int64, C.c_int32, P, _dp], C.c_int),
        "hdm_run": ([P, C.POINTER(HdmRunParams), C.c_int64, _u8p, C.POINTER(HdmRunStats)], C.c_int),
        "hdm_bw_probe": ([P, C.c_int32, C.c_int64, C.c_int32, _dp], C.c_int),
        "hdm_set_timestep": ([P, C.c_int64], C.c_int),
        "hdm_count_broken": ([P, _i64p], C.c_int),
        "hdm_particle_stats": ([P, _i32p, _i32p, _i64p, _i64p], C.c_int),
        "hdm_config_derive": ([C.c_int32, C.POINTER(C.c_char_p), C.POINTER(HdmSetupInfo)], C.c_int),
        "hdm_sim_create": ([C.c_int32, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(P)], C.c_int),
        "hdm_sim_destroy": ([P], C.c_int),
        "hdm_sim_last_error": ([], C.c_char_p),
        "hdm_sim_ctx": ([P, C.POINTER(P)], C.c_int),
        "hdm_sim_run_params": ([P, C.POINTER(HdmRunParams)], C.c_int),
        "hdm_sim_info": ([P, _i64p, _i64p, _i64p, _i64p], C.c_int),
        "hdm_sim_steps": ([P, C.c_int64, _u8p, C.POINTER(HdmRunStats)], C.c_int),
        "hdm_sim_config_dump": ([P], C.c_char_p),
        "hdm_sim_checkpoint": ([P, C.c_char_p], C.c_int),
        "hdm_sched_configure": ([P, C.POINTER(HdmSchedParams)], C.c_int),
        "hdm_sched_set_model": ([P, C.POINTER(HdmMlpWeights)], C.c_int),
        "hdm_sched_features": ([P, C.POINTER(HdmRunParams), C.c_int64, C.c_int32, _i32p, _f32p, _f32p, _u8p],
                               C.c_int),
        "hdm_sched_predict": ([P, C.c_int32, _f32p, _dp], C.c_int),
        "hdm_mlp_forward": ([P, C.POINTER(HdmMlpWeights), _f32p, C.c_int64, _f32p], C.c_int),
        "hdm_mlp_gradient": ([P, C.POINTER(HdmMlpWeights), _f32p, _f32p, _u8p, C.c_int64, _f32p, _dp], C.c_int),
        "hdm_mlp_train": ([P, C.POINTER(HdmMlpWeights), _f32p, _f32p, _u8p, C.c_int64,
                           C.POINTER(HdmMlpTrainParams), _dp], C.c_int),
        "hdm_sched_stats": ([P, _i64p, _i64p, _i64p], C.c_int),
        "hdm_set_exec": ([P, C.POINTER(HdmExecParams)], C.c_int),
        "hdm_set_maps": ([P, C.POINTER(C.POINTER(HdmMap2d))], C.c_int),
        "hdm_reorder": ([P, C.c_double, _i64p, _i64p], C.c_int),
        "hdm_map2d_load": ([C.c_char_p, C.c_int32, C.POINTER(HdmMap2d)], C.c_int),
        "hdm_map2d_free": ([C.POINTER(HdmMap2d)], None),
        "hdm_sched_label_begin": ([P, C.c_int32, C.POINTER(HdmRunParams)], C.c_int),
        "hdm_sched_label_track": ([P], C.c_int),
        "hdm_sched_label_end": ([P, C.c_int32, C.c_int32, _i32p, _i32p], C.c_int),
        "hdm_exec_info_get": ([P, C.POINTER(HdmExecInfo)], C.c_int),
        "hdm_gen_fcc_count": ([C.POINTER(HdmFccSpec), _i64p], C.c_int),
        "hdm_gen_fcc_fill": ([C.POINTER(HdmFccSpec), C.POINTER(HdmParticlesHost)], C.c_int),
        "hdm_gen_links_count": ([C.POINTER(HdmParticlesHost), C.c_double, _i64p], C.c_int),
        "hdm_gen_links_fill": ([C.POINTER(HdmParticlesHost), C.c_double, C.c_double, C.c_double, C.c_uint64,
                                C.POINTER(HdmLinksHost)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def declared_symbols(header: str = HEADER_PATH):
    """Function names declared in hdm_abi.h."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char)\s*\**\s*(hdm_\w+)\s*\(", text, re.M)))


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


PARTICLE_FIELDS = ("xCord", "yCord", "zCord", "xTheta", "yTheta", "zTheta",
                   "xDiff", "yDiff", "zDiff", "uDiff", "vDiff", "wDiff", "radius",
                   "xForce", "yForce", "zForce", "uForce", "vForce", "wForce")


@dataclass
class Particles:
    """Host SoA particle state (HiDEMParticleData fields on the hot path)."""
    n: int
    arrays: dict = field(default_factory=dict)
    type: Optional[np.ndarray] = None

    @classmethod
    def zeros(cls, n):
        p = cls(n=n, arrays={f: np.zeros(n) for f in PARTICLE_FIELDS}, type=np.zeros(n, np.uint8))
        return p

    def __getitem__(self, k):
        return self.arrays[k]

    def copy(self):
        return Particles(self.n, {k: v.copy() for k, v in self.arrays.items()},
                         None if self.type is None else self.type.copy())

    def struct(self):
        s = HdmParticlesHost()
        s.n = self.n
        for f in PARTICLE_FIELDS:
            a = self.arrays.get(f)
            if a is not None:
                assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape == (self.n,)
            setattr(s, f, _ptr(a, C.c_double))
        s.type = _ptr(self.type, C.c_uint8)
        return s


@dataclass
class Links:
    """Host SoA link state (HiDEMLinkData); CT arrays in checkpoint order (2l+end)."""
    n: int
    a: np.ndarray
    b: np.ndarray
    state: np.ndarray
    initialLen: np.ndarray
    maxLoad: Optional[np.ndarray]
    ct: list

    @classmethod
    def zeros(cls, n, with_max_load=True):
        return cls(n, np.zeros(n, np.int32), np.zeros(n, np.int32), np.ones(n, np.uint8), np.zeros(n),
                   np.zeros(n) if with_max_load else None, [np.zeros(2 * n) for _ in range(6)])

    def copy(self):
        return Links(self.n, self.a.copy(), self.b.copy(), self.state.copy(), self.initialLen.copy(),
                     None if self.maxLoad is None else self.maxLoad.copy(), [c.copy() for c in self.ct])

    def struct(self):
        s = HdmLinksHost()
        s.n = self.n
        s.localPartIndA = _ptr(self.a, C.c_int32)
        s.localPartIndB = _ptr(self.b, C.c_int32)
        s.state = _ptr(self.state, C.c_uint8)
        s.initialLen = _ptr(self.initialLen, C.c_double)
        s.maxLoad = _ptr(self.maxLoad, C.c_double)
        for k in range(6):
            s.CT[k] = _ptr(self.ct[k], C.c_double)
        return s


class HdmError(RuntimeError):
    pass


def check(rc, ctx=None, what=""):
    if rc != HDM_OK:
        msg = ""
        if ctx is not None:
            m = lib().hdm_last_error(ctx)
            msg = m.decode() if m else ""
        raise HdmError(f"{what} failed with status {rc}: {msg}")


# ---- generator -----------------------------------------------------------------
def gen_fcc(nx, ny, nz, a, radius, jitter, seed, origin=(0.0, 0.0, 0.0), clamp_x_below=-np.inf,
            order=HDM_ORDER_LATTICE) -> Particles:
    spec = HdmFccSpec(nx, ny, nz, a, (C.c_double * 3)(*origin), radius, jitter, seed, clamp_x_below, order)
    n = C.c_int64()
    check(lib().hdm_gen_fcc_count(C.byref(spec), C.byref(n)), what="hdm_gen_fcc_count")
    p = Particles.zeros(n.value)
    s = p.struct()
    check(lib().hdm_gen_fcc_fill(C.byref(spec), C.byref(s)), what="hdm_gen_fcc_fill")
    return p


def gen_links(p: Particles, link_gen_scale, threshold, spread, seed) -> Links:
    ps = p.struct()
    n = C.c_int64()
    check(lib().hdm_gen_links_count(C.byref(ps), link_gen_scale, C.byref(n)), what="hdm_gen_links_count")
    l = Links.zeros(n.value)
    ls = l.struct()
    check(lib().hdm_gen_links_fill(C.byref(ps), link_gen_scale, threshold, spread, seed, C.byref(ls)),
          what="hdm_gen_links_fill")
    return l


# ---- context -------------------------------------------------------------------
class Context:
    """One hdm_ctx: device mirrors + accelerator structures on one HIP stream."""

    def __init__(self, precision=HDM_PREC_FP64, device=0):
        self._h = C.c_void_p()
        rc = lib().hdm_ctx_create(precision, device, C.byref(self._h))
        if rc != HDM_OK:
            raise HdmError(f"hdm_ctx_create failed with status {rc} (no gfx950 device?)")
        self.precision = precision
        self.n = 0
        self.L = 0

    def close(self):
        if self._h:
            lib().hdm_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        check(rc, self._h, what)

    def upload(self, p: Particles, l: Optional[Links] = None):
        s = p.struct()
        self._c(lib().hdm_particles_upload(self._h, C.byref(s)), "hdm_particles_upload")
        self.n = p.n
        if l is not None:
            ls = l.struct()
            self._c(lib().hdm_links_upload(self._h, C.byref(ls)), "hdm_links_upload")
            self.L = l.n

    def download_particles(self) -> Particles:
        p = Particles.zeros(self.n)
        s = p.struct()
        self._c(lib().hdm_particles_download(self._h, C.byref(s)), "hdm_particles_download")
        return p

    def download_links(self, with_max_load=True) -> Links:
        l = Links.zeros(self.L, with_max_load)
        s = l.struct()
        self._c(lib().hdm_links_download(self._h, C.byref(s)), "hdm_links_download")
        return l

    def grid_build(self, search_distance):
        self._c(lib().hdm_grid_build(self._h, search_distance), "hdm_grid_build")

    def grid_info(self) -> HdmGridInfo:
        g = HdmGridInfo()
        self._c(lib().hdm_grid_info_get(self._h, C.byref(g)), "hdm_grid_info_get")
        return g

    def grid_download(self):
        col = np.zeros(self.n, np.int64)
        order = np.zeros(self.n, np.int32)
        self._c(lib().hdm_grid_download(self._h, _ptr(col, C.c_int64), _ptr(order, C.c_int32)), "hdm_grid_download")
        return col, order

    def neighbours_build(self, interact_dist):
        self._c(lib().hdm_neighbours_build(self._h, interact_dist), "hdm_neighbours_build")

    def neighbours_download(self):
        tot = C.c_int64()
        mx = C.c_int32()
        self._c(lib().hdm_neighbours_total(self._h, C.byref(tot), C.byref(mx)), "hdm_neighbours_total")
        counts = np.zeros(self.n, np.int32)
        starts = np.zeros(self.n + 1, np.int64)
        lst = np.zeros(max(tot.value, 1), np.int32)
        self._c(lib().hdm_neighbours_download(self._h, _ptr(counts, C.c_int32), _ptr(starts, C.c_int64),
                                              _ptr(lst, C.c_int32), lst.size), "hdm_neighbours_download")
        return counts, starts, lst[:tot.value]

    def links_build(self):
        self._c(lib().hdm_links_build(self._h), "hdm_links_build")

    def linklists_download(self):
        counts = np.zeros(self.n, np.int32)
        starts = np.zeros(self.n + 1, np.int64)
        cap = 2 * self.L
        lid = np.zeros(max(cap, 1), np.int32)
        oth = np.zeros(max(cap, 1), np.int32)
        end = np.zeros(max(cap, 1), np.int32)
        self._c(lib().hdm_linklists_download(self._h, _ptr(counts, C.c_int32), _ptr(starts, C.c_int64),
                                             _ptr(lid, C.c_int32), _ptr(oth, C.c_int32), _ptr(end, C.c_int32),
                                             max(cap, 1)), "hdm_linklists_download")
        return counts, starts, lid[:cap], oth[:cap], end[:cap]

    def force(self, fp: HdmForceParams):
        self._c(lib().hdm_force(self._h, C.byref(fp)), "hdm_force")

    def integrate(self, ip: HdmIntegParams):
        self._c(lib().hdm_integrate(self._h, C.byref(ip)), "hdm_integrate")

    def brk(self, bp: HdmBreakParams):
        nb = C.c_int64()
        self._c(lib().hdm_break(self._h, C.byref(bp), C.byref(nb)), "hdm_break")
        return nb.value

    def particle_stats(self):
        """finalizeTs statistics: per-particle active links and reference
        neighbour-list entries, and their sums (hdm_particle_stats)."""
        act = np.zeros(self.n, np.int32)
        nbr = np.zeros(self.n, np.int32)
        sa, sn = C.c_int64(), C.c_int64()
        self._c(lib().hdm_particle_stats(self._h, _ptr(act, C.c_int32), _ptr(nbr, C.c_int32), C.byref(sa),
                                         C.byref(sn)), "hdm_particle_stats")
        return act, nbr, sa.value, sn.value

    def skin_heuristic(self, step_size=10):
        d = C.c_double()
        self._c(lib().hdm_skin_heuristic(self._h, step_size, C.byref(d)), "hdm_skin_heuristic")
        return d.value

    def skin_snapshot(self):
        self._c(lib().hdm_skin_snapshot(self._h), "hdm_skin_snapshot")

    def skin_exact(self):
        d = C.c_double()
        self._c(lib().hdm_skin_exact(self._h, C.byref(d)), "hdm_skin_exact")
        return d.value

    def run(self, rp: HdmRunParams, n_steps, log_rebuilds=False):
        st = HdmRunStats()
        log = np.zeros(max(n_steps, 1), np.uint8) if log_rebuilds else None
        self._c(lib().hdm_run(self._h, C.byref(rp), n_steps, _ptr(log, C.c_uint8), C.byref(st)), "hdm_run")
        return (st, log[:n_steps]) if log_rebuilds else st

    def bw_probe(self, kind, nbytes, iters=10):
        g = C.c_double()
        self._c(lib().hdm_bw_probe(self._h, kind, nbytes, iters, C.byref(g)), "hdm_bw_probe")
        return g.value

    def set_timestep(self, ts):
        self._c(lib().hdm_set_timestep(self._h, ts), "hdm_set_timestep")

    def count_broken(self):
        b = C.c_int64()
        self._c(lib().hdm_count_broken(self._h, C.byref(b)), "hdm_count_broken")
        return b.value

    # ---- learned rebuild scheduler
    def sched_configure(self, block_cells=4, min_batch=8, max_batch=256, safety=0.5):
        sp = HdmSchedParams(block_cells, min_batch, max_batch, safety)
        self._c(lib().hdm_sched_configure(self._h, C.byref(sp)), "hdm_sched_configure")
        self._block_cells = block_cells

    def sched_set_model(self, w):
        self._c(lib().hdm_sched_set_model(self._h, C.byref(w) if w is not None else None), "hdm_sched_set_model")

    def sched_blocks(self):
        """Number of scheduler blocks over the current grid (hdm_sched_features' nBlocks)."""
        g = self.grid_info()
        bc = getattr(self, "_block_cells", 4)
        return (-(-g.X // bc)) * (-(-g.Y // bc)) * (-(-g.Z // bc))

    def sched_features(self, rp, steps_since_rebuild):
        """Features [nBlocks][16], per-block max |x - x0| and validity of the current state."""
        n = self.sched_blocks()
        f = np.zeros((n, SCHED_FEATURES), np.float32)
        d = np.zeros(n, np.float32)
        v = np.zeros(n, np.uint8)
        nb = C.c_int32()
        self._c(lib().hdm_sched_features(self._h, C.byref(rp), steps_since_rebuild, n, C.byref(nb),
                                         _ptr(f, C.c_float), _ptr(d, C.c_float), _ptr(v, C.c_uint8)),
                "hdm_sched_features")
        assert nb.value == n
        return f, d, v

    def sched_predict(self, nblocks):
        y = np.zeros(nblocks, np.float32)
        m = C.c_double()
        self._c(lib().hdm_sched_predict(self._h, nblocks, _ptr(y, C.c_float), C.byref(m)), "hdm_sched_predict")
        return y, m.value

    def mlp_forward(self, w, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(X.shape[0], np.float32)
        self._c(lib().hdm_mlp_forward(self._h, C.byref(w), _ptr(X, C.c_float), X.shape[0], _ptr(y, C.c_float)),
                "hdm_mlp_forward")
        return y

    def mlp_gradient(self, w, X, target, censored=None):
        X = np.ascontiguousarray(X, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        cen = None if censored is None else np.ascontiguousarray(censored, np.uint8)
        g = np.zeros(HdmMlpWeights.N_PARAMS, np.float32)
        loss = C.c_double()
        self._c(lib().hdm_mlp_gradient(self._h, C.byref(w), _ptr(X, C.c_float), _ptr(target, C.c_float),
                                       None if cen is None else _ptr(cen, C.c_uint8), X.shape[0],
                                       _ptr(g, C.c_float), C.byref(loss)), "hdm_mlp_gradient")
        return g, loss.value

    def mlp_train(self, w, X, target, censored=None, epochs=10, batch=4096, lr=1e-3, beta1=0.9, beta2=0.999,
                  eps=1e-8):
        X = np.ascontiguousarray(X, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        cen = None if censored is None else np.ascontiguousarray(censored, np.uint8)
        tp = HdmMlpTrainParams(epochs, batch, lr, beta1, beta2, eps)
        loss = C.c_double()
        self._c(lib().hdm_mlp_train(self._h, C.byref(w), _ptr(X, C.c_float), _ptr(target, C.c_float),
                                    None if cen is None else _ptr(cen, C.c_uint8), X.shape[0], C.byref(tp),
                                    C.byref(loss)), "hdm_mlp_train")
        return loss.value

    def sched_label_begin(self, slot, rp):
        self._c(lib().hdm_sched_label_begin(self._h, slot, C.byref(rp)), "hdm_sched_label_begin")

    def sched_label_track(self):
        self._c(lib().hdm_sched_label_track(self._h), "hdm_sched_label_track")

    def sched_label_end(self, slot, capacity):
        out = np.zeros(capacity, np.int32)
        nb = C.c_int32()
        self._c(lib().hdm_sched_label_end(self._h, slot, capacity, _ptr(out, C.c_int32), C.byref(nb)),
                "hdm_sched_label_end")
        return out[:nb.value]

    def reorder(self, search_distance):
        """hdm_reorder: returns (order, link_order), the old index of each new
        particle / link."""
        order = np.zeros(self.n, np.int64)
        lorder = np.zeros(max(self.L, 1), np.int64)
        self._c(lib().hdm_reorder(self._h, search_distance, _ptr(order, C.c_int64), _ptr(lorder, C.c_int64)),
                "hdm_reorder")
        return order, lorder[:self.L]

    def set_maps(self, bed: "Map2d | None" = None, friction: "Map2d | None" = None, ext=None):
        """GlacierDem map inputs (hdm_set_maps); ext: six maps or None entries
        (x, y, z, u, v, w force).  Replaces every slot."""
        slots = [bed, friction] + list(ext if ext is not None else [None] * 6)
        assert len(slots) == HDM_MAP_COUNT
        keep = [m.struct() if m is not None else None for m in slots]
        arr = (C.POINTER(HdmMap2d) * HDM_MAP_COUNT)()
        for k, kv in enumerate(keep):
            arr[k] = C.pointer(kv[0]) if kv is not None else C.POINTER(HdmMap2d)()
        self._c(lib().hdm_set_maps(self._h, arr), "hdm_set_maps")

    def set_exec(self, autotune=False, graphs=False, variant=-1, list_with_partners=False, persistent=False):
        ep = HdmExecParams(int(autotune), int(graphs), variant, int(list_with_partners), int(persistent))
        self._c(lib().hdm_set_exec(self._h, C.byref(ep)), "hdm_set_exec")

    def exec_info(self):
        i = HdmExecInfo()
        self._c(lib().hdm_exec_info_get(self._h, C.byref(i)), "hdm_exec_info_get")
        return {"variant": i.variant, "name": i.name.decode(), "tuned": bool(i.tuned),
                "tune_ms": [i.tuneMs[v] for v in range(i.variants)], "graph_captures": i.graphCaptures,
                "graph_replays": i.graphReplays, "merged_contacts": bool(i.mergedContacts),
                "persistent_launches": i.persistentLaunches}

    def sched_stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._c(lib().hdm_sched_stats(self._h, C.byref(a), C.byref(b), C.byref(c)), "hdm_sched_stats")
        return {"host_checks": a.value, "steps_launched": b.value, "steps_executed": c.value}

    def sync(self):
        self._c(lib().hdm_sync(self._h), "hdm_sync")

    def set_timing(self, on=True):
        self._c(lib().hdm_set_timing(self._h, 1 if on else 0), "hdm_set_timing")

    def reset_kernel_times(self):
        self._c(lib().hdm_reset_kernel_times(self._h), "hdm_reset_kernel_times")

    def kernel_times(self):
        k = len(TIMER_NAMES)  # HDM_T_COUNT
        ms = (C.c_double * k)()
        cnt = (C.c_int64 * k)()
        self._c(lib().hdm_kernel_times(self._h, ms, cnt), "hdm_kernel_times")
        return {TIMER_NAMES[i]: (ms[i], cnt[i]) for i in range(k)}

    def device_bytes(self):
        b = C.c_int64()
        self._c(lib().hdm_device_bytes(self._h, C.byref(b)), "hdm_device_bytes")
        return b.value


# ---- TOML-driven setup (C++ configure / HiDEMLibHip) ----------------------------
def _argv(args):
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    return len(args), arr


def config_derive(config_path, overrides=()):
    """Setup parameters derived by the C++ TOML path (no GPU needed)."""
    argc, argv = _argv(list(overrides) + [config_path])
    info = HdmSetupInfo()
    rc = lib().hdm_config_derive(argc, argv, C.byref(info))
    if rc != HDM_OK:
        raise HdmError(f"hdm_config_derive: {lib().hdm_sim_last_error().decode()}")
    return info


class Sim:
    """HiDEMLibHip set up from a TOML config (hdm_sim_*)."""

    def __init__(self, config_path, overrides=(), precision=HDM_PREC_FP64):
        argc, argv = _argv(list(overrides) + [config_path])
        self._h = C.c_void_p()
        rc = lib().hdm_sim_create(precision, argc, argv, C.byref(self._h))
        if rc != HDM_OK:
            raise HdmError(f"hdm_sim_create failed ({rc}): {lib().hdm_sim_last_error().decode()}")
        n, L, ts, end = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        lib().hdm_sim_info(self._h, C.byref(n), C.byref(L), C.byref(ts), C.byref(end))
        self.n, self.L, self.end = n.value, L.value, end.value
        ctx = C.c_void_p()
        lib().hdm_sim_ctx(self._h, C.byref(ctx))
        self.ctx = Context.__new__(Context)  # borrowed view of the sim's context
        self.ctx._h = ctx
        self.ctx.precision = precision
        self.ctx.n, self.ctx.L = self.n, self.L
        self.ctx.close = lambda: None

    def run_params(self):
        rp = HdmRunParams()
        check(lib().hdm_sim_run_params(self._h, C.byref(rp)), what="hdm_sim_run_params")
        return rp

    def steps(self, n, log_rebuilds=False):
        st = HdmRunStats()
        log = np.zeros(max(n, 1), np.uint8) if log_rebuilds else None
        check(lib().hdm_sim_steps(self._h, n, _ptr(log, C.c_uint8), C.byref(st)), what="hdm_sim_steps")
        return (st, log[:n]) if log_rebuilds else st

    def checkpoint(self, path):
        """hdm_sim_checkpoint: HDF5 checkpoint of the current state and step."""
        rc = lib().hdm_sim_checkpoint(self._h, os.fsencode(path))
        if rc != HDM_OK:
            raise HdmError(f"hdm_sim_checkpoint failed ({rc}): {lib().hdm_sim_last_error().decode()}")

    def config_dump(self):
        return lib().hdm_sim_config_dump(self._h).decode()

    def close(self):
        if self._h:
            lib().hdm_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- batched form ---------------------------------------------------------------
class HdmBatchSpec(C.Structure):
    _fields_ = [("nBlocks", C.c_int32), ("minParticles", C.c_int64), ("maxParticles", C.c_int64),
                ("a", C.c_double), ("radius", C.c_double), ("jitter", C.c_double), ("gap", C.c_double),
                ("thrLo", C.c_double), ("thrHi", C.c_double), ("spread", C.c_double),
                ("waterLo", C.c_double), ("waterHi", C.c_double), ("clampX", C.c_double), ("seed", C.c_uint64)]


class HdmBatchBlock(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("origin", C.c_double * 3),
                ("particleStart", C.c_int64), ("particleCount", C.c_int64), ("threshold", C.c_double),
                ("waterLevel", C.c_double)]


def _declare_batch(L):
    L.hdm_gen_batch_layout.argtypes = [C.POINTER(HdmBatchSpec), C.POINTER(HdmBatchBlock), _i64p]
    L.hdm_gen_batch_fill.argtypes = [C.POINTER(HdmBatchSpec), C.POINTER(HdmBatchBlock), C.POINTER(HdmParticlesHost),
                                     _i32p]
    L.hdm_gen_batch_thresholds.argtypes = [C.POINTER(HdmBatchSpec), C.POINTER(HdmBatchBlock), _i32p,
                                           C.POINTER(HdmLinksHost)]
    L.hdm_set_blocks.argtypes = [C.c_void_p, _i32p, C.c_int32, _dp]
    for f in (L.hdm_gen_batch_layout, L.hdm_gen_batch_fill, L.hdm_gen_batch_thresholds, L.hdm_set_blocks):
        f.restype = C.c_int


@dataclass
class Batch:
    spec: HdmBatchSpec
    blocks: object          # ctypes array of HdmBatchBlock
    block_id: np.ndarray    # per particle
    water: np.ndarray       # per block (absolute z)
    particles: Particles
    links: Links


def gen_batch(n_blocks=256, min_particles=5000, max_particles=200000, a=1.4707821048680188, radius=0.5,
              jitter=0.01, gap=3.0, thr=(2e-4, 5e-4), spread=0.1, water=(0.2, 0.8), clamp_x=2.0, seed=11,
              link_gen_scale=1.092, link_seed=7) -> Batch:
    L = lib()
    _declare_batch(L)
    spec = HdmBatchSpec(n_blocks, min_particles, max_particles, a, radius, jitter, gap, thr[0], thr[1], spread,
                        water[0], water[1], clamp_x, seed)
    blocks = (HdmBatchBlock * n_blocks)()
    n = C.c_int64()
    check(L.hdm_gen_batch_layout(C.byref(spec), blocks, C.byref(n)), what="hdm_gen_batch_layout")
    p = Particles.zeros(n.value)
    bid = np.zeros(n.value, np.int32)
    ps = p.struct()
    check(L.hdm_gen_batch_fill(C.byref(spec), blocks, C.byref(ps), _ptr(bid, C.c_int32)), what="hdm_gen_batch_fill")
    l = gen_links(p, link_gen_scale, 1.0, spread, link_seed)
    ls = l.struct()
    check(L.hdm_gen_batch_thresholds(C.byref(spec), blocks, _ptr(bid, C.c_int32), C.byref(ls)),
          what="hdm_gen_batch_thresholds")
    water_abs = np.array([blocks[b].waterLevel for b in range(n_blocks)])
    return Batch(spec, blocks, bid, water_abs, p, l)


def set_blocks(ctx: "Context", block_id: np.ndarray, water: np.ndarray):
    L = lib()
    _declare_batch(L)
    ctx._c(L.hdm_set_blocks(ctx._h, _ptr(block_id, C.c_int32), len(water), _ptr(water, C.c_double)),
           "hdm_set_blocks")
