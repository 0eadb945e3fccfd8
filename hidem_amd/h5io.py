"""ctypes bindings of libhidem_io.so (include/hidem/hdm_io.h): the
reference's HDF5 checkpoint file and its particle input file, plus
checkpoint save / load of a context's state."""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import _abi as A

IO_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhidem_io.so")
IO_HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "hidem",
                              "hdm_io.h")

_io = None


def io_lib():
    """Loads libhidem_io.so; raises if it has not been built."""
    global _io
    if _io is None:
        if not os.path.exists(IO_PATH):
            raise RuntimeError(f"{IO_PATH} missing: run __graft_entry__.build() (make -C hidem_amd/csrc)")
        L = C.CDLL(IO_PATH)
        PP, LP = C.POINTER(A.HdmParticlesHost), C.POINTER(A.HdmLinksHost)
        i64p = C.POINTER(C.c_int64)
        sig = {
            "hdm_io_last_error": ([], C.c_char_p),
            "hdm_checkpoint_write": ([C.c_char_p, C.c_int32, PP, LP], C.c_int),
            "hdm_checkpoint_info": ([C.c_char_p, i64p, i64p, C.POINTER(C.c_int32)], C.c_int),
            "hdm_checkpoint_read": ([C.c_char_p, PP, LP], C.c_int),
            "hdm_particle_file_count": ([C.c_char_p, C.c_char_p, i64p], C.c_int),
            "hdm_particle_file_read": ([C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, PP], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _io = L
    return _io


def _check(rc, what):
    if rc != A.HDM_OK:
        raise A.HdmError(f"{what} failed with status {rc}: {io_lib().hdm_io_last_error().decode()}")


def write_checkpoint(path: str, timestep: int, p: A.Particles, l: Optional[A.Links] = None):
    ps = p.struct()
    ls = l.struct() if l is not None else None
    _check(io_lib().hdm_checkpoint_write(os.fsencode(path), int(timestep), C.byref(ps),
                                         C.byref(ls) if ls is not None else None), "hdm_checkpoint_write")


def checkpoint_info(path: str):
    n, L, ts = C.c_int64(), C.c_int64(), C.c_int32()
    _check(io_lib().hdm_checkpoint_info(os.fsencode(path), C.byref(n), C.byref(L), C.byref(ts)),
           "hdm_checkpoint_info")
    return n.value, L.value, ts.value


def read_checkpoint(path: str, with_max_load: bool = False):
    """(timestep, Particles, Links) of a checkpoint; forces read as 0."""
    n, L, ts = checkpoint_info(path)
    p = A.Particles.zeros(n)
    l = A.Links.zeros(L, with_max_load=with_max_load)
    ps, ls = p.struct(), l.struct()
    _check(io_lib().hdm_checkpoint_read(os.fsencode(path), C.byref(ps), C.byref(ls)), "hdm_checkpoint_read")
    return ts, p, l


def read_particle_file(path: str, names=("xCord", "yCord", "zCord", "radius")) -> A.Particles:
    """Positions and radii of a [lattice] type = "File" input, read as float
    like the reference's FileParticleGenerator."""
    n = C.c_int64()
    b = [os.fsencode(x) for x in names]
    _check(io_lib().hdm_particle_file_count(os.fsencode(path), b[0], C.byref(n)), "hdm_particle_file_count")
    p = A.Particles.zeros(n.value)
    ps = p.struct()
    _check(io_lib().hdm_particle_file_read(os.fsencode(path), b[0], b[1], b[2], b[3], C.byref(ps)),
           "hdm_particle_file_read")
    return p


def save(ctx: A.Context, path: str, timestep: int):
    """Checkpoint of a context's current state (downloaded first)."""
    write_checkpoint(path, timestep, ctx.download_particles(), ctx.download_links())


def load(ctx: A.Context, path: str, with_max_load: bool = True):
    """Uploads a checkpoint into a context, builds its link lists and sets its
    step counter; returns the timestep."""
    ts, p, l = read_checkpoint(path, with_max_load=with_max_load)
    ctx.upload(p, l)
    ctx.links_build()
    ctx.set_timestep(ts)
    return ts
