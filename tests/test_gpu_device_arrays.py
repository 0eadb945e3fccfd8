"""Uploads from and downloads into arrays in device memory (as HiDEMLib's
ManVec buffers are in its GPU builds): the same run as from host arrays, bit
for bit, in both forms (the FP32 form converts its force fields on the
device)."""
import ctypes as C

import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


class DeviceMemory:
    """hipMalloc'd buffers from the HIP runtime the library itself links (a
    second runtime in the process, such as torch's bundled one, would not see
    the GPU once the library's runtime holds it)."""

    H2D, D2H = 1, 2

    def __init__(self):
        A.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def check(self, err, what):
        assert err == 0, f"{what}: hip error {err}"

    def upload(self, a, dtype):
        a = np.ascontiguousarray(a)
        ptr = C.c_void_p()
        self.check(self.hip.hipMalloc(C.byref(ptr), C.c_size_t(max(a.nbytes, 1))), "hipMalloc")
        self.ptrs.append(ptr)
        self.check(self.hip.hipMemcpy(ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), self.H2D), "hipMemcpy")
        return C.cast(ptr, C.POINTER(dtype))

    def zeros(self, n):
        return self.upload(np.zeros(n, dtype=np.float64), C.c_double)

    def download(self, ptr, n):
        out = np.empty(n, dtype=np.float64)
        self.check(self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.cast(ptr, C.c_void_p), C.c_size_t(out.nbytes),
                                      self.D2H), "hipMemcpy")
        return out

    def free(self):
        for ptr in self.ptrs:
            self.check(self.hip.hipFree(ptr), "hipFree")
        self.ptrs = []


@pytest.mark.parametrize("prec", [A.HDM_PREC_FP64, A.HDM_PREC_MIXED])
def test_device_arrays_match_host_arrays(prec):
    print("device arrays", prec, flush=True)
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"), [8, 6, 5])
    r = np.random.default_rng(2)
    for f in ("xForce", "yForce", "zForce", "uForce", "vForce", "wForce"):
        p.arrays[f][:] = r.normal(scale=1e3, size=p.n)
    mem = DeviceMemory()
    dptr = mem.upload

    ps = A.HdmParticlesHost()
    ps.n = p.n
    for f in A.PARTICLE_FIELDS:
        setattr(ps, f, dptr(p.arrays[f], C.c_double))
    ps.type = dptr(p.type, C.c_uint8)
    ls = A.HdmLinksHost()
    ls.n = l.n
    ls.localPartIndA, ls.localPartIndB = dptr(l.a, C.c_int32), dptr(l.b, C.c_int32)
    ls.state, ls.initialLen, ls.maxLoad = dptr(l.state, C.c_uint8), dptr(l.initialLen, C.c_double), dptr(l.maxLoad, C.c_double)
    for k in range(6):
        ls.CT[k] = dptr(l.ct[k], C.c_double)
    with A.Context(prec) as ctx:
        ctx.upload(p, l)
        up_host = ctx.download_particles()
        ctx.links_build()
        ctx.run(s.run, 20)
        want = ctx.download_particles()
    with A.Context(prec) as ctx:
        ctx._c(A.lib().hdm_particles_upload(ctx._h, C.byref(ps)), "upload particles")
        ctx._c(A.lib().hdm_links_upload(ctx._h, C.byref(ls)), "upload links")
        ctx.n, ctx.L = p.n, l.n
        up_dev = ctx.download_particles()
        ctx.links_build()
        ctx.run(s.run, 20)
        # download into fresh device arrays
        out = {f: mem.zeros(p.n) for f in A.PARTICLE_FIELDS}
        qs = A.HdmParticlesHost()
        qs.n = p.n
        for f in A.PARTICLE_FIELDS:
            setattr(qs, f, out[f])
        ctx._c(A.lib().hdm_particles_download(ctx._h, C.byref(qs)), "download particles")
        got = {f: mem.download(out[f], p.n) for f in A.PARTICLE_FIELDS}
    mem.free()
    for f in A.PARTICLE_FIELDS:
        assert bits_equal(up_dev[f], up_host[f]), ("upload", f)
        assert bits_equal(got[f], want[f]), ("run", f)
