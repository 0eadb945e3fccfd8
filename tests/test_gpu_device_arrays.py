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


@pytest.mark.parametrize("prec", [A.HDM_PREC_FP64, A.HDM_PREC_MIXED])
def test_device_arrays_match_host_arrays(prec):
    import torch
    print("device arrays", prec, flush=True)
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"), [8, 6, 5])
    r = np.random.default_rng(2)
    for f in ("xForce", "yForce", "zForce", "uForce", "vForce", "wForce"):
        p.arrays[f][:] = r.normal(scale=1e3, size=p.n)
    dev = torch.device("cuda:0")
    keep = []

    def dptr(a, dtype):
        t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        keep.append(t)
        return C.cast(t.data_ptr(), C.POINTER(dtype))

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
    torch.cuda.synchronize()
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
        out = {f: torch.zeros(p.n, dtype=torch.float64, device=dev) for f in A.PARTICLE_FIELDS}
        qs = A.HdmParticlesHost()
        qs.n = p.n
        for f in A.PARTICLE_FIELDS:
            setattr(qs, f, C.cast(out[f].data_ptr(), C.POINTER(C.c_double)))
        torch.cuda.synchronize()
        ctx._c(A.lib().hdm_particles_download(ctx._h, C.byref(qs)), "download particles")
        got = {f: out[f].cpu().numpy() for f in A.PARTICLE_FIELDS}
    for f in A.PARTICLE_FIELDS:
        assert bits_equal(up_dev[f], up_host[f]), ("upload", f)
        assert bits_equal(got[f], want[f]), ("run", f)
