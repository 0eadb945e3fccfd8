"""The allocation-free entry point hdm_skin_heuristic_noalloc, called through
the C ABI on Diff arrays in device memory, against hdm_skin_heuristic (the
context form, HiDEMLib::getMaxDistOwned + accumulateDist's stepSize * |max|,
HiDEMLib.h:424-474) on the same values: bit for bit, on two calls in a row,
the second with a smaller input, and against the maxima numpy takes."""
import ctypes as C

import numpy as np
import pytest

from hidem_amd import _abi as A
from test_gpu_device_arrays import DeviceMemory

pytestmark = pytest.mark.gpu

NA_MAX_N = 134_217_728


def diffs(n, seed, positive=True):
    r = np.random.default_rng(seed)
    d = [r.normal(scale=1e-4, size=n) for _ in range(3)]
    if not positive:   # every component <= 0: the reduction's 0.0 start wins
        d = [-np.abs(v) for v in d]
    return d


def reference(d, step):
    """max+ of each component from 0.0, then stepSize * sqrt(x^2 + y^2 + z^2)."""
    m = [max(0.0, float(v.max())) for v in d]
    return step * np.sqrt((m[0] * m[0]) + (m[1] * m[1]) + (m[2] * m[2]))


def ctx_form(d, step):
    p = A.Particles.zeros(len(d[0]))
    for f, v in zip(("xDiff", "yDiff", "zDiff"), d):
        p.arrays[f][:] = v
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p)
        return ctx.skin_heuristic(step)


def noalloc_form(mem, d, step):
    ptrs = [C.cast(mem.upload(v, C.c_double), C.c_void_p) for v in d]
    out = C.c_double(-1.0)
    rc = A.lib().hdm_skin_heuristic_noalloc(ptrs[0], ptrs[1], ptrs[2], len(d[0]), step, None, C.byref(out))
    assert rc == A.HDM_OK, A.lib().hdm_last_error(None)
    return out.value


def test_noalloc_matches_context_form_twice():
    mem = DeviceMemory()
    try:
        # two calls in a row, the second smaller (the first call's partials
        # beyond the second's workgroups stay in the shared scratch)
        for n, seed, step, pos in ((1_000_003, 1, 10, True), (77_777, 2, 10, True), (5_000, 3, 7, False)):
            print(f"noalloc n={n} seed={seed} step={step}", flush=True)
            d = diffs(n, seed, pos)
            want = ctx_form(d, step)
            got = noalloc_form(mem, d, step)
            assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (n, got, want)
            assert got == reference(d, step), (n, got, reference(d, step))
            if not pos:
                assert got == 0.0
    finally:
        mem.free()


def test_noalloc_rejects_oversize_and_empty():
    out = C.c_double(-1.0)
    L = A.lib()
    # sizes are checked before anything touches the arrays (none are passed)
    dummy = C.c_void_p(16)
    assert L.hdm_skin_heuristic_noalloc(dummy, dummy, dummy, NA_MAX_N + 1, 10, None, C.byref(out)) == -5
    assert b"exceeds the maximum" in L.hdm_last_error(None)
    assert L.hdm_skin_heuristic_noalloc(None, None, None, 0, 10, None, C.byref(out)) == A.HDM_OK
    assert out.value == 0.0
    assert L.hdm_skin_heuristic_noalloc(None, None, None, 5, 10, None, C.byref(out)) == -1
