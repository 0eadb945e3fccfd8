"""CPU-only checks of the C-ABI library: it loads, exports every symbol the
header declares, rejects use without a gfx950 device, and its (host-side)
seeded generator produces the documented lattice and links."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    declared = A.declared_symbols()
    assert len(declared) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared if s not in exported]
    assert not missing, f"declared in hdm_abi.h but not exported: {missing}"
    for s in declared:
        getattr(lib, s)
    assert lib.hdm_abi_version() == 1


def test_no_cuda_or_thrust_in_library():
    out = subprocess.run(["nm", "-D", A.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "cuda" not in out.lower()
    assert "thrust" not in out.lower()


def test_ctx_create_rejects_bad_args():
    h = C.c_void_p()
    assert A.lib().hdm_ctx_create(7, 0, C.byref(h)) == -1  # unknown precision
    assert A.lib().hdm_ctx_create(0, -1, C.byref(h)) != 0


def test_fcc_generator_lattice_and_jitter():
    a, r, j = 1.4707821048680188, 0.5, 0.01
    p = A.gen_fcc(4, 3, 2, a, r, j, 1)
    assert p.n == 4 * 3 * 2 * 4
    q = A.gen_fcc(4, 3, 2, a, r, 0.0, 1)
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a
    for idx in range(q.n):
        cell, b = divmod(idx, 4)
        i, jj, k = cell % 4, (cell // 4) % 3, cell // 12
        want = np.array([i * a, jj * a, k * a]) + basis[b]
        got = np.array([q["xCord"][idx], q["yCord"][idx], q["zCord"][idx]])
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    d = np.stack([p["xCord"] - q["xCord"], p["yCord"] - q["yCord"], p["zCord"] - q["zCord"]])
    assert np.all(np.abs(d) <= j) and np.abs(d).max() > 0.9 * j
    p2 = A.gen_fcc(4, 3, 2, a, r, j, 1)
    assert np.array_equal(p["xCord"], p2["xCord"])  # deterministic
    p3 = A.gen_fcc(4, 3, 2, a, r, j, 2)
    assert not np.array_equal(p["xCord"], p3["xCord"])  # seeded


def test_iceblock_config_links():
    cfg = params.load_config("iceblock_16k")
    s, p, l = params.build_inputs(cfg)
    assert p.n == 16000
    # every pair within 1.05 x the NN distance and nothing else: 12 per interior particle
    assert 88000 < l.n < 96000
    assert np.all(l.a < l.b)
    key = l.a.astype(np.int64) * p.n + l.b
    assert np.all(np.diff(key) > 0)  # sorted by (A, B), unique
    deg = np.bincount(l.a, minlength=p.n) + np.bincount(l.b, minlength=p.n)
    x, y, z = p["xCord"], p["yCord"], p["zCord"]
    inner = (x > 2) & (x < 26) & (y > 2) & (y < 26) & (z > 2) & (z < 12)
    assert np.all(deg[inner] == 12)
    d = np.sqrt((x[l.a] - x[l.b]) ** 2 + (y[l.a] - y[l.b]) ** 2 + (z[l.a] - z[l.b]) ** 2)
    np.testing.assert_allclose(l.initialLen, d, rtol=1e-15)
    assert d.max() <= 1.05 * 1.04 and d.min() > 1.04 - 2 * math.sqrt(3) * 0.01
    assert l.maxLoad.min() >= 3e-4 * 0.9 and l.maxLoad.max() <= 3e-4 * 1.1
    assert abs(l.maxLoad.mean() / 3e-4 - 1) < 0.01
    # clamp: every particle with x < 2 m is fixed
    assert np.array_equal(p.type == 1, x < 2.0)
    assert s.search_distance == pytest.approx(1.2) and s.interact_dist == pytest.approx(0.2)


def test_glacier_config_size():
    cfg = params.load_config("glacier_8m")
    s = params.derive(cfg)
    nx, ny, nz = cfg["hdm"]["cells"]
    assert nx * ny * nz * 4 == 8_000_000
    assert s.a == pytest.approx(1.0, rel=1e-6)
    assert 2 * s.radius == pytest.approx(1 / math.sqrt(2) - 0.04, rel=1e-4)


def test_morton_order_is_a_permutation_of_lattice_order():
    a = 1.4707821048680188
    lat = A.gen_fcc(6, 5, 4, a, 0.5, 0.01, 3, order=A.HDM_ORDER_LATTICE)
    mor = A.gen_fcc(6, 5, 4, a, 0.5, 0.01, 3, order=A.HDM_ORDER_MORTON)
    P = lambda p: np.stack([p["xCord"], p["yCord"], p["zCord"]], 1)  # noqa: E731
    pl, pm = P(lat), P(mor)
    assert not np.array_equal(pl, pm)
    key = lambda x: np.lexsort((x[:, 2], x[:, 1], x[:, 0]))  # noqa: E731
    assert np.array_equal(pl[key(pl)], pm[key(pm)])  # same particles, bit for bit
    # Morton order is spatially local: consecutive particles stay close
    step = np.linalg.norm(np.diff(pm, axis=0), axis=1)
    assert np.median(step) < 1.5 * a


def test_timer_names_match_header_enum():
    hdr = open(A.HEADER_PATH).read()
    body = re.search(r"enum \{ HDM_T_FORCE = 0,(.*?)HDM_T_COUNT \};", hdr, re.S).group(1)
    names = ["HDM_T_FORCE"] + [s.strip() for s in body.split(",") if s.strip()]
    assert len(names) == len(A.TIMER_NAMES)
    assert [n[6:].lower() for n in names] == list(A.TIMER_NAMES)


def test_io_library_exports_every_declared_symbol():
    from hidem_amd import h5io
    h5io.io_lib()
    declared = A.declared_symbols(h5io.IO_HEADER_PATH)
    assert len(declared) >= 6
    out = subprocess.run(["nm", "-D", "--defined-only", h5io.IO_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared if s not in exported]
    assert not missing, f"declared in hdm_io.h but not exported: {missing}"


CALLSITES = os.path.join(ROOT, "tests", "adapter_callsites.cpp")


def build_callsites(out, *extra):
    cmd = ["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), CALLSITES, *extra]
    if out is None:
        cmd.append("-fsyntax-only")
    else:
        cmd += ["-o", str(out), "-L", os.path.dirname(A.LIB_PATH), "-lhidem_amd",
                "-Wl,-rpath," + os.path.dirname(A.LIB_PATH)]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_adapter_takes_the_reference_call_expressions(tmp_path):
    """HiDEMLib.h's accelerator calls (generateNeighbours(pData, lData, s, i),
    iterateCombinedPre(policy, pData, lData, nF, lF, Pack4), iterateLinks(
    policy, pData, lData, bF, Pack3), lAccel.sortSideData, grid.toVTKHDF ...)
    compile against CombinedAccelHip and link against the library."""
    r = build_callsites(tmp_path / "callsites")
    assert r.returncode == 0, r.stderr


def test_adapter_rejects_other_functors():
    r = build_callsites(None, "-DHDM_WRONG_FUNCTOR")
    assert r.returncode != 0
    assert "runs the default link force functor only" in r.stderr
