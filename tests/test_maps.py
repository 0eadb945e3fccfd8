"""GlacierDem map inputs on the CPU: the oracle's Map2d restatement
(oracle/hdm_oracle.c: hdmo_map_*) against the reference's Map2d tests
(CatchTests/Map2DTester.cpp) and closed forms, and the library's CSV map
loader (hdm_map2d_load, host code) against the reference's sample DEM."""
import os

import numpy as np
import pytest

from hidem_amd import _abi as A

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_index_cases(orc):
    # Map2DTester.cpp:11-35, 80-89
    m, _ = orc.map_struct(np.zeros((2, 2)), 0, 0, 1, 1)
    assert orc.map_indexes(m, 0, 1) == (0, 1, 1, 1)
    m, _ = orc.map_struct(np.zeros((4, 4)), 0, 0, 1, 1)
    assert orc.map_indexes(m, 0, 1) == (0, 3, 1, 3)
    assert orc.map_indexes(m, 0.25, 0.74) == (0, 2, 1, 3)
    m, _ = orc.map_struct(np.zeros((3, 3)), -2, -2, 2, 2)
    assert orc.map_indexes(m, 0.25, -1.74) == (1, 0, 2, 1)


def test_reference_interp_cases(orc):
    # Map2DTester.cpp:37-78 (setPointInd(x, y, v) is values[y, x]); the
    # reference checks |result - expected| < 0.001
    v = np.zeros((2, 2))
    v[0, 0], v[0, 1], v[1, 0], v[1, 1] = 91, 210, 162, 95
    m, _ = orc.map_struct(v, 14, 20, 15, 21)
    assert abs(orc.map_interp(m, 14.5, 20.2) - 146.1) < 1e-3
    v = np.zeros((3, 3))
    v[1, 0], v[1, 1] = 5, 5
    m, _ = orc.map_struct(v, 0, 0, 2, 2)
    assert abs(orc.map_interp(m, 1, 1) - 5) < 1e-3
    v = np.zeros((2, 2))
    v[1, :] = 1
    m, _ = orc.map_struct(v, 0, 0, 1, 1)
    assert abs(orc.map_interp(m, 0.5, 2) - 1) < 1e-3     # y beyond the grid: clamped
    assert abs(orc.map_interp(m, 2, 0.5) - 0.5) < 1e-3   # x beyond the grid: clamped


def test_bilinear_function_is_reproduced(orc):
    """f = a + b x + c y + d x y is bilinear, so interpolation reproduces it
    up to FP32 rounding anywhere inside the grid; exact grid values at the
    nodes; the nearest edge value beyond it."""
    a, b, c, d = 3.0, 0.5, -0.25, 0.125
    xs, ys = np.linspace(-2, 6, 9), np.linspace(1, 4, 7)
    vals = a + b * xs[None, :] + c * ys[:, None] + d * xs[None, :] * ys[:, None]
    m, _ = orc.map_struct(vals, -2, 1, 6, 4)
    r = np.random.default_rng(0)
    for x, y in zip(r.uniform(-2, 6, 200), r.uniform(1, 4, 200)):
        fx, fy = float(np.float32(x)), float(np.float32(y))
        want = a + b * fx + c * fy + d * fx * fy
        assert abs(orc.map_interp(m, x, y) - want) < 4e-6 * 12, (x, y)
    for ix in range(9):
        for iy in range(7):
            got = orc.map_interp(m, xs[ix], ys[iy])
            assert abs(got - vals[iy, ix]) < 2e-6 * 12
    assert orc.map_interp(m, 100.0, 100.0) == vals[-1, -1]   # both axes clamp: the value itself
    assert orc.map_interp(m, -100.0, -100.0) == vals[0, 0]
    # one axis clamped: linear along the other
    got = orc.map_interp(m, 2.0, 100.0)
    assert abs(got - (a + b * 2 + c * 4 + d * 2 * 4)) < 1e-5


def test_single_point_map_is_the_value(orc):
    """A 1 x 1 map (the reference's single-value force) returns its FP64
    value unrounded, wherever it is read."""
    m, _ = orc.map_struct(np.array([[0.1]]), 0, 0, 1, 1)
    for x, y in ((0.5, 0.5), (-1e6, 3e7), (float("nan"), 0.0)):
        assert orc.map_interp(m, x, y) == 0.1


def test_offset_shifts_the_map(orc):
    vals = np.arange(12.0).reshape(3, 4)
    m0, _ = orc.map_struct(vals, 0, 0, 3, 2)
    m1, _ = orc.map_struct(vals, 0, 0, 3, 2, 1.0, 0.5)
    for x, y in ((0.25, 0.5), (1.5, 1.0), (0.0, 0.0)):
        assert orc.map_interp(m1, x, y) == orc.map_interp(m0, x + 1.0, y + 0.5)


def test_loader_reads_reference_dem():
    """HiDEM/GlacierDem/dem.csv: four corner points of a 2 x 2 grid over
    [-200, 1200]^2; columns x, y, top, base, friction."""
    path = os.path.join(GOLDEN, "glacierdem_dem.csv")
    top = A.load_map(path, 2)
    assert top.values.shape == (2, 2)
    assert (top.x_min, top.y_min, top.x_max, top.y_max) == (-200.0, -200.0, 1200.0, 1200.0)
    assert top.values[0, 0] == 250.0 and top.values[0, 1] == 100.0 and top.values[1, 0] == 100.0
    assert np.all(A.load_map(path, 3).values == 0.0)
    assert np.all(A.load_map(path, 4).values == 1.0)
    with pytest.raises(A.HdmError):
        A.load_map(path, 5)


def test_loader_grid_placement(tmp_path, orc):
    """Map2DTester.cpp:92-104 through the loader: corners (-2, -2, 0),
    (-2, 2, 0), (2, -2, 1), (2, 2, 1) (setPointCord) interpolate to 0.5 at the
    centre; rows in any order land on their rounded grid index."""
    f = tmp_path / "m.csv"
    f.write_text("2 2 1\n-2 -2 0\n2 -2 1\n-2 2 0\n")
    mp = A.load_map(str(f), 2)
    assert mp.values.tolist() == [[0.0, 1.0], [0.0, 1.0]]
    m, _ = orc.map_struct(mp.values, mp.x_min, mp.y_min, mp.x_max, mp.y_max)
    assert abs(orc.map_interp(m, 0, 0) - 0.5) < 1e-3
    # Map2DTester.cpp:106-119: a 4 x 8 grid of ones
    rows = [f"{i} {j} 1" for j in np.arange(-2, 2, 0.5) for i in range(-2, 6, 2)]
    f.write_text("\n".join(rows) + "\n")
    mp = A.load_map(str(f), 2)
    assert mp.values.shape == (8, 4) and np.all(mp.values == 1.0)
    m, _ = orc.map_struct(mp.values, mp.x_min, mp.y_min, mp.x_max, mp.y_max)
    assert abs(orc.map_interp(m, 0, 0) - 1.0) < 1e-3


def test_loader_rejects_malformed(tmp_path):
    f = tmp_path / "bad.csv"
    f.write_text("0 0 1\n1 0 2 3\n")           # a row with another column count
    with pytest.raises(A.HdmError):
        A.load_map(str(f), 2)
    f.write_text("0 0 1\n1 0 1\n0 1 1\n")      # rows of unequal length
    with pytest.raises(A.HdmError):
        A.load_map(str(f), 2)
    f.write_text("0 0 1\n1 0 2x\n")            # not a number
    with pytest.raises(A.HdmError):
        A.load_map(str(f), 2)
    with pytest.raises(A.HdmError):
        A.load_map(str(tmp_path / "missing.csv"), 2)


def test_loader_extent_quirk(tmp_path):
    """The reference starts the max extents at the smallest positive normal
    float (Map2dLoader.cpp:16-19), so an all-negative x column reads
    xMax = 1.1754944e-38; the loader keeps that."""
    f = tmp_path / "neg.csv"
    f.write_text("-4 1 1\n-2 1 1\n-4 3 1\n-2 3 1\n")
    mp = A.load_map(str(f), 2)
    assert mp.x_min == -4.0 and mp.x_max == float(np.finfo(np.float32).tiny)
    assert mp.y_min == 1.0 and mp.y_max == 3.0
