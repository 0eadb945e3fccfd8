"""TOML-driven setup (C++ configure / ArgParser) against the Python parameter
derivation of the same configs, and the reference's CLI override rules."""
import os
import tempfile

import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params


def fields(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if hasattr(v, "_fields_"):
            for k, w in fields(v).items():
                out[f"{name}.{k}"] = w
        elif hasattr(v, "__len__") and not isinstance(v, (bytes, str)):
            out[name] = tuple(v)
        else:
            out[name] = v
    return out


@pytest.mark.parametrize("name", ["iceblock_16k", "glacier_8m"])
def test_cpp_and_python_derivations_agree(name):
    path = os.path.join(params.CONFIG_DIR, name + ".toml")
    info = A.config_derive(path)
    py = params.derive(params.load_config(name))
    a, b = fields(info.run), fields(py.run)
    assert a.keys() == b.keys()
    for k in a:
        assert a[k] == b[k], k   # bit-identical doubles
    cfg = params.load_config(name)
    assert (info.fcc.nx, info.fcc.ny, info.fcc.nz) == tuple(cfg["hdm"]["cells"])
    assert info.fcc.a == py.a and info.fcc.radius == py.radius
    assert info.linkGenScale == py.link_gen_scale
    assert info.endTimestep == py.time_steps


def test_cli_overrides_like_reference():
    path = os.path.join(params.CONFIG_DIR, "iceblock_16k.toml")
    info = A.config_derive(path, ["--general.dt=0.0002", "--grid.gridRegenInterval", "25",
                                  "--hdm.cells=[4,5,6]", "--iteration.timeSteps=77"])
    assert info.run.dt == 0.0002 and info.run.integ.invDt == 1.0 / 0.0002
    assert info.run.policy == A.HDM_REBUILD_INTERVAL and info.run.regenInterval == 25
    assert (info.fcc.nx, info.fcc.ny, info.fcc.nz) == (4, 5, 6)
    assert info.endTimestep == 77


def test_reference_sample_config_parses():
    # the reference's own sample layout: [[iceParameters]], [[geometry]], inline comments, 2e+9
    text = '''
[general]
scale = 5.0
E = 2e+9   # Young's modulus
dt = 0.001
[grid]
searchDistanceScale = 1.4
[lattice]
type = "FCC"
particleRadius = 0.5
fccSpacing = 1.41421
linkGenScale = 1.2
[[iceParameters]]
linkGenerateProbability = 1.0
boundingGeometry = "block"
linkWith = [0]
[[geometry]]
type = "box"
name = "block"
lower = [-0,-0,-20]
upper = [80,80,60]
[hdm]
cells = [2, 2, 2]
'''
    with tempfile.NamedTemporaryFile("w", suffix=".toml", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        info = A.config_derive(path)
        assert info.run.force.scaleE == 2e9 * 5.0
        assert info.fcc.a == 1.41421 * 5.0
        assert info.run.searchDistance == 1.4 * (0.5 * 5.0 * 2.0)
    finally:
        os.unlink(path)


def test_bad_config_reports_error():
    with tempfile.NamedTemporaryFile("w", suffix=".toml", delete=False) as f:
        f.write("[lattice]\ntype = \"HCP\"\n")
        path = f.name
    try:
        with pytest.raises(A.HdmError, match="FCC"):
            A.config_derive(path)
    finally:
        os.unlink(path)
