"""GlacierDem map inputs on the GPU (hdm_set_maps): bed height, friction and
external force maps in the HDM_INTEG_GLACIER integrator, bit-identical to
the CPU oracle, per call and over fused steps; 1 x 1 maps reproduce the
analytic inputs bit for bit."""
import os

import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def glacier(cells=(10, 6, 5), **hdm):
    cfg = params.load_config("glacier_8m")
    cfg["hdm"].update(hdm)
    return params.build_inputs(cfg, list(cells))


def make_maps(p, seed=5):
    """Map slots over the block's footprint (one cell of margin): a tilted,
    bumpy bed crossing the bottom layer, a varying friction, external force
    maps for x and w, a single-value map for z, analytic y, u, v."""
    r = np.random.default_rng(seed)
    x, y = np.asarray(p["xCord"]), np.asarray(p["yCord"])
    ext = (float(x.min()) - 1.0, float(y.min()) - 1.0, float(x.max()) + 1.0, float(y.max()) + 1.0)
    gx, gy = np.meshgrid(np.linspace(ext[0], ext[2], 7), np.linspace(ext[1], ext[3], 5))
    bed = 0.03 * gx - 0.02 * gy - 0.05 + r.uniform(-0.02, 0.02, gx.shape)
    fric = r.uniform(0.2, 1.5, (3, 4))
    fx = r.uniform(-50.0, 50.0, (4, 3))
    fw = r.uniform(-5.0, 5.0, (2, 6))
    slots = [(bed, ext), (fric, ext), (fx, ext), None, (np.array([[-30.0]]), (0.0, 0.0, 1.0, 1.0)),
             None, None, (fw, ext)]
    oracle_maps = [None if s is None else (s[0],) + s[1] for s in slots]
    gpu_maps = [None if s is None else A.Map2d(s[0], *s[1]) for s in slots]
    return oracle_maps, gpu_maps


def set_maps(ctx, gpu_maps):
    ctx.set_maps(bed=gpu_maps[0], friction=gpu_maps[1], ext=gpu_maps[2:])


def test_integrate_with_maps_bit_identical(orc):
    s, p0, l0 = glacier()
    p = p0.copy()
    rng = np.random.default_rng(7)
    for f in ("xDiff", "yDiff", "zDiff", "uDiff", "vDiff", "wDiff"):
        p.arrays[f][:] = rng.normal(scale=1e-4, size=p.n)
    for f in ("xForce", "yForce", "zForce", "uForce", "vForce", "wForce"):
        p.arrays[f][:] = rng.normal(scale=1e4, size=p.n)
    om, gm = make_maps(p)
    say("integrate with maps", p.n)
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l0)
        set_maps(ctx, gm)
        ctx.integrate(s.integ)
        gp = ctx.download_particles()
    sim = orc.OracleSim(p, l0, s, maps=om)
    sim.integrate()
    for f in A.PARTICLE_FIELDS:
        assert bits_equal(gp[f], sim.p[f]), f
    # the maps changed the result (bed, friction and forces are not the analytic ones)
    plain = orc.OracleSim(p, l0, s)
    plain.integrate()
    assert not bits_equal(plain.p["zCord"], sim.p["zCord"])
    assert not bits_equal(plain.p["xCord"], sim.p["xCord"])


@pytest.mark.parametrize("prec", [A.HDM_PREC_FP64, A.HDM_PREC_MIXED])
def test_fused_steps_with_maps(prec, orc):
    """FP64: 150 fused steps bit-identical to the oracle; mixed: the
    graph-replayed, autotuned HDM_REBUILD_LEARNED run with the maps ends
    bit-identical to the eager run with them."""
    s, p, l = glacier()
    om, gm = make_maps(p)
    steps = 150
    say("fused steps with maps", prec)
    with A.Context(prec) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        set_maps(ctx, gm)
        st, log = ctx.run(s.run, steps, log_rebuilds=True)
        gp = ctx.download_particles()
        gl = ctx.download_links()
    if prec == A.HDM_PREC_FP64:
        sim = orc.OracleSim(p, l, s, maps=om)
        rlog, _ = sim.run(steps, log=True)
        assert np.array_equal(log, rlog)
        assert np.array_equal(gl.state, sim.l.state)
        for f in ("xCord", "yCord", "zCord", "xTheta", "yTheta", "zTheta", "xDiff", "wDiff"):
            assert bits_equal(gp[f], sim.p[f]), f
        plain = orc.OracleSim(p, l, s)
        plain.run(steps)
        assert not bits_equal(plain.p["zCord"], sim.p["zCord"])  # the maps took effect
    else:
        rp = A.HdmRunParams.from_buffer_copy(s.run)
        with A.Context(prec) as ctx:
            ctx.upload(p, l)
            ctx.links_build()
            set_maps(ctx, gm)
            ctx.set_exec(autotune=True, graphs=True)
            rp.policy = A.HDM_REBUILD_LEARNED
            ctx.run(rp, steps)
            info = ctx.exec_info()
            gq = ctx.download_particles()
        assert info["graph_replays"] > 0
        for f in ("xCord", "yCord", "zCord", "xTheta", "xDiff", "wDiff"):
            assert bits_equal(gp[f], gq[f]), f


def test_single_value_maps_equal_analytic_inputs():
    """1 x 1 maps holding the analytic bed height, friction 1 and the constant
    external force give the analytic run's bits."""
    s, p, l = glacier(bedXEnd=1e300, frictionXEnd=1e300, bedZ=-0.02)
    for k in range(6):
        s.run.integ.ext[k] = [3.0, -2.0, -40.0, 0.5, 0.0, -0.25][k]
    one = lambda v: A.Map2d.constant(v)  # noqa: E731
    out = []
    for use_maps in (False, True):
        with A.Context(A.HDM_PREC_FP64) as ctx:
            ctx.upload(p, l)
            ctx.links_build()
            if use_maps:
                ctx.set_maps(bed=one(-0.02), friction=one(1.0), ext=[one(v) for v in s.run.integ.ext])
            ctx.run(s.run, 60)
            out.append(ctx.download_particles())
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(out[0][f], out[1][f]), f


def test_toml_named_maps_match_python_run(tmp_path):
    """A GlacierDem TOML naming a bed table ([[collisionGeometry]] type =
    "bed") and an external force map ({path, fieldID}) run by the C++ driver
    (hdm_sim_*, maps loaded by hdm_map2d_load) ends bit-identical to the same
    maps loaded and bound from Python."""
    base = open(os.path.join(params.CONFIG_DIR, "glacier_8m.toml")).read()
    assert "cells = [400, 100, 50]" in base
    s0, p0, _ = glacier()
    x, y = np.asarray(p0["xCord"]), np.asarray(p0["yCord"])
    xs = np.linspace(float(x.min()) - 1.0, float(x.max()) + 1.0, 6)
    ys = np.linspace(float(y.min()) - 1.0, float(y.max()) + 1.0, 4)
    r = np.random.default_rng(9)
    rows = []
    for yv in ys:
        for xv in xs:
            rows.append(f"{xv:.6f} {yv:.6f} {0.02 * xv - 0.03 + r.uniform(-0.01, 0.01):.6f} "
                        f"{r.uniform(0.3, 1.2):.6f} {r.uniform(-20, 20):.6f}")
    csv = tmp_path / "bed.csv"
    csv.write_text("\n".join(rows) + "\n")
    text = base.replace("cells = [400, 100, 50]", "cells = [10, 6, 5]")
    text += (f'\n[externalForce]\nxForce = {{ path = "{csv}", fieldID = 4 }}\nzForce = -25.0\n'
             f'\n[[collisionGeometry]]\ntype = "bed"\nname = "bed"\npath = "{csv}"\nbaseFieldID = 2\n'
             f'frictionFieldID = 3\n')
    cfg_path = tmp_path / "glacier_maps.toml"
    cfg_path.write_text(text)
    say("TOML-named maps")
    with A.Sim(str(cfg_path), [], precision=A.HDM_PREC_FP64) as sim:
        sim.steps(40)
        q = sim.ctx.download_particles()
    cfg = params.load_config(str(cfg_path))
    setup, p, l = params.build_inputs(cfg)
    assert p.n == sim.n
    maps = params.map_inputs(cfg)
    assert maps[A.HDM_MAP_BED] is not None and maps[A.HDM_MAP_EXT] is not None and maps[A.HDM_MAP_EXT + 2] is None
    assert setup.integ.ext[2] == -25.0
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.set_maps(bed=maps[0], friction=maps[1], ext=maps[2:])
        ctx.run(setup.run, 40)
        g = ctx.download_particles()
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(q[f], g[f]), f
