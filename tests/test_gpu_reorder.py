"""hdm_reorder on the GPU: the redistribution order of the device (colour
keys, hipCUB radix passes, gathers) equal to the CPU oracle's, array for
array, and runs after a reorder -- from the start and in the middle of a
run, with CT already in the side store -- bit-identical to the oracle doing
the same."""
import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params
from tests.test_reorder import shuffled

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def system(cells=(10, 8, 6), seed=4):
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"), list(cells))
    r = np.random.default_rng(seed)
    for c in range(6):
        l.ct[c][:] = r.normal(size=2 * l.n)
    q, m = shuffled(p, l, seed)
    return s, q, m


def test_device_order_equals_oracle(orc):
    s, p, l = system()
    say("reorder vs oracle", p.n, l.n)
    q, m, order, lorder = orc.reorder(p, l, s.search_distance)
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        go, glo = ctx.reorder(s.search_distance)
        gp = ctx.download_particles()
        gl = ctx.download_links()
    assert np.array_equal(go, order)
    assert np.array_equal(glo, lorder)
    for f in A.PARTICLE_FIELDS:
        assert bits_equal(gp[f], q[f]), f
    assert np.array_equal(gl.a, m.a) and np.array_equal(gl.b, m.b) and np.array_equal(gl.state, m.state)
    assert bits_equal(gl.initialLen, m.initialLen)
    for c in range(6):
        assert bits_equal(gl.ct[c], m.ct[c]), c
    # consecutive particles are spatial neighbours again (the input was shuffled)
    def step(d):
        xyz = np.stack([np.asarray(d[f]) for f in ("xCord", "yCord", "zCord")], 1)
        return np.mean(np.linalg.norm(np.diff(xyz, axis=0), axis=1))
    assert step(gp) < 0.2 * step(p)


@pytest.mark.parametrize("prec", [A.HDM_PREC_FP64, A.HDM_PREC_MIXED])
def test_reorder_mid_run(prec, orc):
    """50 steps, reorder (CT in the side store goes through the checkpoint
    order and back), 50 more: FP64 bit-identical to the oracle doing the
    same; the FP32 form bit-identical to its own run with the reorder done
    through a download / oracle reorder / upload."""
    s, p, l = system(seed=6)
    say("reorder mid run", prec)
    with A.Context(prec) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.run(s.run, 50)
        ctx.reorder(s.search_distance)
        ctx.run(s.run, 50)
        gp = ctx.download_particles()
        gl = ctx.download_links()
    if prec == A.HDM_PREC_FP64:
        sim = orc.OracleSim(p, l, s)
        sim.run(50)
        q, m, _, _ = orc.reorder(sim.p, sim.l, s.search_distance)
        sim2 = orc.OracleSim(q, m, s)
        sim2.s.ts = 50
        sim2.run(50)
        want_p, want_l = sim2.p, sim2.l
    else:
        with A.Context(prec) as ctx:
            ctx.upload(p, l)
            ctx.links_build()
            ctx.run(s.run, 50)
            hp, hl = ctx.download_particles(), ctx.download_links()
        q, m, _, _ = orc.reorder(hp, hl, s.search_distance)
        with A.Context(prec) as ctx:
            ctx.upload(q, m)
            ctx.links_build()
            ctx.set_timestep(50)
            ctx.run(s.run, 50)
            want_p, want_l = ctx.download_particles(), ctx.download_links()
    for f in ("xCord", "yCord", "zCord", "xTheta", "zTheta", "xDiff", "wDiff"):
        assert bits_equal(gp[f], want_p[f]), f
    assert np.array_equal(gl.state, want_l.state)
    for c in range(6):
        assert bits_equal(gl.ct[c], want_l.ct[c]), c


def test_driver_redistribute_matches_python():
    """hdm.redistribute = true in the TOML driver: the reorder before the
    first step, then the same run as hdm_reorder + hdm_run from Python."""
    import os
    cfg = os.path.join(params.CONFIG_DIR, "iceblock_16k.toml")
    say("driver redistribute")
    with A.Sim(cfg, ["--hdm.redistribute=true"], precision=A.HDM_PREC_FP64) as sim:
        sim.steps(60)
        q = sim.ctx.download_particles()
        ql = sim.ctx.download_links(with_max_load=False)
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        order, _ = ctx.reorder(s.search_distance)
        ctx.run(s.run, 60)
        g = ctx.download_particles()
        gl = ctx.download_links(with_max_load=False)
    assert not np.array_equal(order, np.arange(p.n))   # the reorder did something
    for f in ("xCord", "yCord", "zCord", "xTheta", "xDiff"):
        assert bits_equal(q[f], g[f]), f
    assert np.array_equal(ql.state, gl.state) and np.array_equal(ql.a, gl.a)


def test_reorder_batched_blocks(orc):
    """Batched form: the per-particle block ids move with their particles, so
    each particle keeps its block's waterline; FP64 run after the reorder
    bit-identical to the oracle run on the reordered state."""
    b = A.gen_batch(n_blocks=4, min_particles=800, max_particles=2500, seed=5)
    s = params.derive(params.load_config("iceblock_16k"))
    p, l = shuffled(b.particles, b.links, 8)
    r = np.random.default_rng(8)
    perm = r.permutation(b.particles.n)   # the same permutation shuffled() drew
    bid = b.block_id[perm]
    say("reorder batched", p.n)
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        A.set_blocks(ctx, bid, b.water)
        ctx.links_build()
        order, _ = ctx.reorder(s.search_distance)
        ctx.run(s.run, 80)
        gp = ctx.download_particles()
    q, m, o2, _ = orc.reorder(p, l, s.search_distance)
    assert np.array_equal(order, o2)
    sim = orc.OracleSim(q, m, s, block_id=bid[order].astype(np.int32), block_water=b.water)
    sim.run(80)
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(gp[f], sim.p[f]), f


def test_driver_interval_redistribution():
    """hdm.redistribute with dynamicRedistribution = false: the driver reorders
    on the first step and every redistributionInterval steps (HiDEMLib.h
    :560-587); the per-call loop and the fused runs (which end at those
    steps) agree bit for bit, and match hdm_reorder + hdm_run from Python."""
    import os
    cfg = os.path.join(params.CONFIG_DIR, "iceblock_16k.toml")
    over = ["--hdm.redistribute=true", "--distribution.dynamicRedistribution=false",
            "--distribution.redistributionInterval=25"]
    say("driver interval redistribution")
    out = {}
    for name, extra in (("per_call", ["--hdm.fusedSteps=false"]), ("fused", [])):
        with A.Sim(cfg, over + extra, precision=A.HDM_PREC_FP64) as sim:
            sim.steps(60)
            out[name] = sim.ctx.download_particles()
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        for k in (25, 25, 10):
            ctx.reorder(s.search_distance)
            ctx.run(s.run, k)
        g = ctx.download_particles()
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(out["per_call"][f], out["fused"][f]), f
        assert bits_equal(out["fused"][f], g[f]), f
