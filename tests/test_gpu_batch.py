"""Batched form on the GPU: many independent ice blocks, each with its own
waterline and fracture thresholds, in one state.

- FP64: bit-identical to the CPU oracle run with the same per-block
  waterlines over 150 steps (positions, link states);
- mixed: the graph-replayed, autotuned HDM_REBUILD_LEARNED run ends
  bit-identical to the eager HDM_REBUILD_EXACT run.
"""
import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.fixture(scope="module")
def batch():
    b = A.gen_batch(n_blocks=6, min_particles=1000, max_particles=4000, seed=3)
    s = params.derive(params.load_config("iceblock_16k"))
    return s, b


def test_batch_fp64_matches_oracle(batch, orc):
    s, b = batch
    say("batch fp64 vs oracle", b.particles.n, b.links.n)
    sim = orc.OracleSim(b.particles, b.links, s, block_id=b.block_id, block_water=b.water)
    sim.run(150)
    with A.Context(A.HDM_PREC_FP64) as c:
        c.upload(b.particles, b.links)
        A.set_blocks(c, b.block_id, b.water)
        st = c.run(s.run, 150)
        q = c.download_particles()
        ql = c.download_links()
    assert st.rebuilds == sim.rebuilds
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(q[f], sim.p[f]), f
    assert np.array_equal(ql.state, sim.l.state)
    # the waterlines differ per block, so the blocks' vertical motion differs
    dz = q["zCord"] - b.particles["zCord"]
    per_block = [dz[b.block_id == k].mean() for k in range(6)]
    assert np.ptp(per_block) > 0


def test_batch_graph_autotuned_matches_eager(batch):
    s, b = batch
    say("batch learned+graph vs exact")
    out = {}
    for name, policy, exec_ in (("exact", A.HDM_REBUILD_EXACT, None),
                                ("learned", A.HDM_REBUILD_LEARNED, dict(autotune=True, graphs=True))):
        rp = A.HdmRunParams.from_buffer_copy(s.run)
        rp.policy = policy
        rp.interactDist = rp.force.attractDistance + 0.004     # frequent rebuilds: recaptures and guard trips
        with A.Context(A.HDM_PREC_MIXED) as c:
            c.upload(b.particles, b.links)
            A.set_blocks(c, b.block_id, b.water)
            c.sched_configure(min_batch=16, max_batch=64)
            if exec_:
                c.set_exec(**exec_)
            logs = [c.run(rp, k, log_rebuilds=True)[1] for k in (120, 180)]
            out[name] = (np.concatenate(logs), c.download_particles(), c.download_links(), c.exec_info())
    (le, pe, ke, _), (ll, pl, kl, info) = out["exact"], out["learned"]
    say("  ", info)
    assert le.sum() >= 2
    assert np.array_equal(le, ll)
    for f in A.PARTICLE_FIELDS:
        assert bits_equal(pl[f], pe[f]), f
    assert np.array_equal(kl.state, ke.state)
    for k in range(6):
        assert bits_equal(kl.ct[k], ke.ct[k])
    assert info["tuned"] and info["graph_replays"] > 0
