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


@pytest.mark.parametrize("exec_", [dict(autotune=True, graphs=True), dict(persistent=True)],
                         ids=["graphs", "persistent"])
def test_batch_graph_autotuned_matches_eager(batch, exec_):
    s, b = batch
    say("batch learned vs exact", exec_)
    out = {}
    for name, policy, ex in (("exact", A.HDM_REBUILD_EXACT, None), ("learned", A.HDM_REBUILD_LEARNED, exec_)):
        rp = A.HdmRunParams.from_buffer_copy(s.run)
        rp.policy = policy
        rp.interactDist = rp.force.attractDistance + 0.004     # frequent rebuilds: recaptures and guard trips
        with A.Context(A.HDM_PREC_MIXED) as c:
            c.upload(b.particles, b.links)
            A.set_blocks(c, b.block_id, b.water)
            c.sched_configure(min_batch=16, max_batch=64)
            if ex:
                c.set_exec(**ex)
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
    if exec_.get("graphs"):
        assert info["tuned"] and info["graph_replays"] > 0
    else:
        assert info["persistent_launches"] > 0


def test_batch_at_config_sizes_against_fp64_oracle(orc):
    """Blocks of the batched config's sizes (log-uniform in [5k, 200k], four
    blocks, each with its own threshold mean and waterline), 300 steps against
    the FP64 oracle of the same batch: the FP64 form bit-identical, and the
    FP32 form (graph-replayed, autotuned, device guard) within position
    max|dx| / max|x| <= 1e-5 and broken-link count within 0.1% (or one link)."""
    b = A.gen_batch(n_blocks=4, min_particles=5000, max_particles=200000, seed=5)
    s = params.derive(params.load_config("iceblock_16k"))
    counts = np.bincount(b.block_id)
    say("batch at config sizes", counts.tolist(), b.links.n)
    assert counts.min() >= 5000 and counts.max() <= 200000
    steps = 300
    sim = orc.OracleSim(b.particles, b.links, s, block_id=b.block_id, block_water=b.water)
    sim.run(steps)
    x64 = np.stack([sim.p["xCord"], sim.p["yCord"], sim.p["zCord"]], 1)
    b64 = int((sim.l.state == 0).sum())
    with A.Context(A.HDM_PREC_FP64) as c:
        c.upload(b.particles, b.links)
        A.set_blocks(c, b.block_id, b.water)
        c.run(s.run, steps)
        q = c.download_particles()
        ql = c.download_links()
    for f in ("xCord", "yCord", "zCord", "xTheta", "yTheta", "zTheta"):
        assert bits_equal(q[f], sim.p[f]), f
    assert np.array_equal(ql.state, sim.l.state)
    rp = A.HdmRunParams.from_buffer_copy(s.run)
    rp.policy = A.HDM_REBUILD_LEARNED
    with A.Context(A.HDM_PREC_MIXED) as c:
        c.upload(b.particles, b.links)
        A.set_blocks(c, b.block_id, b.water)
        c.sched_configure(min_batch=16, max_batch=64)
        c.set_exec(autotune=True, graphs=True)
        c.run(rp, steps)
        q = c.download_particles()
        b32 = c.count_broken()
        info = c.exec_info()
    x32 = np.stack([q["xCord"], q["yCord"], q["zCord"]], 1)
    pos = float(np.linalg.norm(x32 - x64, axis=1).max() / np.abs(x64).max())
    disp = float(np.abs(x64 - np.stack([b.particles[f] for f in ("xCord", "yCord", "zCord")], 1)).max())
    say(f"  FP32 batched vs FP64 oracle: position {pos:.3e} (max displacement {disp:.3e} m), "
        f"broken {b32} / {b64}, graph replays {info['graph_replays']}")
    sim.close()
    assert info["tuned"] and info["graph_replays"] > 0
    assert pos <= 1e-5
    assert abs(b32 - b64) <= max(1, 1e-3 * b64)
