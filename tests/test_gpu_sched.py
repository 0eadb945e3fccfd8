"""GPU checks of the learned rebuild scheduler through the C ABI.

- the BF16 MFMA network (forward, gradient) against the CPU oracle's
  bfloat16-emulating form: the same operands are rounded at the same points,
  so only float32 accumulation order differs (tolerances below);
- device training lowers the loss of a known regression target;
- block features match the per-block displacement computed on the host;
- HDM_REBUILD_LEARNED (device-side exact test, batched launches, with and
  without a network) advances the state bit-identically to
  HDM_REBUILD_EXACT, with the same rebuild steps and broken links.
"""
import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def weights_struct(flat):
    return A.HdmMlpWeights.from_numpy(np.asarray(flat, np.float32))


def rand_inputs(n, seed):
    r = np.random.default_rng(seed)
    return (r.normal(0.3, 1.5, (n, M.NF))).astype(np.float32)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.fixture(scope="module")
def ctx():
    with A.Context(A.HDM_PREC_MIXED) as c:
        yield c


def test_mlp_forward_matches_oracle(ctx):
    say("mlp forward")
    w = M.init_weights(21, mean=np.full(M.NF, 0.3), std=np.full(M.NF, 1.5))
    M.unpack(w)["b1"][...] = 0.05
    X = rand_inputs(1000, 1)                     # not a multiple of the 32-row tile
    y = ctx.mlp_forward(weights_struct(w), X)
    ref = M.forward(w, X, emulate_bf16=True)
    err = np.abs(y - ref)
    # float32 accumulation order only; a bf16 rounding of a hidden unit may
    # flip by one ulp (2^-8 relative) for values on a rounding boundary
    assert np.max(err) <= 2e-2 * (1 + np.max(np.abs(ref))), np.max(err)
    assert np.median(err) <= 1e-5 * (1 + np.max(np.abs(ref)))
    exact = M.forward(w.astype(np.float64), X)
    assert np.max(np.abs(y - exact)) <= 5e-2 * (1 + np.max(np.abs(exact)))


def test_mlp_gradient_matches_oracle(ctx):
    say("mlp gradient")
    w = M.init_weights(22, mean=np.full(M.NF, 0.3), std=np.full(M.NF, 1.5))
    X = rand_inputs(777, 2)
    r = np.random.default_rng(3)
    t = r.normal(2.0, 1.0, 777).astype(np.float32)
    cen = (r.random(777) < 0.3).astype(np.uint8)
    g, loss = ctx.mlp_gradient(weights_struct(w), X, t, cen)
    lref, gref = M.loss_and_grad(w, X, t, cen, emulate_bf16=True)
    assert loss == pytest.approx(lref, rel=1e-3)
    o = 0
    for name, shp in M.SHAPES:
        size = int(np.prod(shp)) if shp else 1
        a, b = g[o:o + size], gref[o:o + size]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-6, name
        o += size
    # and the gradient points where the exact float64 gradient points
    _, gex = M.loss_and_grad(w.astype(np.float64), X, t, cen)
    cos = float(g @ gex / (np.linalg.norm(g) * np.linalg.norm(gex)))
    assert cos > 0.999


def test_mlp_training_lowers_loss(ctx):
    say("mlp training")
    X = rand_inputs(20000, 4)
    t = (3.0 + 0.8 * X[:, 0] - 0.4 * np.abs(X[:, 1]) + 0.3 * X[:, 2] * X[:, 3]).astype(np.float32)
    w0 = M.init_weights(23, mean=X.mean(0), std=X.std(0))
    _, loss0 = ctx.mlp_gradient(weights_struct(w0), X, t)
    w = weights_struct(w0)
    loss = ctx.mlp_train(w, X, t, epochs=30, batch=1024, lr=3e-3)
    say("loss", loss0 / len(t), "->", loss)
    assert loss < 0.05 * loss0 / len(t)
    # the returned weights are the ones the device used
    y = ctx.mlp_forward(w, X[:512])
    ref = M.forward(w.to_numpy(), X[:512], emulate_bf16=True)
    assert np.max(np.abs(y - ref)) <= 2e-2 * (1 + np.max(np.abs(ref)))


def test_block_features_match_host():
    say("block features")
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    with A.Context(A.HDM_PREC_MIXED) as c:
        c.upload(p, l)
        c.links_build()
        st, log = c.run(s.run, 30, log_rebuilds=True)
        assert log[0] == 1 and log[1:].sum() == 0   # one rebuild, at step 0: x0 = initial positions
        f, d, v = c.sched_features(s.run, 30)
        g = c.grid_info()
        q = c.download_particles()
    x = np.stack([q["xCord"], q["yCord"], q["zCord"]], 1)
    x0 = np.stack([p["xCord"], p["yCord"], p["zCord"]], 1)
    edge = 4 * g.spacing
    BX, BY, BZ = -(-g.X // 4), -(-g.Y // 4), -(-g.Z // 4)
    assert f.shape == (BX * BY * BZ, 16)
    ijk = np.floor((x - np.array([g.xMin, g.yMin, g.zMin])) / edge).astype(np.int64)
    ijk = np.clip(ijk, 0, np.array([BX, BY, BZ]) - 1)
    blk = ijk[:, 0] + BX * (ijk[:, 1] + BY * ijk[:, 2])
    disp = np.linalg.norm(x - x0, axis=1)
    want = np.zeros(BX * BY * BZ)
    np.maximum.at(want, blk, disp)
    cnt = np.bincount(blk, minlength=BX * BY * BZ)
    assert np.array_equal(v != 0, cnt > 0)
    np.testing.assert_allclose(d, want, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(f[:, 10], cnt / 150.0, rtol=1e-6)
    budget = 0.5 * (s.run.interactDist - s.run.force.attractDistance)
    np.testing.assert_allclose(f[v != 0, 4], (want / budget)[v != 0], rtol=1e-5)
    assert np.all(np.isfinite(f))


def _weak_block(skin=0.004):
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    l.maxLoad[:] *= 0.25
    # a thin skin beyond the attraction range: the exact test trips every few dozen steps
    s.run.interactDist = s.run.force.attractDistance + skin
    return s, p, l


def _run(prec, s, p, l, policy, chunks, sched=None, model=None, exec_=None):
    rp = A.HdmRunParams.from_buffer_copy(s.run)
    rp.policy = policy
    logs = []
    with A.Context(prec) as c:
        c.upload(p, l)
        c.links_build()
        if sched:
            c.sched_configure(**sched)
        if model is not None:
            c.sched_set_model(model)
        if exec_:
            c.set_exec(**exec_)
        for k in chunks:
            st, log = c.run(rp, k, log_rebuilds=True)
            logs.append(log)
        stats = c.sched_stats()
        stats.update(c.exec_info())
        return np.concatenate(logs), c.download_particles(), c.download_links(), stats


def _biased(w, b3):
    """An untrained network whose predictions sit near b3 (log2 steps), so the
    horizons vary around 0.5 * 2^b3 instead of collapsing to minBatch."""
    w = w.copy()
    w[M.N_PARAMS - 1] = b3
    return w


@pytest.mark.parametrize("prec", [A.HDM_PREC_FP64, A.HDM_PREC_MIXED])
def test_learned_policy_is_bit_identical_to_exact(prec):
    s, p, l = _weak_block()
    say("learned vs exact", prec)
    rl_e, pe, le, _ = _run(prec, s, p, l, A.HDM_REBUILD_EXACT, [400])
    assert 2 <= rl_e.sum() <= 100, rl_e.sum()   # guard trips inside batches and rebuilds between them
    cases = [
        dict(chunks=[400], sched=dict(min_batch=64, max_batch=64)),                       # fixed batches
        dict(chunks=[150, 1, 249], sched=dict(min_batch=1, max_batch=37, safety=0.5),     # learned, odd sizes
             model=weights_struct(_biased(M.init_weights(31, mean=np.zeros(16), std=np.ones(16)), 5.0))),
        dict(chunks=[200, 200], sched=dict(min_batch=16, max_batch=48),                   # autotuned, graphs
             exec_=dict(autotune=True, graphs=True)),
        dict(chunks=[150, 1, 249], sched=dict(min_batch=16, max_batch=64),                # one launch per batch
             exec_=dict(persistent=True)),
    ]
    for case in cases:
        rl, pl, ll, stats = _run(prec, s, p, l, A.HDM_REBUILD_LEARNED, **case)
        say("  ", case["sched"], stats)
        assert np.array_equal(rl, rl_e)
        for f in A.PARTICLE_FIELDS:
            assert bits_equal(pl[f], pe[f]), f
        assert np.array_equal(ll.state, le.state)
        for k in range(6):
            assert bits_equal(ll.ct[k], le.ct[k])
        assert stats["steps_executed"] == 400
        assert stats["steps_launched"] >= 400
        assert stats["host_checks"] < 400
        if case.get("exec_", {}).get("graphs"):
            assert stats["tuned"] and stats["graph_replays"] > 10 and stats["graph_captures"] >= 1
        if case.get("exec_", {}).get("persistent"):
            assert stats["persistent_launches"] > 3


def test_label_tracking_matches_host():
    say("label tracking")
    s, p, l = _weak_block(skin=0.0003)
    rp = A.HdmRunParams.from_buffer_copy(s.run)
    rp.policy = A.HDM_REBUILD_EXACT
    with A.Context(A.HDM_PREC_FP64) as c:
        c.upload(p, l)
        c.links_build()
        st, log = c.run(rp, 5, log_rebuilds=True)
        g = c.grid_info()
        f, d, v = c.sched_features(rp, 5)
        q0 = c.download_particles()
        c.sched_label_begin(3, rp)
        xs = []
        for k in range(100):
            c.run(rp, 1)
            c.sched_label_track()
            q = c.download_particles()
            xs.append(np.stack([q["xCord"], q["yCord"], q["zCord"]], 1))
        first = c.sched_label_end(3, len(v))
    x0 = np.stack([q0["xCord"], q0["yCord"], q0["zCord"]], 1)
    edge = 4 * g.spacing
    BX, BY, BZ = -(-g.X // 4), -(-g.Y // 4), -(-g.Z // 4)
    ijk = np.clip(np.floor((x0 - np.array([g.xMin, g.yMin, g.zMin])) / edge).astype(np.int64), 0,
                  np.array([BX, BY, BZ]) - 1)
    blk = ijk[:, 0] + BX * (ijk[:, 1] + BY * ijk[:, 2])
    budget = np.float32(0.5 * (rp.interactDist - rp.force.attractDistance))
    remaining = np.maximum(budget - d, np.float32(0))
    want = np.full(len(v), -1)
    for k, x in enumerate(xs, start=1):
        moved = np.linalg.norm(x - x0, axis=1)
        hit = np.zeros(len(v), bool)
        np.logical_or.at(hit, blk, moved >= remaining[blk].astype(np.float64))
        want = np.where((want < 0) & hit, k, want)
    assert (want > 0).sum() > 0   # free-falling particles cover the budget in ~50 steps
    assert np.array_equal(first, want)


def test_trained_network_gains_over_fixed_batches():
    """What the trained network gains: a network trained on the device
    (sched_train.train) on the first steps of a small glacier slab, run on
    the following steps, against the same device-side exact test with fixed
    batches (no network) and with a host check after every step.  All
    three end bit-identical; the network needs far fewer host checks than
    the per-step test and wastes fewer launches (steps launched after the
    guard tripped) than fixed batches."""
    from hidem_amd import sched_train
    cfg = params.load_config("glacier_8m")
    setup, p, l = params.build_inputs(cfg, [40, 6, 8])
    # a thin skin: rebuilds every few dozen steps, so both the labels and the
    # evaluation see many trips
    setup.run.interactDist = setup.run.force.attractDistance + 0.004
    say("trained network gain", p.n, l.n)
    X, T, Cn, rebuilds, state, ts = sched_train.generate(setup, p, l, 1500, 25)
    say("  rows", len(X), "censored", float(Cn.mean()), "rebuilds", rebuilds)
    assert rebuilds >= 5 and (Cn == 0).sum() > 100
    w, loss0, loss = sched_train.train(X, T, Cn, epochs=20)
    say("  loss", loss0, "->", loss)
    assert loss < loss0
    ev = sched_train.evaluate(setup, state, ts, 1500, w, chunk=500)
    for k, v in ev.items():
        say("  ", k, v)
    assert ev["exact"]["rebuilds"] >= 5
    assert ev["learned"]["bit_identical_to_exact"] and ev["learned_no_model"]["bit_identical_to_exact"]
    assert ev["learned"]["rebuilds"] == ev["exact"]["rebuilds"]
    assert ev["learned"]["host_checks"] * 4 < ev["exact"]["host_checks"]
    assert ev["learned"]["wasted_launches"] < ev["learned_no_model"]["wasted_launches"]
