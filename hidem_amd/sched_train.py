#!/usr/bin/env python3
"""Trains and evaluates the learned rebuild scheduler on the GPU.

1. Trajectory: the mixed-precision form of the config (default: the 8M
   glacier slab) runs steps [0, train_steps) under the exact rebuild test;
   every `every` steps the 16 block features are recorded.
2. Labels: for a decision after step t and a block b, the number of steps
   until some particle of b has moved b's remaining skin budget (the budget
   (interactDist - attractDistance)/2 minus b's displacement since the
   rebuild at t) from where it was at t -- by the triangle inequality a lower
   bound of the steps until b alone trips the exact test, tracked on the
   device for every block (hdm_sched_label_*).  Rows whose block has not moved
   that far after 32 decision intervals are censored at that age (the true
   value is only known to be larger).
3. Training: hdm_mlp_train (BF16 MFMA forward/backward + Adam on the device)
   on the standardised, shuffled rows.
4. Evaluation: from the state after train_steps, eval_steps more steps under
   (a) the reference heuristic (HiDEMLib.h:548-559), (b) the exact test with a
   host check every step, (c) HDM_REBUILD_LEARNED with fixed batches and no
   network, (d) HDM_REBUILD_LEARNED with the trained network, (e) as (d)
   with captured graphs and the autotuned launch shape (the bench's form).
   Reports rebuilds, host checks, launched vs executed steps and wall time;
   (c), (d) and (e) must end bit-identical to (b) -- the device-side exact
   test decides every rebuild, so no step ever runs on a stale neighbour
   list (missed contacts: 0).

With --eval-only the saved network is evaluated again (steps 1 and 4 only;
the state after train_steps comes from a plain run under the exact test)
and the report's evaluation is replaced.

Usage: python -m hidem_amd.sched_train [--config glacier_8m] [--out PATH] [--eval-only]
Prints progress to stderr and one JSON report line to stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from hidem_amd import _abi as A, params  # noqa: E402
else:
    from . import _abi as A, params

MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def save_model(path, w: A.HdmMlpWeights, meta: dict):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    w.to_numpy().astype("<f4").tofile(path)
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)


def load_model(path) -> A.HdmMlpWeights:
    flat = np.fromfile(path, dtype="<f4")
    return A.HdmMlpWeights.from_numpy(flat)


def generate(setup, p, l, steps, every, slots=32):
    """Trajectory under the exact test with a label reference at every decision.

    Every `every` steps the block features are evaluated and a label reference
    is taken (hdm_sched_label_begin); after each step the device records, per
    block, the first step at which a particle of the block has moved the
    block's remaining skin budget from its reference position.  A reference
    closes after slots * every steps (censoring its open blocks at that age)
    or at the end of the trajectory."""
    rp = A.HdmRunParams.from_buffer_copy(setup.run)
    rp.policy = A.HDM_REBUILD_EXACT
    X, T, Cn = [], [], []
    open_refs = {}          # slot -> (features, valid, start step)
    rebuilds = 0
    since = 0
    t0 = time.perf_counter()

    def close(c, slot, step):
        f, v, start = open_refs.pop(slot)
        first = c.sched_label_end(slot, len(v))
        age = step - start
        cens = first < 0
        k = np.where(cens, age, first)
        keep = (v != 0) & (k > 0)
        X.append(f[keep])
        T.append(np.log2(k[keep]).astype(np.float32))
        Cn.append(cens[keep].astype(np.uint8))

    with A.Context(A.HDM_PREC_MIXED) as c:
        c.upload(p, l)
        c.links_build()
        for step in range(steps):
            st, rl = c.run(rp, 1, log_rebuilds=True)
            if rl[0]:
                rebuilds += 1
                since = 0
            since += 1
            c.sched_label_track()
            if step % every == every - 1:
                slot = (step // every) % slots
                if slot in open_refs:
                    close(c, slot, step + 1)
                f, d, v = c.sched_features(rp, since)
                c.sched_label_begin(slot, rp)
                open_refs[slot] = (f, v, step + 1)
            if step % 250 == 249:
                log(f"# trajectory step {step + 1}/{steps}, {rebuilds} rebuilds, {time.perf_counter() - t0:.0f} s")
        for slot in list(open_refs):
            close(c, slot, steps)
        state = (c.download_particles(), c.download_links())
        ts = st.timestep
    return np.concatenate(X), np.concatenate(T), np.concatenate(Cn), rebuilds, state, ts


def evaluate(setup, state, ts, steps, model, chunk=1000):
    p, l = state
    results = {}
    runs = [("heuristic", A.HDM_REBUILD_HEURISTIC, None), ("exact", A.HDM_REBUILD_EXACT, None),
            ("learned_no_model", A.HDM_REBUILD_LEARNED, None), ("learned", A.HDM_REBUILD_LEARNED, model),
            ("learned_graphs", A.HDM_REBUILD_LEARNED, model)]
    final = {}
    for name, policy, w in runs:
        rp = A.HdmRunParams.from_buffer_copy(setup.run)
        rp.policy = policy
        with A.Context(A.HDM_PREC_MIXED) as c:
            c.upload(p, l)
            c.links_build()
            c.set_timestep(ts)
            c.sched_configure(block_cells=4, min_batch=8, max_batch=256, safety=0.5)
            if w is not None:
                c.sched_set_model(w)
            if name == "learned_graphs":
                c.set_exec(autotune=True, graphs=True)
            rebuilds, secs = 0, 0.0
            done = 0
            while done < steps:
                k = min(chunk, steps - done)
                st = c.run(rp, k)
                rebuilds += st.rebuilds
                secs += st.seconds
                done += k
                log(f"# eval {name}: step {ts + done}, rebuilds {rebuilds}, {secs:.1f} s")
            stats = c.sched_stats()
            q = c.download_particles()
            broken = c.count_broken()
        r = {"rebuilds": rebuilds, "seconds": secs, "ms_per_step": 1e3 * secs / steps, "broken": broken}
        if policy == A.HDM_REBUILD_LEARNED:
            r.update(stats)
            r["wasted_launches"] = stats["steps_launched"] - stats["steps_executed"]
        else:
            r["host_checks"] = steps if policy == A.HDM_REBUILD_EXACT else steps // 10
        final[name] = q
        results[name] = r
    x = lambda q: np.stack([q["xCord"], q["yCord"], q["zCord"]], 1)  # noqa: E731
    for name in ("learned_no_model", "learned", "learned_graphs"):
        results[name]["bit_identical_to_exact"] = bool(
            np.array_equal(x(final[name]).view(np.uint64), x(final["exact"]).view(np.uint64)))
        results[name]["missed_contacts"] = 0 if results[name]["bit_identical_to_exact"] else None
    return results


def advance(setup, p, l, steps, chunk=1000):
    """The state after `steps` steps under the exact test (the trajectory
    generate() follows, without its label tracking)."""
    rp = A.HdmRunParams.from_buffer_copy(setup.run)
    rp.policy = A.HDM_REBUILD_EXACT
    with A.Context(A.HDM_PREC_MIXED) as c:
        c.upload(p, l)
        c.links_build()
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            st = c.run(rp, k)
            done += k
            log(f"# advance: step {done}/{steps}")
        return (c.download_particles(), c.download_links()), st.timestep


def train(X, T, Cn, epochs=8, max_rows=4_000_000, seed=0):
    """Standardises the rows, initialises the network (He normal weights, the
    output bias at the mean uncensored target) and trains it on the device
    (hdm_mlp_train).  Returns (weights, mean loss before, mean loss after)."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(X))[:max_rows]
    X, T, Cn = X[order], T[order], Cn[order]
    mean, std = X.mean(0), X.std(0)
    std = np.where(std > 1e-6, std, 1.0).astype(np.float32)
    w0 = np.zeros(A.C.sizeof(A.HdmMlpWeights) // 4, np.float32)
    r = np.random.default_rng(seed + 1)
    w0[:16 * 64] = r.normal(0, np.sqrt(2 / 16), 16 * 64)
    o = 16 * 64 + 64
    w0[o:o + 64 * 64] = r.normal(0, np.sqrt(2 / 64), 64 * 64)
    o += 64 * 64 + 64
    w0[o:o + 64] = r.normal(0, np.sqrt(1 / 64), 64)
    w0[o + 64] = float(T[Cn == 0].mean()) if (Cn == 0).any() else float(T.mean())
    w0[o + 65:o + 81] = mean
    w0[o + 81:o + 97] = std
    w = A.HdmMlpWeights.from_numpy(w0)
    with A.Context(A.HDM_PREC_MIXED) as c:
        _, loss0 = c.mlp_gradient(w, X, T, Cn)
        loss = c.mlp_train(w, X, T, Cn, epochs=epochs, batch=4096, lr=1e-3)
    return w, loss0 / len(X), loss


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", default="glacier_8m")
    ap.add_argument("--cells", type=int, nargs=3, default=None)
    ap.add_argument("--train-steps", type=int, default=5000)
    ap.add_argument("--eval-steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=50, help="steps between decision rows (label references)")
    ap.add_argument("--max-rows", type=int, default=4_000_000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--out", default=None, help="model file (default hidem_amd/models/sched_<config>.bin)")
    ap.add_argument("--eval-only", action="store_true", help="evaluate the saved network again")
    args = ap.parse_args(argv)

    cfg = params.load_config(args.config)
    setup, p, l = params.build_inputs(cfg, args.cells)
    log(f"# {args.config}: {p.n} particles, {l.n} links")
    t0 = time.perf_counter()
    out = args.out or os.path.join(MODEL_DIR, f"sched_{args.config}.bin")
    if args.eval_only:
        with open(os.path.splitext(out)[0] + ".json") as f:
            report = json.load(f)
        state, ts = advance(setup, p, l, report["train_steps"])
        ev = evaluate(setup, state, ts, report["eval_steps"], load_model(out))
        report["evaluation"] = ev
        with open(os.path.splitext(out)[0] + ".json", "w") as f:
            json.dump(report, f, indent=1)
        print(json.dumps(report), flush=True)
        return report
    X, T, Cn, rebuilds, state, ts = generate(setup, p, l, args.train_steps, args.every)
    log(f"# {len(X)} rows, {Cn.mean():.3f} censored, {rebuilds} rebuilds, {time.perf_counter() - t0:.0f} s")
    w, loss0, loss = train(X, T, Cn, epochs=args.epochs, max_rows=args.max_rows)
    log(f"# training: mean loss {loss0:.4f} -> {loss:.4f}")
    ev = evaluate(setup, state, ts, args.eval_steps, w)
    report = {"config": args.config, "particles": p.n, "links": l.n, "train_steps": args.train_steps,
              "eval_steps": args.eval_steps, "rows": int(min(len(X), args.max_rows)), "censored_fraction": float(Cn.mean()),
              "rebuilds_in_training": rebuilds, "loss_initial": loss0,
              "loss_trained": loss, "evaluation": ev, "seconds": time.perf_counter() - t0}
    save_model(out, w, report)
    print(json.dumps(report), flush=True)
    return report


if __name__ == "__main__":
    main()
