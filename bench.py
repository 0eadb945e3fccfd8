#!/usr/bin/env python3
"""Timestep throughput of the MI355X HiDEM hot path on the 8M-particle glacier slab.

A "step" is one full timestep of the hot path -- rebuild check (and rebuild
when due), contact + beam force gather, time integration, link break -- over
all 8,000,000 particles / ~48M links resident in HBM.  Prints one JSON line.

Each form runs on its own context from the same generated inputs:
  1. launch-shape autotuning: whole tuning batches (every shape three times,
     identical results by construction) until a shape is picked -- never
     inside the warmup or the timed steps, whatever --warmup is;
  2. W warmup steps and K timed steps with the form's own rebuild policy,
     bracketed by device synchronisation;
  3. the state after the window (positions, the last step's force, broken
     links) for the error report, and a timed forced rebuild.
Every form has run the same number of steps when its state is read, so the
mixed forms' error is taken against the FP64 form at the same step.

value      whole-job particle-steps/s of the headline form, mixed+learned
           (FP32 force arithmetic, FP64 integration and strain, learned
           rebuild scheduler with the exact skin test on the device):
           N_particles * K / wall time of the K timed steps.  The plain FP64
           method (reference rebuild heuristic) and the mixed form with a
           host-checked exact test run beside it under "forms".
roofline   for the dominant kernel (k_step: link break of the previous step,
           force gather and integration in one launch): algorithmic bytes per
           launch (compulsory traffic of the arrays it must touch, see
           step_bytes) / its mean launch time from HIP events recorded on the
           context's stream during the timed steps, against 8 TB/s.  traffic
           is the HBM bytes per launch measured with rocprofv3 counters;
           copy_probe_GBps the streaming copy probe on the same box, the
           practical ceiling (traffic_over_copy_probe: the step's measured
           traffic rate against it).
error      each mixed form against the FP64 form after the window: position
           max|dx| / max|x|, force ||dF||/||F|| (translational, last step),
           broken-link count difference.
amortised  the window of an intact slab holds no rebuild; the amortised
           step adds a forced rebuild (timed here) at the rebuild rate of the
           policy measured over steps 5k-15k of the same slab.
cpu_baseline  the CPU oracle (a C/OpenMP restatement of the reference's CPU
           path) timed on this host: a bounded sample of the 8M slab built
           with the reference's release flags (-O3 -ffast-math), the same
           sample without -ffast-math, and the 16k block (2,000 steps) with
           both builds.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from hidem_amd import _abi as A  # noqa: E402
from hidem_amd import params  # noqa: E402
from hidem_amd.roofline import HBM_PEAK_GBS, reference_list_entries, step_bytes  # noqa: E402

METRIC = "particle-steps/s at 8M particles (mixed+learned vs plain FP64); % HBM peak"
TUNE_BATCHES_MAX = 6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Forms: (name, precision, rebuild policy, uses the trained scheduler network)
FORMS = {
    "fp64": (A.HDM_PREC_FP64, A.HDM_REBUILD_HEURISTIC, False),      # plain FP64 method, reference gridService
    "mixed": (A.HDM_PREC_MIXED, A.HDM_REBUILD_EXACT, False),        # precision form, exact test checked each step
    "mixed+learned": (A.HDM_PREC_MIXED, A.HDM_REBUILD_LEARNED, True),  # precision form + learned scheduler
}
# rebuild-rate source of each policy: the 10,000-step evaluation (steps 5k-15k) of the same slab
POLICY_EVAL = {A.HDM_REBUILD_HEURISTIC: "heuristic", A.HDM_REBUILD_EXACT: "exact", A.HDM_REBUILD_LEARNED: "learned"}


def sched_model(config):
    path = os.path.join(ROOT, "hidem_amd", "models", f"sched_{config}.bin")
    if not os.path.exists(path):
        return None, None, None
    from hidem_amd import sched_train
    report = None
    rj = path[:-4] + ".json"
    if os.path.exists(rj):
        with open(rj) as f:
            report = json.load(f)
    return sched_train.load_model(path), os.path.relpath(path, ROOT), report


def tune_shape(ctx, rp):
    """Runs whole tuning batches of the learned policy (the only one that
    tunes) until the context has picked a launch shape; returns the steps
    this advanced the state by."""
    rw = A.HdmRunParams.from_buffer_copy(rp)
    rw.policy = A.HDM_REBUILD_LEARNED
    batch = 3 * len(ctx.exec_info()["tune_ms"])
    steps = 0
    while not ctx.exec_info()["tuned"]:
        if steps >= TUNE_BATCHES_MAX * batch:
            raise RuntimeError(f"launch-shape autotuning did not settle in {steps} steps")
        ctx.run(rw, batch)
        steps += batch
    return steps


def timed_rebuild(ctx, rp, repeats=3):
    """Wall time of a forced rebuild (grid, neighbour list, rebuild snapshot),
    device-synchronised; best of `repeats`."""
    best = None
    for _ in range(repeats):
        ctx.sync()
        t0 = time.perf_counter()
        ctx.grid_build(rp.searchDistance)
        ctx.neighbours_build(rp.interactDist)
        ctx.skin_snapshot()
        ctx.sync()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best


def run_gpu(setup, p, l, prec, policy, model, steps, warmup):
    rp = A.HdmRunParams.from_buffer_copy(setup.run)
    rp.policy = policy
    with A.Context(prec) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.sched_configure(block_cells=4, min_batch=8, max_batch=256, safety=0.5)
        if model is not None:
            ctx.sched_set_model(model)
        # every form autotunes its launch shape first (all shapes give identical results)
        ctx.set_exec(autotune=True, graphs=policy == A.HDM_REBUILD_LEARNED)
        tune_steps = tune_shape(ctx, rp)
        if warmup > 0:
            ctx.run(rp, warmup)
        ctx.sync()
        s0 = ctx.sched_stats()
        ctx.set_timing(True)
        ctx.reset_kernel_times()
        t0 = time.perf_counter()
        st = ctx.run(rp, steps)
        ctx.sync()
        t1 = time.perf_counter()
        kt = ctx.kernel_times()
        ctx.set_timing(False)
        s1 = ctx.sched_stats()
        exec_info = ctx.exec_info()
        counts, _, _ = ctx.neighbours_download()
        n_neigh = reference_list_entries(int(counts.sum()), l.n, exec_info)
        broken = ctx.count_broken()
        dev_bytes = ctx.device_bytes()
        q = ctx.download_particles()
        state = {"x": np.stack([q["xCord"], q["yCord"], q["zCord"]], 1),
                 "f": np.stack([q["xForce"], q["yForce"], q["zForce"]], 1).astype(np.float64),
                 "broken": broken, "step": st.timestep}
        del q
        rebuild_s = timed_rebuild(ctx, rp)
    sched = {k: s1[k] - s0[k] for k in s1}
    return dict(wall=t1 - t0, stats=st, kernels=kt, n_neigh=n_neigh, broken=broken, dev_bytes=dev_bytes,
                sched=sched, exec=exec_info, state=state, tune_steps=tune_steps, rebuild_s=rebuild_s)


def error_vs_fp64(s64, s):
    """Error of a form's state against the FP64 form's state at the same step."""
    if s["step"] != s64["step"]:
        return {"comparable": False, "step": s["step"], "step_fp64": s64["step"]}
    err = np.linalg.norm(s["x"] - s64["x"], axis=1)
    fnorm = float(np.linalg.norm(s64["f"]))
    b64, b = s64["broken"], s["broken"]
    return {"step": int(s["step"]),
            "pos_max_rel": float(err.max() / np.abs(s64["x"]).max()),
            "pos_max_abs_m": float(err.max()),
            "force_rel_l2": float(np.linalg.norm(s["f"] - s64["f"]) / fnorm) if fnorm > 0 else None,
            "broken_fp64": int(b64), "broken": int(b),
            "broken_rel_diff": float(abs(b - b64) / b64) if b64 else float(b != 0)}


def cpu_sample(setup, p, l, steps, fast):
    from oracle import oracle as orc  # checker / baseline only
    t0 = time.perf_counter()
    sim = orc.OracleSim(p, l, setup, fast=fast)
    sim.run(1)  # builds grid, neighbour list and link lists, one step
    t1 = time.perf_counter()
    sim.run(steps)
    t2 = time.perf_counter()
    threads = orc.num_threads(fast)
    sim.close()
    return {"value": p.n * steps / (t2 - t1), "steps": steps, "seconds": t2 - t1, "setup_seconds": t1 - t0,
            "threads": threads}


def cpu_baseline(setup, p, l, steps, exact_steps, block_steps):
    """The oracle on this host's cores: the slab with both builds, the 16k block with both."""
    import platform
    out = {}
    log(f"cpu baseline: oracle -ffast-math, {steps} steps of the slab")
    out["slab_fast_math"] = cpu_sample(setup, p, l, steps, True)
    if exact_steps > 0:
        log(f"cpu baseline: oracle without -ffast-math, {exact_steps} steps of the slab")
        out["slab_ieee"] = cpu_sample(setup, p, l, exact_steps, False)
    if block_steps > 0:
        cfg0 = params.load_config("iceblock_16k")
        s0, p0, l0 = params.build_inputs(cfg0)
        for key, fast in (("block16k_fast_math", True), ("block16k_ieee", False)):
            log(f"cpu baseline: {key}, {block_steps} steps")
            out[key] = cpu_sample(s0, p0, l0, block_steps, fast)
    cpu_model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    out["cpu_model"] = cpu_model
    out["host_cpus"] = os.cpu_count()
    return out


def probe_traffic(setup, p, l, forms):
    """Kernels of known size for calibrating the FETCH_SIZE / WRITE_SIZE
    counters, then fused step launches on the real workload."""
    for prec in forms:
        # the headline FP32 form runs the learned policy's guarded step (here
        # as per-step launches, so every dispatch is counted)
        rp = A.HdmRunParams.from_buffer_copy(setup.run)
        if prec == A.HDM_PREC_MIXED:
            rp.policy = A.HDM_REBUILD_LEARNED
        with A.Context(prec) as ctx:
            ctx.upload(p, l)
            ctx.links_build()
            # the same autotuned launch shape the timed bench steps use
            ctx.set_exec(autotune=True)
            tune_shape(ctx, rp)
            if prec == forms[0]:
                g = ctx.bw_probe(A.HDM_PROBE_COPY, 1 << 30, 1)  # k_probe_copy: 512 MiB read + 512 MiB written
                log(f"copy probe 512 MiB read + 512 MiB written: {g:.0f} GB/s")
            ctx.run(rp, 3)
            log(f"launch shape {ctx.exec_info()['name']}")
            ctx.sync()
            counts, _, _ = ctx.neighbours_download()
            n_neigh = reference_list_entries(int(counts.sum()), l.n, ctx.exec_info())
            b = step_bytes(p.n, n_neigh, l.n, prec)
            print(json.dumps({"precision": "fp64" if prec == A.HDM_PREC_FP64 else "mixed",
                              "step_algorithmic_bytes": b, "neighbours": n_neigh, "links": l.n}),
                  flush=True)


def ratio(a, b):
    return a / b if b else None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--config", default="glacier_8m")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="oracle steps of the slab in the CPU baseline, -ffast-math build (0: no CPU baseline)")
    ap.add_argument("--cpu-ieee-steps", type=int, default=20,
                    help="oracle steps of the slab, build without -ffast-math (0: skip)")
    ap.add_argument("--cpu-block-steps", type=int, default=2000,
                    help="oracle steps of the 16k block, both builds (0: skip)")
    ap.add_argument("--cells", type=int, nargs=3, default=None, help="override unit cells (smaller runs)")
    ap.add_argument("--probe-traffic", action="store_true",
                    help="for rocprofv3 --pmc runs: a 512 MiB copy kernel of known traffic, then 3 force launches "
                         "per precision form on the workload; prints the launch sizes")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; runs on more than one GPU are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print(json.dumps({"error": "--steps must be >= 1 and --warmup >= 0"}))
        sys.exit(2)

    cfg = params.load_config(args.config)
    t0 = time.perf_counter()
    setup, p, l = params.build_inputs(cfg, args.cells)
    log(f"generated {p.n} particles, {l.n} links in {time.perf_counter() - t0:.1f}s")

    names = {"fp64": ["fp64"], "mixed": ["mixed", "mixed+learned"],
             "both": ["fp64", "mixed", "mixed+learned"]}[args.precision]
    if args.probe_traffic:
        probe_traffic(setup, p, l, sorted({FORMS[n][0] for n in names}))
        return
    model, model_path, model_report = sched_model(args.config)
    results = {}
    for name in names:
        prec, policy, learned = FORMS[name]
        log(f"running {name}: autotune, warmup {args.warmup}, timed {args.steps}")
        r = run_gpu(setup, p, l, prec, policy, model if learned else None, args.steps, args.warmup)
        r["value"] = p.n * args.steps / r["wall"]
        sms, scnt = r["kernels"]["step"]
        executed = r["sched"]["steps_executed"] if policy == A.HDM_REBUILD_LEARNED else scnt
        if sms <= 0 or executed <= 0:
            raise RuntimeError(f"{name}: no timed step launches recorded ({sms} ms over {executed} steps)")
        r["step_avg_ms"] = sms / executed     # skipped (guarded) launches cost ~µs and are not steps
        r["step_bytes"] = step_bytes(p.n, r["n_neigh"], l.n, prec)
        r["policy"] = policy
        results[name] = r
        log(f"  {name}: {r['value']:.4e} p-steps/s, {1e3 * r['wall'] / args.steps:.3f} ms/step, "
            f"step kernel {r['step_avg_ms']:.3f} ms ({r['step_bytes'] / r['step_avg_ms'] / 1e6:.0f} GB/s), "
            f"shape {r['exec']['name']} (tuned in {r['tune_steps']} steps), rebuilds {r['stats'].rebuilds}, "
            f"broken {r['broken']}, forced rebuild {1e3 * r['rebuild_s']:.1f} ms, sched {r['sched']}")

    # the practical HBM ceiling on this box: the streaming copy probe
    # (k_probe_copy over 2 x 2 GiB, best of the timed passes), outside any
    # timed window
    with A.Context(A.HDM_PREC_MIXED) as ctx:
        copy_gbs = max(ctx.bw_probe(A.HDM_PROBE_COPY, 4 << 30, 3) for _ in range(3))
    log(f"copy probe: {copy_gbs:.0f} GB/s")

    errors = {}
    if "fp64" in results:
        for name in ("mixed", "mixed+learned"):
            if name in results:
                errors[name] = error_vs_fp64(results["fp64"]["state"], results[name]["state"])
                log(f"  {name} vs fp64: {errors[name]}")
    for r in results.values():
        del r["state"]

    head = next(n for n in ("mixed+learned", "mixed", "fp64") if n in results)
    h = results[head]
    achieved = h["step_bytes"] / (h["step_avg_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"step_{'fp64' if head == 'fp64' else 'mixed'}_{args.config}_traffic.json")
    if os.path.exists(tf) and args.cells is None:
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")

    def amortised(r):
        ev = (model_report or {}).get("evaluation", {}).get(POLICY_EVAL[r["policy"]])
        if not ev or args.cells is not None:
            return None
        rate = ev["rebuilds"] / model_report["eval_steps"]
        ms = 1e3 * r["wall"] / args.steps + 1e3 * r["rebuild_s"] * rate
        return {"ms_per_step": ms, "value": p.n / (ms * 1e-3), "rebuild_ms": 1e3 * r["rebuild_s"],
                "rebuilds_per_step": rate,
                "rate_source": f"{model_report['eval_steps']}-step evaluation of the "
                               f"{POLICY_EVAL[r['policy']]} policy, steps 5k-15k ({model_path[:-4]}.json)"}

    cpu = None
    if args.cpu_steps > 0:
        c = cpu_baseline(setup, p, l, args.cpu_steps, args.cpu_ieee_steps, args.cpu_block_steps)
        s = c["slab_fast_math"]
        cpu = {"value": s["value"], "unit": "particle-steps/s", "cores": s["threads"], "kind": "port",
               "sample": f"{s['steps']} timesteps of the full {p.n}-particle workload (after 1 setup step), "
                         f"C/OpenMP oracle built -O3 -ffast-math (the reference's release flags), "
                         f"{s['seconds']:.1f} s on {s['threads']} threads of {c['cpu_model']}",
               "runs": c}

    kern = {}
    for name, r in results.items():
        kern[name] = {k: {"ms_total": v[0], "launches": v[1]} for k, v in r["kernels"].items() if v[1]}
    out = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * h["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: FCC glacier slab, {p.n} particles, {l.n} links, "
                        + {"fp64": "plain FP64 method, reference rebuild heuristic",
                           "mixed": "FP32 force / FP64 integration, exact skin test checked every step",
                           "mixed+learned": "FP32 force / FP64 integration, learned rebuild scheduler "
                                            "with the exact skin test on the device"}[head],
            "particles": p.n, "links": l.n, "form": head,
            "lattice": "a = 1.0 m, nearest-neighbour distance 0.70711 m, particle diameter 0.66711 m: "
                       "a 4 cm gap, so the intact slab starts free of contact pre-stress and its "
                       "neighbour pairs are all beyond contact range in the timed window",
            "steps_before_window": int(h["tune_steps"] + args.warmup),
            "rebuilds_in_timed_steps": int(h["stats"].rebuilds),
            "scheduler_model": model_path if head == "mixed+learned" else None,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_over_algorithmic": ratio(traffic, h["step_bytes"]) if traffic else None,
                     "traffic_GBps": traffic / (h["step_avg_ms"] * 1e-3) / 1e9 if traffic else None,
                     "copy_probe_GBps": copy_gbs,
                     "traffic_over_copy_probe": (traffic / (h["step_avg_ms"] * 1e-3) / 1e9 / copy_gbs
                                                 if traffic else None),
                     "kernel": "k_step (link break + force gather + integration)",
                     "bytes_per_launch": h["step_bytes"], "avg_launch_ms": h["step_avg_ms"]},
        "error_vs_fp64": errors.get(head) if head != "fp64" else None,
        "amortised_with_rebuilds": amortised(h),
        "cpu_baseline": cpu,
        "forms": {name: {"value": r["value"], "ms_per_step": 1e3 * r["wall"] / args.steps,
                         "step_avg_ms": r["step_avg_ms"],
                         "step_GBps": r["step_bytes"] / (r["step_avg_ms"] * 1e-3) / 1e9,
                         "rebuilds": int(r["stats"].rebuilds), "broken_links": r["broken"],
                         "device_bytes": r["dev_bytes"], "kernels": kern[name], "scheduler": r["sched"],
                         "launch_shape": r["exec"], "tune_steps": r["tune_steps"],
                         "forced_rebuild_ms": 1e3 * r["rebuild_s"],
                         "error_vs_fp64": errors.get(name), "amortised_with_rebuilds": amortised(r)}
                  for name, r in results.items()},
    }
    for name in ("mixed", "mixed+learned"):
        if "fp64" in results and name in results:
            out[f"{name}_over_fp64"] = ratio(results[name]["value"], results["fp64"]["value"])
            a, b = out["forms"][name]["amortised_with_rebuilds"], out["forms"]["fp64"]["amortised_with_rebuilds"]
            if a and b:
                out[f"{name}_over_fp64_amortised"] = ratio(a["value"], b["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
