#!/usr/bin/env python3
"""Timestep throughput of the MI355X HiDEM hot path on the 8M-particle glacier slab.

A "step" is one full timestep of the hot path -- rebuild check (and rebuild
when due), contact + beam force gather, time integration, link break -- over
all 8,000,000 particles / ~48M links resident in HBM.  Prints one JSON line.

value      whole-job particle-steps/s of the headline form, mixed+learned
           (FP32 force arithmetic, FP64 integration and strain, learned
           rebuild scheduler with the exact skin test on the device):
           N_particles * K / wall time of the K timed steps, bracketed by
           device synchronisation.  The plain FP64 method (reference rebuild
           heuristic) and the mixed form with a host-checked exact test run
           beside it under "forms".
roofline   for the dominant kernel (k_step: link break of the previous step,
           force gather and integration in one launch): algorithmic bytes per
           launch (compulsory traffic of the arrays it must touch, see
           step_bytes) / its mean launch time from HIP events recorded on the
           context's stream during the timed steps, against 8 TB/s.
cpu_baseline  the CPU oracle (a C/OpenMP restatement of the reference's CPU
           path, built with the reference's release flags) timed on this host
           on a bounded sample of the same workload.
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

METRIC = "particle-steps/s at 8M particles (mixed+learned vs plain FP64); % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def step_bytes(n, n_neigh_total, n_links, prec, exact=True):
    """Compulsory HBM bytes of one fused step launch (k_step: link break of
    the previous step, contact + beam force gather, integration).

    Per particle: coordinates, rotations, Diff (12 x 8), radius (8), type (1),
    neighbour and link counts (2 x 4), and for the exact rebuild test the
    coordinates at the last rebuild (3 x 8) read; coordinates, rotations and
    Diff (12 x 8) and the force (6 x 8 FP64 / 6 x 4 mixed) written.  Per
    neighbour-list entry: its 4 B index.  Per link side: CT0..5 read and
    written (2 x 6 x 8 FP64 / 2 x 6 x 4 mixed), link word and other endpoint
    (2 x 4).  Per link: initialLen and maxLoad (2 x 8) for the break test.
    Partner state is gathered from particles already counted once (cache
    reuse); the rare writes of a breaking link are not counted."""
    fo = 8 if prec == A.HDM_PREC_FP64 else 4
    per_particle = 96 + 8 + 1 + 8 + (24 if exact else 0) + 96 + 6 * fo
    per_side = 12 * fo + 8
    return n * per_particle + 4 * n_neigh_total + 2 * n_links * per_side + 16 * n_links


def reference_list_entries(n_list, n_links, exec_info):
    """Entries of the reference's neighbour list: the mixed form's list leaves
    out link partners (their contacts run in the link loop), which the
    reference lists, two entries per link."""
    return n_list + (2 * n_links if exec_info["merged_contacts"] else 0)


def break_bytes(n_links):
    """Compulsory bytes of the standalone k_break: A, B, state, initialLen,
    maxLoad per link (endpoint coordinates come through the cache)."""
    return n_links * (4 + 4 + 1 + 8 + 8)


# Forms: (name, precision, rebuild policy, uses the trained scheduler network)
FORMS = {
    "fp64": (A.HDM_PREC_FP64, A.HDM_REBUILD_HEURISTIC, False),      # plain FP64 method, reference gridService
    "mixed": (A.HDM_PREC_MIXED, A.HDM_REBUILD_EXACT, False),        # precision form, exact test checked each step
    "mixed+learned": (A.HDM_PREC_MIXED, A.HDM_REBUILD_LEARNED, True),  # precision form + learned scheduler
}


def sched_model(config):
    path = os.path.join(ROOT, "hidem_amd", "models", f"sched_{config}.bin")
    if not os.path.exists(path):
        return None, None
    from hidem_amd import sched_train
    return sched_train.load_model(path), os.path.relpath(path, ROOT)


def run_gpu(setup, p, l, prec, policy, model, steps, warmup):
    rp = A.HdmRunParams.from_buffer_copy(setup.run)
    rp.policy = policy
    with A.Context(prec) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.sched_configure(block_cells=4, min_batch=8, max_batch=256, safety=0.5)
        if model is not None:
            ctx.sched_set_model(model)
        # every form autotunes the step's launch shape during warmup (all shapes give
        # identical results); the timed steps then use the form's own rebuild policy
        ctx.set_exec(autotune=True, graphs=policy == A.HDM_REBUILD_LEARNED)
        if warmup > 0:
            rw = A.HdmRunParams.from_buffer_copy(rp)
            rw.policy = A.HDM_REBUILD_LEARNED
            ctx.run(rw, warmup)
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
    sched = {k: s1[k] - s0[k] for k in s1}
    return dict(wall=t1 - t0, stats=st, kernels=kt, n_neigh=n_neigh, broken=broken, dev_bytes=dev_bytes,
                sched=sched, exec=exec_info)


def cpu_baseline(setup, p, l, steps, fast=True):
    from oracle import oracle as orc  # checker / baseline only
    t0 = time.perf_counter()
    sim = orc.OracleSim(p, l, setup, fast=fast)
    sim.run(1)  # builds grid, neighbour list and link lists, one step
    t1 = time.perf_counter()
    sim.run(steps)
    t2 = time.perf_counter()
    threads = orc.num_threads(fast)
    sim.close()
    return dict(value=p.n * steps / (t2 - t1), seconds=t2 - t1, setup_seconds=t1 - t0, threads=threads)


def probe_traffic(setup, p, l, forms):
    """Kernels of known size for calibrating the FETCH_SIZE / WRITE_SIZE
    counters, then fused step launches on the real workload."""
    for prec in forms:
        with A.Context(prec) as ctx:
            ctx.upload(p, l)
            ctx.links_build()
            # the same autotuned launch shape the timed bench steps use
            ctx.set_exec(autotune=True)
            rw = A.HdmRunParams.from_buffer_copy(setup.run)
            rw.policy = A.HDM_REBUILD_LEARNED
            ctx.run(rw, 30)
            if prec == forms[0]:
                g = ctx.bw_probe(A.HDM_PROBE_COPY, 1 << 30, 1)  # k_probe_copy: 512 MiB read + 512 MiB written
                log(f"copy probe 512 MiB read + 512 MiB written: {g:.0f} GB/s")
            ctx.run(setup.run, 3)
            log(f"launch shape {ctx.exec_info()['name']}")
            ctx.sync()
            counts, _, _ = ctx.neighbours_download()
            n_neigh = reference_list_entries(int(counts.sum()), l.n, ctx.exec_info())
            b = step_bytes(p.n, n_neigh, l.n, prec)
            print(json.dumps({"precision": "fp64" if prec == A.HDM_PREC_FP64 else "mixed",
                              "step_algorithmic_bytes": b, "neighbours": n_neigh, "links": l.n}),
                  flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--config", default="glacier_8m")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--cpu-steps", type=int, default=40, help="oracle steps in the CPU baseline sample (0: skip)")
    ap.add_argument("--cells", type=int, nargs=3, default=None, help="override unit cells (smaller runs)")
    ap.add_argument("--probe-traffic", action="store_true",
                    help="for rocprofv3 --pmc runs: a 512 MiB copy kernel of known traffic, then 3 force launches "
                         "per precision form on the workload; prints the launch sizes")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
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
    model, model_path = sched_model(args.config)
    results = {}
    for name in names:
        prec, policy, learned = FORMS[name]
        log(f"running {name}: warmup {args.warmup}, timed {args.steps}")
        r = run_gpu(setup, p, l, prec, policy, model if learned else None, args.steps, args.warmup)
        r["value"] = p.n * args.steps / r["wall"]
        sms, scnt = r["kernels"]["step"]
        executed = r["sched"]["steps_executed"] if policy == A.HDM_REBUILD_LEARNED else scnt
        r["step_avg_ms"] = sms / max(executed, 1)     # skipped (guarded) launches cost ~µs and are not steps
        r["step_bytes"] = step_bytes(p.n, r["n_neigh"], l.n, prec)
        results[name] = r
        log(f"  {name}: {r['value']:.4e} p-steps/s, {1e3 * r['wall'] / args.steps:.3f} ms/step, "
            f"step kernel {r['step_avg_ms']:.3f} ms ({r['step_bytes'] / r['step_avg_ms'] / 1e6:.0f} GB/s), "
            f"rebuilds {r['stats'].rebuilds}, broken {r['broken']}, sched {r['sched']}")

    head = next(n for n in ("mixed+learned", "mixed", "fp64") if n in results)
    h = results[head]
    achieved = h["step_bytes"] / (h["step_avg_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"step_{'fp64' if head == 'fp64' else 'mixed'}_{args.config}_traffic.json")
    if os.path.exists(tf) and args.cells is None:
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")

    cpu = None
    if args.cpu_steps > 0:
        log(f"cpu baseline: oracle, {args.cpu_steps} steps")
        c = cpu_baseline(setup, p, l, args.cpu_steps)
        cpu = {"value": c["value"], "unit": "particle-steps/s", "cores": c["threads"], "kind": "port",
               "sample": f"{args.cpu_steps} timesteps of the full {p.n}-particle workload (after 1 setup step), "
                         f"C/OpenMP oracle built -O3 -ffast-math, {c['seconds']:.1f} s"}

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
            "rebuilds_in_timed_steps": int(h["stats"].rebuilds),
            "scheduler_model": model_path if head == "mixed+learned" else None,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_step (link break + force gather + integration)",
                     "bytes_per_launch": h["step_bytes"], "avg_launch_ms": h["step_avg_ms"]},
        "cpu_baseline": cpu,
        "forms": {name: {"value": r["value"], "ms_per_step": 1e3 * r["wall"] / args.steps,
                         "step_avg_ms": r["step_avg_ms"],
                         "step_GBps": r["step_bytes"] / (r["step_avg_ms"] * 1e-3) / 1e9,
                         "rebuilds": int(r["stats"].rebuilds), "broken_links": r["broken"],
                         "device_bytes": r["dev_bytes"], "kernels": kern[name], "scheduler": r["sched"],
                         "launch_shape": r["exec"]}
                  for name, r in results.items()},
    }
    for name in ("mixed", "mixed+learned"):
        if "fp64" in results and name in results:
            out[f"{name}_over_fp64"] = results[name]["value"] / results["fp64"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
