#!/usr/bin/env python3
"""Batched form: many independent ice blocks advanced together on one GPU.

256 blocks (default) with particle counts log-uniform in [5k, 200k], each
with its own seeded fracture-threshold mean (uniform in [2e-4, 5e-4]) and
waterline (uniform in [0.2, 0.8] of its height), laid out side by side in one
state (the 16k block's material and integrator, a clamp at each block's left
end).  Every step is one launch of the fused step kernel over all blocks
(FP32 force arithmetic, FP64 integration and strain), replayed from a
captured hipGraph in groups of 8 (or 2), with the launch shape picked by
the autotuner; the exact rebuild test runs on the device
(HDM_REBUILD_LEARNED without a network: fixed batches of host checks).  A
third leg runs each batch as one co-resident launch that loops over its
steps with a grid barrier between them (hdm_exec_params.persistent).

Usage: python -m hidem_amd.batched [--blocks 256] [--steps 5000]
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
    from hidem_amd.roofline import HBM_PEAK_GBS, reference_list_entries, step_bytes  # noqa: E402
else:
    from . import _abi as A, params
    from .roofline import HBM_PEAK_GBS, reference_list_entries, step_bytes


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(b, setup, steps, graphs, autotune, variant=-1, chunk=500, prec=A.HDM_PREC_MIXED, persistent=False):
    rp = A.HdmRunParams.from_buffer_copy(setup.run)
    rp.policy = A.HDM_REBUILD_LEARNED
    with A.Context(prec) as c:
        c.upload(b.particles, b.links)
        c.links_build()
        A.set_blocks(c, b.block_id, b.water)
        c.sched_configure(block_cells=4, min_batch=8, max_batch=256, safety=0.5)
        c.set_exec(autotune=autotune, graphs=graphs, variant=variant)
        c.run(rp, 20)                                   # setup step, first rebuild, autotuning
        if persistent:
            c.set_exec(persistent=True, variant=c.exec_info()["variant"])
        c.sync()
        s0 = c.sched_stats()
        c.set_timing(True)
        c.reset_kernel_times()
        t0 = time.perf_counter()
        rebuilds, done = 0, 0
        while done < steps:
            k = min(chunk, steps - done)
            st = c.run(rp, k)
            rebuilds += st.rebuilds
            done += k
            log(f"#   step {done}/{steps}: {time.perf_counter() - t0:.1f} s, rebuilds {rebuilds}")
        c.sync()
        wall = time.perf_counter() - t0
        kt = c.kernel_times()
        c.set_timing(False)
        info = c.exec_info()
        stats = c.sched_stats()
        executed = stats["steps_executed"] - s0["steps_executed"]
        counts, _, _ = c.neighbours_download()
        n_neigh = reference_list_entries(int(counts.sum()), b.links.n, info)
        ql = c.download_links(with_max_load=False)
    broken = ql.state == 0
    per_block = np.bincount(b.block_id[b.links.a[broken]], minlength=len(b.water))
    # roofline of the step launches (k_step, or k_step_loop per batch): the
    # compulsory bytes of a step (list entries as at the end of the run) over
    # the mean step time from the step timer's HIP events
    sms, _ = kt["step"]
    avg = sms / executed if executed > 0 else float("nan")
    nbytes = step_bytes(b.particles.n, n_neigh, b.links.n, prec)
    achieved = nbytes / (avg * 1e-3) / 1e9
    return {"steps": steps, "seconds": wall, "ms_per_step": 1e3 * wall / steps,
            "particle_steps_per_s": b.particles.n * steps / wall, "rebuilds": rebuilds,
            "broken_links": int(broken.sum()), "blocks_fractured": int((per_block > 0).sum()),
            "exec": info, "scheduler": stats,
            "roofline": {"bound": "hbm", "kernel": "k_step_loop" if persistent else "k_step",
                         "avg_step_ms": avg, "bytes_per_step": nbytes, "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS}}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--blocks", type=int, default=256)
    ap.add_argument("--min-particles", type=int, default=5000)
    ap.add_argument("--max-particles", type=int, default=200000)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--baseline-steps", type=int, default=None,
                    help="steps of the same run without graphs or autotuning (default: --steps; 0: skip)")
    ap.add_argument("--persistent-steps", type=int, default=None,
                    help="steps of the one-launch-per-batch leg (default: --steps; 0: skip)")
    ap.add_argument("--seed", type=int, default=11)
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    b = A.gen_batch(n_blocks=args.blocks, min_particles=args.min_particles, max_particles=args.max_particles,
                    seed=args.seed)
    setup = params.derive(params.load_config("iceblock_16k"))
    counts = np.bincount(b.block_id, minlength=args.blocks)
    log(f"# {args.blocks} blocks, {b.particles.n} particles ({counts.min()}..{counts.max()} per block), "
        f"{b.links.n} links, generated in {time.perf_counter() - t0:.1f} s")
    report = {"blocks": args.blocks, "particles": int(b.particles.n), "links": int(b.links.n),
              "particles_per_block": [int(counts.min()), int(counts.max())]}
    if args.baseline_steps is None:
        args.baseline_steps = args.steps
    if args.baseline_steps > 0:
        log("# baseline: per-step launches, default shape")
        report["baseline"] = run(b, setup, args.baseline_steps, graphs=False, autotune=False, variant=0)
    log("# graphs + autotuned shape")
    report["batched"] = run(b, setup, args.steps, graphs=True, autotune=True)
    log(f"#   {report['batched']['ms_per_step']:.3f} ms/step, step {report['batched']['roofline']['avg_step_ms']:.3f} ms "
        f"({report['batched']['roofline']['achieved']:.0f} GB/s)")
    if args.persistent_steps is None:
        args.persistent_steps = args.steps
    if args.persistent_steps > 0:
        log("# one co-resident launch per batch")
        report["persistent"] = run(b, setup, args.persistent_steps, graphs=False, autotune=True, persistent=True)
        r = report["persistent"]
        log(f"#   {r['ms_per_step']:.3f} ms/step, step {r['roofline']['avg_step_ms']:.3f} ms "
            f"({r['roofline']['achieved']:.0f} GB/s), {r['exec']['persistent_launches']} launches")
    report["seconds"] = time.perf_counter() - t0
    print(json.dumps(report), flush=True)
    return report


if __name__ == "__main__":
    main()
