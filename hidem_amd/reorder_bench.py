#!/usr/bin/env python3
"""Effect of the redistribution order (hdm_reorder) on the fused step.

The 8M-particle glacier slab is generated, its particles shuffled into a
random order (links renamed), and the FP32-form step timed on that order,
then again after hdm_reorder has put the particles back into four-level cell
order -- the order the reference's redistribution produces.  The generator's
own Morton order is timed for comparison.  Prints one JSON line.

Usage: python -m hidem_amd.reorder_bench [--config glacier_8m] [--steps 50]
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def shuffle(p, l, seed):
    r = np.random.default_rng(seed)
    perm = r.permutation(p.n)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(p.n)
    for f in p.arrays:
        p.arrays[f][:] = p.arrays[f][perm]
    p.type[:] = p.type[perm]
    l.a[:] = inv[l.a].astype(np.int32)
    l.b[:] = inv[l.b].astype(np.int32)


def step_ms(ctx, rp, steps, warmup):
    ctx.run(rp, warmup)
    ctx.sync()
    ctx.set_timing(True)
    ctx.reset_kernel_times()
    ctx.run(rp, steps)
    ctx.sync()
    ms, launches = ctx.kernel_times()["step"]
    ctx.set_timing(False)
    return ms / max(launches, 1)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", default="glacier_8m")
    ap.add_argument("--cells", type=int, nargs=3, default=None)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    args = ap.parse_args(argv)
    setup, p, l = params.build_inputs(params.load_config(args.config), args.cells)
    log(f"# {p.n} particles, {l.n} links")
    rp = A.HdmRunParams.from_buffer_copy(setup.run)
    rp.policy = A.HDM_REBUILD_EXACT
    out = {"config": args.config, "particles": p.n, "links": l.n, "steps": args.steps}
    with A.Context(A.HDM_PREC_MIXED) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        out["morton_ms"] = step_ms(ctx, rp, args.steps, args.warmup)
        log(f"# generator (Morton) order: {out['morton_ms']:.3f} ms per step")
    shuffle(p, l, 1)
    with A.Context(A.HDM_PREC_MIXED) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        out["shuffled_ms"] = step_ms(ctx, rp, args.steps, args.warmup)
        log(f"# shuffled order: {out['shuffled_ms']:.3f} ms per step")
        ctx.sync()
        t0 = time.perf_counter()
        ctx.reorder(setup.search_distance)
        ctx.sync()
        out["reorder_seconds"] = time.perf_counter() - t0
        out["reordered_ms"] = step_ms(ctx, rp, args.steps, args.warmup)
        log(f"# after hdm_reorder ({out['reorder_seconds']:.3f} s): {out['reordered_ms']:.3f} ms per step")
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
