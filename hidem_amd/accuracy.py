#!/usr/bin/env python3
"""Error of the mixed-precision form against the plain FP64 form.

Runs both forms side by side on one GPU from the same seeded inputs (FP32
contact and beam arithmetic and CT storage; FP64 positions, integration and
link strain) and reports, at each checkpoint step:

  position   max |x_mixed - x_fp64| over particles (Euclidean, metres), that
             maximum relative to the largest coordinate magnitude, and the
             largest per-particle error relative to that particle's
             displacement since step 0 (floored at 1 mm)
  force      ||F_mixed - F_fp64||_2 / ||F_fp64||_2 over all particles
             (translational components, forces of the checkpoint step)
  broken     broken-link counts of both forms and their relative difference

With --perturb, a third run (FP64, every coordinate moved by one ulp towards
+inf at step 0) is compared with the FP64 run the same way: the divergence a
one-ulp difference grows into, i.e. the floor below which no change of
arithmetic can stay once the dynamics amplify differences (fracture).  With
--perturb-mixed, a mixed run from the same perturbed coordinates is compared
with the mixed run: the spread of the mixed form itself, which tells a
systematic difference between the forms from the chaos of one realisation.
With --ensemble K, K more runs of each form start from coordinates moved by
2..K+1 ulps, and each checkpoint reports the broken-link counts of every
realisation of both forms (mean and spread per form).

Usage: python -m hidem_amd.accuracy [--config glacier_8m] [--checkpoints 100 1000 10000]
Prints one JSON line per checkpoint and a final summary line.
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


def _xyz(p):
    return np.stack([p["xCord"], p["yCord"], p["zCord"]], 1)


def _force(p):
    return np.stack([p["xForce"], p["yForce"], p["zForce"]], 1)


def compare(p64, pmx, x0, broken64, brokenmx):
    x64, xmx = _xyz(p64), _xyz(pmx)
    err = np.linalg.norm(xmx - x64, axis=1)
    scale = float(np.abs(x64).max())
    disp = np.maximum(np.linalg.norm(x64 - x0, axis=1), 1e-3)
    f64, fmx = _force(p64), _force(pmx)
    fnorm = float(np.linalg.norm(f64))
    return {
        "pos_max_abs_m": float(err.max()),
        "pos_max_rel": float(err.max() / scale),
        "pos_max_rel_to_displacement": float((err / disp).max()),
        "pos_rms_m": float(np.sqrt(np.mean(err ** 2))),
        "force_rel_l2": float(np.linalg.norm(fmx - f64) / fnorm) if fnorm > 0 else 0.0,
        "broken_fp64": int(broken64),
        "broken_mixed": int(brokenmx),
        "broken_rel_diff": float(abs(brokenmx - broken64) / broken64) if broken64 else float(brokenmx != 0),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", default="glacier_8m")
    ap.add_argument("--cells", type=int, nargs=3, default=None, help="override unit cells (smaller runs)")
    ap.add_argument("--checkpoints", type=int, nargs="+", default=[100, 1000, 10000])
    ap.add_argument("--chunk", type=int, default=500, help="steps per hdm_run call between progress lines")
    ap.add_argument("--perturb", action="store_true", help="also run FP64 from one-ulp perturbed coordinates")
    ap.add_argument("--perturb-mixed", action="store_true",
                    help="also run the mixed form from one-ulp perturbed coordinates")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="K more runs of each form from coordinates moved by 2..K+1 ulps (broken counts only)")
    args = ap.parse_args(argv)

    cfg = params.load_config(args.config)
    setup, p, l = params.build_inputs(cfg, args.cells)
    x0 = _xyz(p)
    print(f"# {args.config}: {p.n} particles, {l.n} links", file=sys.stderr, flush=True)
    out = []
    t0 = time.perf_counter()
    ctxs = [A.Context(A.HDM_PREC_FP64), A.Context(A.HDM_PREC_MIXED)]
    role = {}
    if args.perturb:
        role["fp64_ulp_perturbed"] = len(ctxs)
        ctxs.append(A.Context(A.HDM_PREC_FP64))
    if args.perturb_mixed:
        role["mixed_ulp_perturbed"] = len(ctxs)
        ctxs.append(A.Context(A.HDM_PREC_MIXED))
    def moved(ulps):
        q = p.copy()
        for f in ("xCord", "yCord", "zCord"):
            v = q[f]
            for _ in range(ulps):
                v = np.nextafter(v, np.inf)
            q.arrays[f] = v
        return q

    start = [p, p] + [moved(1)] * len(role)
    ens = []  # (precision name, context index)
    for k in range(args.ensemble):
        for prec, name in ((A.HDM_PREC_FP64, "fp64"), (A.HDM_PREC_MIXED, "mixed")):
            ens.append((name, len(ctxs)))
            ctxs.append(A.Context(prec))
            start.append(moved(k + 2))
    try:
        for i, c in enumerate(ctxs):
            c.upload(start[i], l)
            c.links_build()
        done = 0
        rebuilds = [0] * len(ctxs)
        for cp in sorted(args.checkpoints):
            while done < cp:
                k = min(args.chunk, cp - done)
                for i, c in enumerate(ctxs):
                    st = c.run(setup.run, k)
                    rebuilds[i] += st.rebuilds
                done += k
                print(f"# step {done}  ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
            p64 = ctxs[0].download_particles()
            b64 = ctxs[0].count_broken()
            pmx = ctxs[1].download_particles()
            bmx = ctxs[1].count_broken()
            r = compare(p64, pmx, x0, b64, bmx)
            r.update(step=cp, rebuilds_fp64=rebuilds[0], rebuilds_mixed=rebuilds[1])
            for name, i in role.items():
                base_p, base_b = (p64, b64) if name.startswith("fp64") else (pmx, bmx)
                r[name] = compare(base_p, ctxs[i].download_particles(), x0, base_b, ctxs[i].count_broken())
            if ens:
                counts = {"fp64": [b64], "mixed": [bmx]}
                for name, i in role.items():
                    counts["fp64" if name.startswith("fp64") else "mixed"].append(ctxs[i].count_broken())
                for name, i in ens:
                    counts[name].append(ctxs[i].count_broken())
                r["ensemble_broken"] = {k: {"counts": v, "mean": float(np.mean(v)), "std": float(np.std(v))}
                                        for k, v in counts.items()}
            out.append(r)
            print(json.dumps(r), flush=True)
    finally:
        for c in ctxs:
            c.close()
    summary = {"config": args.config, "particles": p.n, "links": l.n, "checkpoints": out,
               "seconds": time.perf_counter() - t0}
    print(json.dumps(summary), flush=True)
    return summary


if __name__ == "__main__":
    main()
