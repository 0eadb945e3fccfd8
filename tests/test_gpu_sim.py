"""The TOML-driven C++ driver (HiDEMLibHip through hdm_sim_*) on the GPU.

The reference-shaped per-call loop (gridService, computePartForce,
applyForce, linkUpdate, finalizeTs: one kernel each) and the fused loop
(hdm_run) reach bit-identical states from the same TOML, and both match an
hdm_run driven from Python with the same derived inputs.
"""
import os

import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params

pytestmark = pytest.mark.gpu

CFG = os.path.join(params.CONFIG_DIR, "iceblock_16k.toml")


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("prec", [A.HDM_PREC_FP64, A.HDM_PREC_MIXED])
def test_sim_loops_match_run(prec):
    print("sim loops", prec, flush=True)
    runs = {}
    for name, over in (("per_call", ["--hdm.fusedSteps=false"]), ("fused", [])):
        with A.Sim(CFG, over, precision=prec) as s:
            st, log = s.steps(80, log_rebuilds=True)
            assert st.timestep == 80
            runs[name] = (log, s.ctx.download_particles(), s.ctx.download_links(with_max_load=False))
    setup, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    with A.Context(prec) as c:
        c.upload(p, l)
        c.links_build()
        st, log = c.run(setup.run, 80, log_rebuilds=True)
        runs["python"] = (log, c.download_particles(), c.download_links(with_max_load=False))
    ref_log, ref_p, ref_l = runs["python"]
    for name in ("per_call", "fused"):
        log, q, ql = runs[name]
        assert np.array_equal(log, ref_log), name
        for f in ("xCord", "yCord", "zCord", "xTheta", "yTheta", "zTheta", "xDiff", "wDiff"):
            assert bits_equal(q[f], ref_p[f]), (name, f)
        assert np.array_equal(ql.state, ref_l.state), name
