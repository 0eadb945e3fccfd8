"""Restart from an HDF5 checkpoint on the GPU: 60 steps, checkpoint to a file
in the reference's layout, load it into a fresh context, 60 more steps --
bit-identical to the same restart through an in-memory download / upload
(FP64 and the FP32 form; CT and link states survive the file)."""
import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import h5io
from hidem_amd import params

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("prec", [A.HDM_PREC_FP64, A.HDM_PREC_MIXED])
def test_restart_from_checkpoint(prec, tmp_path):
    print("restart from checkpoint", prec, flush=True)
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"), [10, 8, 6])
    path = str(tmp_path / "checkpoint_60.h5")
    with A.Context(prec) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        st = ctx.run(s.run, 60)
        h5io.save(ctx, path, st.timestep)
        mem_p, mem_l = ctx.download_particles(), ctx.download_links()
    assert h5io.checkpoint_info(path) == (p.n, l.n, 60)
    out = []
    for source in ("file", "memory"):
        with A.Context(prec) as ctx:
            if source == "file":
                assert h5io.load(ctx, path) == 60
            else:
                ctx.upload(mem_p, mem_l)
                ctx.links_build()
                ctx.set_timestep(60)
            ctx.run(s.run, 60)
            out.append((ctx.download_particles(), ctx.download_links()))
    (fp, fl), (mp, ml) = out
    for f in ("xCord", "yCord", "zCord", "xTheta", "zTheta", "xDiff", "wDiff"):
        assert bits_equal(fp[f], mp[f]), f
    assert np.array_equal(fl.state, ml.state)
    for c in range(6):
        assert bits_equal(fl.ct[c], ml.ct[c]), c
    assert (fl.state == 0).sum() > 0 or (mem_l.state == 0).sum() == 0
