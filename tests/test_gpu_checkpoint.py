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


GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden")


def test_driver_file_lattice(tmp_path):
    """[lattice] type = "File" in the TOML driver: the reference's 160k
    benchmark particle file read through libhidem_io.so, links generated,
    20 steps -- bit-identical to the same inputs assembled from Python."""
    import os
    path = os.path.join(GOLDEN, "benchmark_160k_particles.h5")
    cfg = tmp_path / "file_lattice.toml"
    cfg.write_text(f"""
[general]
scale = 12.0
E = 2e9
dt = 0.002
waterLevel = 300.0
maxLoad = 5e-4
bendBreak = 0.01
damping = 10000.0
dampingRotation = 10000.0
breakTime = 0.0
[iteration]
timeSteps = 1000
[grid]
searchDistanceScale = 1.4
[lattice]
type = "File"
path = "{path}"
particleRadius = 0.5
linkGenScale = 1.2
[hdm]
integrator = "iceblock"
rebuildPolicy = "exact"
maxLoadSpread = 0.1
linkSeed = 7
""")
    print("driver file lattice", flush=True)
    with A.Sim(str(cfg), [], precision=A.HDM_PREC_FP64) as sim:
        rp = sim.run_params()
        sim.steps(20)
        q = sim.ctx.download_particles()
        n, L = sim.n, sim.L
    p = h5io.read_particle_file(path)
    l = A.gen_links(p, 1.2, 5e-4, 0.1, 7)
    assert (p.n, l.n) == (n, L) and l.n > 5 * p.n
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.run(rp, 20)
        g = ctx.download_particles()
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(q[f], g[f]), f
    assert not bits_equal(q["zCord"], p["zCord"])  # the block moved


def test_driver_restart(tmp_path):
    """[restart] fromRestart in the TOML driver continues a checkpoint at its
    step, bit-identical to h5io.load + hdm_run."""
    import os
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    path = str(tmp_path / "checkpoint_30.h5")
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.run(s.run, 30)
        h5io.save(ctx, path, 30)
    cfg = os.path.join(params.CONFIG_DIR, "iceblock_16k.toml")
    print("driver restart", flush=True)
    with A.Sim(cfg, [f"--restart.fromRestart={path}"], precision=A.HDM_PREC_FP64) as sim:
        st = sim.steps(40)
        q = sim.ctx.download_particles()
    assert st.timestep == 70
    with A.Context(A.HDM_PREC_FP64) as ctx:
        assert h5io.load(ctx, path) == 30
        ctx.run(s.run, 40)
        g = ctx.download_particles()
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(q[f], g[f]), f


def test_driver_writes_checkpoint(tmp_path):
    """hdm_sim_checkpoint after 25 driver steps holds the same state as a
    checkpoint written from Python after the same 25 steps."""
    import os
    cfg = os.path.join(params.CONFIG_DIR, "iceblock_16k.toml")
    a, b = str(tmp_path / "driver.h5"), str(tmp_path / "python.h5")
    print("driver checkpoint", flush=True)
    with A.Sim(cfg, [], precision=A.HDM_PREC_FP64) as sim:
        sim.steps(25)
        sim.checkpoint(a)
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.run(s.run, 25)
        h5io.save(ctx, b, 25)
    ta, pa, la = h5io.read_checkpoint(a, with_max_load=True)
    tb, pb, lb = h5io.read_checkpoint(b, with_max_load=True)
    assert ta == tb == 25
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff", "radius"):
        assert bits_equal(pa[f], pb[f]), f
    assert np.array_equal(la.a, lb.a) and np.array_equal(la.state, lb.state)
    assert bits_equal(la.maxLoad, lb.maxLoad)
    for c in range(6):
        assert bits_equal(la.ct[c], lb.ct[c]), c


def test_driver_periodic_and_end_checkpoints(tmp_path):
    """[restart] checkpointInterval / timestepCheckpoint / writeEndCheckpoint
    in the driver (HiDEMLib.h:476-484, 536-541): files at steps 20, 40 and the
    end (50) under outWorkingPath/checkpoints, each holding the state of that
    step."""
    import os
    cfg = os.path.join(params.CONFIG_DIR, "iceblock_16k.toml")
    over = ["--restart.checkpointInterval=20", "--restart.timestepCheckpoint=true",
            "--restart.writeEndCheckpoint=true", f"--output.outWorkingPath={tmp_path}", "--iteration.timeSteps=50"]
    print("driver periodic checkpoints", flush=True)
    with A.Sim(cfg, over, precision=A.HDM_PREC_FP64) as sim:
        st = sim.steps(80)
    assert st.timestep == 50 and st.steps == 50
    files = sorted(os.listdir(tmp_path / "checkpoints"))
    assert files == ["checkpoint_20.h5", "checkpoint_40.h5", "checkpoint_50.h5"], files
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"))
    with A.Context(A.HDM_PREC_FP64) as ctx:
        ctx.upload(p, l)
        ctx.links_build()
        ctx.run(s.run, 20)
        ctx.run(s.run, 20)
        g = ctx.download_particles()
    ts, q, _ = h5io.read_checkpoint(str(tmp_path / "checkpoints" / "checkpoint_40.h5"))
    assert ts == 40
    for f in ("xCord", "yCord", "zCord", "xTheta", "zDiff"):
        assert bits_equal(q[f], g[f]), f
