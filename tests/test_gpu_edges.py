"""Edge cases of the GPU path through the C ABI, each against the CPU oracle:

- a link naming a particle out of range is rejected (HiDEMLib's "Dangling
  links" error) and leaves the context usable;
- no links at all (contacts only);
- a duplicated link (two sides of one particle to the same partner), which
  also makes the FP32 form keep link partners in its neighbour list;
- an isolated particle (no neighbours, no links) beside a block.
"""
import numpy as np
import pytest

from hidem_amd import _abi as A
from hidem_amd import params

pytestmark = pytest.mark.gpu

XYZ = ("xCord", "yCord", "zCord")


def say(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.fixture(scope="module")
def squeezed_small():
    """A 6x5x4-cell block squeezed to NN distance 0.97 m: contacts active."""
    cfg = params.load_config("iceblock_16k")
    cfg["lattice"]["fccSpacing"] = 0.97 * 2 ** 0.5
    cfg["lattice"]["linkGenScale"] = 1.05 * 0.97
    cfg["hdm"]["cells"] = [6, 5, 4]
    return params.build_inputs(cfg)


def cat_links(l, extra_idx):
    """l with the links at extra_idx appended again (CT zero)."""
    def cat(v):
        return None if v is None else np.concatenate([v, v[extra_idx]])
    n = l.n + len(extra_idx)
    return A.Links(n, cat(l.a), cat(l.b), cat(l.state), cat(l.initialLen), cat(l.maxLoad),
                   [np.zeros(2 * n) for _ in range(6)])


def run_both(s, p, l, steps, orc):
    sim = orc.OracleSim(p, l, s)
    sim.run(steps)
    out = {}
    for prec in (A.HDM_PREC_FP64, A.HDM_PREC_MIXED):
        with A.Context(prec) as c:
            c.upload(p, l)
            c.links_build()
            c.run(s.run, steps)
            out[prec] = (c.download_particles(), c.download_links(), c.exec_info())
    return sim, out


def check(sim, out, p0, rel=1e-3):
    gp, gl, _ = out[A.HDM_PREC_FP64]
    for f in XYZ + ("xTheta", "yTheta", "zTheta"):
        assert bits_equal(gp[f], sim.p[f]), f
    assert np.array_equal(gl.state, sim.l.state)
    mp, ml, _ = out[A.HDM_PREC_MIXED]
    ref = np.stack([sim.p[f] for f in XYZ])
    disp = np.max(np.abs(ref - np.stack([p0[f] for f in XYZ])))
    err = np.max(np.abs(np.stack([mp[f] for f in XYZ]) - ref))
    say("  mixed max |dx|", err, "displacement", disp)
    assert np.isfinite(err) and err <= rel * disp


def test_dangling_link_is_rejected(squeezed_small):
    s, p, l = squeezed_small
    bad = cat_links(l, [0])
    bad.b[-1] = p.n  # one past the last particle
    say("dangling link")
    with A.Context(A.HDM_PREC_FP64) as c:
        c.upload(p, bad)
        with pytest.raises(A.HdmError, match="Dangling"):
            c.links_build()
        c.upload(p, l)  # the context stays usable
        c.links_build()
        st = c.run(s.run, 3)
        assert st.timestep == 3


def test_no_links(squeezed_small, orc):
    s, p, _ = squeezed_small
    say("no links")
    sim, out = run_both(s, p, A.Links.zeros(0), 20, orc)
    check(sim, out, p)
    assert out[A.HDM_PREC_MIXED][2]["merged_contacts"]  # nothing to leave out, list unchanged


def test_duplicated_link(squeezed_small, orc):
    s, p, l = squeezed_small
    say("duplicated link")
    sim, out = run_both(s, p, cat_links(l, [0, 5]), 5, orc)
    check(sim, out, p, rel=1e-2)
    # two sides of one particle to one partner: the FP32 form keeps the
    # reference's list (the link loop would apply that pair's contact twice)
    assert not out[A.HDM_PREC_MIXED][2]["merged_contacts"]


def test_isolated_particle(orc):
    cfg = params.load_config("iceblock_16k")
    cfg["hdm"]["cells"] = [6, 5, 4]
    s, p, l = params.build_inputs(cfg)
    q = A.Particles.zeros(p.n + 1)
    for f in A.PARTICLE_FIELDS:
        q.arrays[f][:p.n] = p[f]
    q.type[:p.n] = p.type
    q.arrays["radius"][p.n] = p["radius"][0]
    q.arrays["xCord"][p.n] = float(np.max(p["xCord"])) + 20.0  # far beyond the search distance
    q.arrays["yCord"][p.n] = float(np.mean(p["yCord"]))
    q.arrays["zCord"][p.n] = float(np.mean(p["zCord"]))
    say("isolated particle", q.n)
    sim, out = run_both(s, q, l, 30, orc)
    check(sim, out, q)
    assert sim.p["zCord"][p.n] != q["zCord"][p.n]  # it moved (gravity / buoyancy)
