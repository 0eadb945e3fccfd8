"""Edge cases of the GPU path through the C ABI, each against the CPU oracle:

- a link naming a particle out of range is rejected (HiDEMLib's "Dangling
  links" error) and leaves the context usable;
- no links at all (contacts only);
- a duplicated link (two sides of one particle to the same partner), which
  also makes the FP32 form keep link partners in its neighbour list;
- an isolated particle (no neighbours, no links) beside a block;
- a pair joined by a single link beside a block (one link side per end);
- a particle with more than 64 links (wider link slots than the FP32 form's
  merged list allows).
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


def add_links(l, a, b, length):
    """l with links a[i]-b[i] of rest length length[i] appended (CT zero)."""
    m = len(a)
    n = l.n + m
    return A.Links(n, np.concatenate([l.a, np.asarray(a, dtype=l.a.dtype)]),
                   np.concatenate([l.b, np.asarray(b, dtype=l.b.dtype)]),
                   np.concatenate([l.state, np.ones(m, dtype=l.state.dtype)]),
                   np.concatenate([l.initialLen, np.asarray(length, dtype=l.initialLen.dtype)]),
                   None if l.maxLoad is None else np.concatenate([l.maxLoad, np.full(m, l.maxLoad[0])]),
                   [np.zeros(2 * n) for _ in range(6)])


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


def test_single_link_pair(squeezed_small, orc):
    s, p, l = squeezed_small
    q = A.Particles.zeros(p.n + 2)
    for f in A.PARTICLE_FIELDS:
        q.arrays[f][:p.n] = p[f]
    q.type[:p.n] = p.type
    x = float(np.max(p["xCord"])) + 20.0  # far beyond the search distance of the block
    for i, y in enumerate((float(np.mean(p["yCord"])), float(np.mean(p["yCord"])) + 0.97)):
        q.arrays["radius"][p.n + i] = p["radius"][0]
        q.arrays["xCord"][p.n + i] = x
        q.arrays["yCord"][p.n + i] = y
        q.arrays["zCord"][p.n + i] = float(np.mean(p["zCord"]))
    ql = add_links(l, [p.n], [p.n + 1], [0.97])
    say("single link pair", q.n, ql.n)
    sim, out = run_both(s, q, ql, 30, orc)
    check(sim, out, q)
    assert out[A.HDM_PREC_MIXED][2]["merged_contacts"]


def test_particle_with_more_than_64_links(squeezed_small, orc):
    s, p, l = squeezed_small
    linked = set(l.b[l.a == 0].tolist()) | set(l.a[l.b == 0].tolist())
    partners = [j for j in range(1, p.n) if j not in linked][:70]
    xyz = np.stack([p[f] for f in XYZ])
    length = np.linalg.norm(xyz[:, partners] - xyz[:, [0]], axis=0)  # at rest
    ql = add_links(l, [0] * len(partners), partners, length)
    say("particle with", len(partners) + len(linked), "links")
    sim, out = run_both(s, p, ql, 10, orc)
    check(sim, out, p, rel=1e-2)
    assert not out[A.HDM_PREC_MIXED][2]["merged_contacts"]  # 64 sides at most in the merged list
