"""Pins the oracle's per-particle link lists (hdmo_link_lists, restating
LinkAccelSide::generateLinks / generatePerParticleIndexes,
PLiB/Accel/LinkAccelSide.h:43-143) and two loop-level closed forms of the
oracle's timestep (hdmo_sim_run):

* link lists: every link appears exactly once in each endpoint's list and
  nowhere else; each list ascends by link id (the reference's bubble sort,
  LinkAccelSide.h:123-137); `other` is the opposite end; the side index is
  2l for end A and 2l+1 for end B (LinkDataHandler.h:229-236 hands
  processInitLink sides outSideInd, outSideInd+1; HiDEMLinkData.h:58-67);
  the side indices over all lists are a permutation of 0..2L-1;
* the reference's FCC link cases (CatchTests/LinkGeneratorTester.cpp:84-129,
  202-256): a centre particle of an FCC lattice has 12 links, each to a
  nearest neighbour;
* a single link pulled at constant speed (weak material, no damping or drag)
  breaks at the first step whose stretch (t+1)*v*dt/L0 plus
  bendBreak * (t+1)*omega*dt reaches the threshold, and stays broken;
* Newton's third law over the whole gather: contact and beam forces of an
  isolated block sum to zero up to rounding.

Link sets here are built by brute force in numpy with the reference's
criterion dist <= (r_i + r_j) * linkGenScale (LinkGenerator.h:83-89), in a
shuffled order with A/B swapped at random, so that a list order or side
convention that only holds for sorted input would fail.
"""
import numpy as np
import pytest

from hidem_amd import params
from hidem_amd._abi import Links, Particles


def fcc_points(n, a):
    """FCC lattice points of n x n x n unit cells with spacing a (4 basis atoms)."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    cells = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3) * a
    return (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)


def particles_at(xyz, radius):
    p = Particles.zeros(len(xyz))
    for k, f in enumerate(("xCord", "yCord", "zCord")):
        p.arrays[f][:] = xyz[:, k]
    p.arrays["radius"][:] = radius
    return p


def brute_links(xyz, radius, scale, rng=None):
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    ii, jj = np.nonzero(np.triu(d <= 2 * radius * scale, 1))
    if rng is not None:
        perm = rng.permutation(len(ii))
        ii, jj = ii[perm], jj[perm]
        swap = rng.random(len(ii)) < 0.5
        ii, jj = np.where(swap, jj, ii), np.where(swap, ii, jj)
    l = Links.zeros(len(ii))
    l.a[:] = ii
    l.b[:] = jj
    l.initialLen[:] = d[ii, jj]
    return l


def check_lists(n, l, counts, starts, lid, oth, side):
    a, b = l.a.astype(np.int64), l.b.astype(np.int64)
    # counts = number of links touching each particle
    expect = np.bincount(a, minlength=n) + np.bincount(b, minlength=n)
    assert np.array_equal(counts, expect)
    assert np.array_equal(starts, np.concatenate([[0], np.cumsum(expect)]))
    owner = np.repeat(np.arange(n), counts)
    # each list strictly ascending by link id
    for p in range(n):
        seg = lid[starts[p]:starts[p + 1]]
        assert np.all(np.diff(seg) > 0), f"particle {p}: link ids not ascending {seg}"
    # every link once in each endpoint's list and nowhere else
    is_a = owner == a[lid]
    is_b = owner == b[lid]
    assert np.all(is_a | is_b), "a link listed at a particle that is not one of its ends"
    assert np.array_equal(np.bincount(lid[is_a], minlength=l.n), np.ones(l.n, np.int64))
    assert np.array_equal(np.bincount(lid[is_b], minlength=l.n), np.ones(l.n, np.int64))
    # other end and side: 2l at end A, 2l+1 at end B
    assert np.array_equal(oth, np.where(is_a, b[lid], a[lid]))
    assert np.array_equal(side, np.where(is_a, 2 * lid, 2 * lid + 1))
    # the side indices over all lists are a bijection onto 0..2L-1
    assert np.array_equal(np.sort(side), np.arange(2 * l.n))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_link_lists_invariants_shuffled(orc, seed):
    rng = np.random.default_rng(seed)
    xyz = fcc_points(4, 1.0) + rng.uniform(-0.02, 0.02, (256, 3))
    l = brute_links(xyz, 0.3535, 1.05, rng)
    assert l.n > 500
    counts, starts, lid, oth, side = orc.link_lists(len(xyz), l)
    check_lists(len(xyz), l, counts, starts, lid, oth, side)


def test_link_lists_sorted_input_and_isolated_particles(orc):
    # links sorted by (A, B) as the generator emits them; particles 0 and 5 unlinked
    l = Links.zeros(4)
    l.a[:] = [1, 1, 2, 3]
    l.b[:] = [2, 4, 3, 4]
    counts, starts, lid, oth, side = orc.link_lists(6, l)
    check_lists(6, l, counts, starts, lid, oth, side)
    assert counts[0] == 0 and counts[5] == 0
    # particle 4 is end B of links 1 and 3: sides 3 and 7, others 1 and 3
    s4 = slice(starts[4], starts[5])
    assert list(lid[s4]) == [1, 3] and list(oth[s4]) == [1, 3] and list(side[s4]) == [3, 7]


def test_link_lists_reject_dangling(orc):
    l = Links.zeros(1)
    l.a[0], l.b[0] = 0, 7
    with pytest.raises(RuntimeError):
        orc.link_lists(4, l)


@pytest.mark.parametrize("case", ["spacing1_centre111", "spacing05_centre222"])
def test_fcc_centre_has_12_links(orc, case):
    # LinkGeneratorTester.cpp:84-129: FCC spacing 1 on [0,2]^3, radius 0.25, linkGenScale 1.57,
    # particle (1,1,1) has 12 links; :202-256: spacing 0.5 on [0,4]^3, radius 0.125, scale 1.6,
    # particle (2,2,2) has 12 links, all within spacing/2 per axis
    a, n, radius, scale, centre = {"spacing1_centre111": (1.0, 2, 0.25, 1.57, 1.0),
                                   "spacing05_centre222": (0.5, 8, 0.125, 1.6, 2.0)}[case]
    xyz = fcc_points(n, a)
    l = brute_links(xyz, radius, scale, np.random.default_rng(3))
    counts, starts, lid, oth, side = orc.link_lists(len(xyz), l)
    check_lists(len(xyz), l, counts, starts, lid, oth, side)
    c = int(np.nonzero(np.all(xyz == centre, axis=1))[0][0])
    assert counts[c] == 12
    others = xyz[oth[starts[c]:starts[c + 1]]]
    assert np.all(np.abs(others - xyz[c]) <= a / 2)
    np.testing.assert_allclose(np.linalg.norm(others - xyz[c], axis=1), a / np.sqrt(2), rtol=1e-15)


# ---- loop-level closed forms ---------------------------------------------------------
def quiet_setup(E=1e-3, **general):
    """The 16k block's setup with no gravity, drag or damping and a weak
    material, so a particle given a Diff keeps it to rounding."""
    cfg = params.load_config("iceblock_16k")
    g = cfg["general"]
    g.update(E=E, g=0.0, airDrag=0.0, waterDrag=0.0, damping=0.0, dampingRotation=0.0, breakTime=-1.0)
    g.update(general)
    return params.derive(cfg)


@pytest.mark.parametrize("omega", [0.0, 3e-4])
def test_single_link_pulled_breaks_at_predicted_step(orc, omega):
    s = quiet_setup(bendBreak=0.01)
    L0, v_dt, thr, bend = 1.04, 1e-5, 3e-4, 0.01
    p = particles_at(np.array([[10.0, 10.0, 10.0], [10.0 + L0, 10.0, 10.0]]), 0.5)
    p.type[0] = 1                    # end A clamped
    p.arrays["xDiff"][1] = v_dt      # end B pulled along the link at constant speed
    p.arrays["vDiff"][1] = omega     # and turned about y at a constant rate
    l = Links.zeros(1)
    l.a[0], l.b[0] = 0, 1
    l.initialLen[0] = L0
    l.maxLoad[0] = thr
    # after step t the stretch is (t+1)*v_dt and the bend (t+1)*omega
    per_step = v_dt / L0 + bend * omega
    t_expect = int(np.ceil(thr / per_step)) - 1
    frac = thr / per_step - np.floor(thr / per_step)
    assert 0.05 < frac < 0.95, "the crossing must not sit on a step boundary"
    sim = orc.OracleSim(p, l, s)
    _, broken = sim.run(t_expect + 40, log=True)
    assert broken.tolist() == [[t_expect, 0]]
    assert sim.l.state[0] == 0       # never un-breaks
    np.testing.assert_allclose(sim.p["xCord"][1] - sim.p["xCord"][0], L0 + (t_expect + 40) * v_dt, rtol=1e-12)
    sim.close()


def test_gather_forces_sum_to_zero(orc):
    # isolated block, no clamp, overlapping neighbours (diameter > nearest-neighbour distance)
    rng = np.random.default_rng(11)
    s = quiet_setup(E=1e9, damping=1e4, dampingRotation=1e4)
    xyz = fcc_points(4, np.sqrt(2.0)) + rng.uniform(-0.01, 0.01, (256, 3)) + 20.0
    p = particles_at(xyz, 0.505)
    l = brute_links(xyz, 0.505, 1.05, rng)
    l.maxLoad[:] = 1.0               # no link breaks
    for f in ("xDiff", "yDiff", "zDiff"):
        p.arrays[f][:] = rng.uniform(-1e-5, 1e-5, len(xyz))
    for f in ("uDiff", "vDiff", "wDiff"):
        p.arrays[f][:] = rng.uniform(-1e-5, 1e-5, len(xyz))
    sim = orc.OracleSim(p, l, s)
    sim.run(20)
    sim.force()
    F = np.stack([sim.p["xForce"], sim.p["yForce"], sim.p["zForce"]], 1)
    scale = np.abs(F).sum(axis=0)
    assert np.all(scale > 0)
    assert np.all(np.abs(F.sum(axis=0)) <= 1e-12 * scale), (F.sum(axis=0), scale)
    # the block is in contact: the balance covers the contact law too
    counts, starts, lst = orc.neigh_build(sim.p, s.search_distance, s.interact_dist)
    x = np.stack([sim.p["xCord"], sim.p["yCord"], sim.p["zCord"]], 1)
    own = np.repeat(np.arange(len(x)), counts)
    overlap = np.linalg.norm(x[own] - x[lst], axis=1) < 2 * 0.505
    assert overlap.sum() > len(x)
    sim.close()
