"""Redistribution order on the CPU: the oracle's hdmo_reorder (restating
HiDEMDistribution.h:150-181 on one rank) checked against what fixes it --
the four-level colour order recomputed independently, links in (old A,
old B) order with their endpoints renamed, the physical state unchanged as
a set, a line of particles coming out sorted, and the force field of the
reordered state equal to the original's up to summation order."""
import numpy as np

from hidem_amd import params

FIELDS = ("xCord", "yCord", "zCord", "xTheta", "yTheta", "zTheta", "xDiff", "yDiff", "zDiff",
          "uDiff", "vDiff", "wDiff", "radius")


def shuffled(p, l, seed):
    """The same system with particles in a random order (links renamed)."""
    r = np.random.default_rng(seed)
    perm = r.permutation(p.n)          # new i = old perm[i]
    inv = np.empty_like(perm)
    inv[perm] = np.arange(p.n)
    q, m = p.copy(), l.copy()
    for f in q.arrays:
        q.arrays[f][:] = p.arrays[f][perm]
    q.type[:] = p.type[perm]
    m.a[:] = inv[l.a]
    m.b[:] = inv[l.b]
    return q, m


def colours(x, y, z, s):
    """NeighbourGrid colours (bbox of all particles, 1 + floor(extent / s))."""
    mn = [v.min() for v in (x, y, z)]
    mx = [v.max() for v in (x, y, z)]
    X = 1 + int(np.floor((mx[0] - mn[0]) / s))
    Y = 1 + int(np.floor((mx[1] - mn[1]) / s))
    ix = np.floor((x - mn[0]) / s).astype(np.int64)
    iy = np.floor((y - mn[1]) / s).astype(np.int64)
    iz = np.floor((z - mn[2]) / s).astype(np.int64)
    return ix + iy * X + iz * X * Y


def system():
    s, p, l = params.build_inputs(params.load_config("iceblock_16k"), [8, 6, 5])
    r = np.random.default_rng(3)
    for c in range(6):
        l.ct[c][:] = r.normal(size=2 * l.n)
    l.state[r.random(l.n) < 0.1] = 0
    q, m = shuffled(p, l, 4)
    return s, q, m


def test_particle_order_is_four_level_colour_order(orc):
    s, p, l = system()
    q, m, order, lorder = orc.reorder(p, l, s.search_distance)
    assert np.array_equal(np.sort(order), np.arange(p.n))
    x, y, z = (np.asarray(p[f]) for f in ("xCord", "yCord", "zCord"))
    keys = [colours(x, y, z, s.search_distance * 2 ** k) for k in (3, 2, 1, 0)] + [np.arange(p.n)]
    want = np.lexsort(keys[::-1])      # primary key last for lexsort
    assert np.array_equal(order, want)
    for f in FIELDS:
        assert np.array_equal(q[f], np.asarray(p[f])[order]), f
    assert np.array_equal(q.type, p.type[order])


def test_links_follow_in_old_endpoint_order(orc):
    s, p, l = system()
    q, m, order, lorder = orc.reorder(p, l, s.search_distance)
    assert np.array_equal(lorder, np.lexsort((np.arange(l.n), l.b, l.a)))
    inv = np.empty_like(order)
    inv[order] = np.arange(p.n)
    assert np.array_equal(m.a, inv[l.a[lorder]]) and np.array_equal(m.b, inv[l.b[lorder]])
    assert np.array_equal(m.state, l.state[lorder])
    assert np.array_equal(m.initialLen, l.initialLen[lorder])
    assert np.array_equal(m.maxLoad, l.maxLoad[lorder])
    for c in range(6):
        sides = np.stack([2 * lorder, 2 * lorder + 1], 1).ravel()
        assert np.array_equal(m.ct[c], l.ct[c][sides]), c
    # each link still joins the same two points in space
    for f in ("xCord", "yCord", "zCord"):
        assert np.array_equal(np.asarray(q[f])[m.a], np.asarray(p[f])[l.a[lorder]])
        assert np.array_equal(np.asarray(q[f])[m.b], np.asarray(p[f])[l.b[lorder]])


def test_line_of_particles_comes_out_sorted(orc):
    """Particles 10 search distances apart on a line, in random order: every
    one has its own coarsest cell, so the order is by x."""
    from hidem_amd import _abi as A
    n = 50
    p = A.Particles.zeros(n)
    r = np.random.default_rng(1)
    xs = r.permutation(n) * 10.0
    p.arrays["xCord"][:] = xs
    p.arrays["radius"][:] = 0.5
    q, m, order, _ = orc.reorder(p, A.Links.zeros(0), 1.0)
    assert np.all(np.diff(np.asarray(q["xCord"])) > 0)


def test_force_is_invariant_up_to_summation_order(orc):
    s, p, l = system()
    r = np.random.default_rng(5)
    for f in ("xDiff", "yDiff", "zDiff", "uDiff", "vDiff", "wDiff"):
        p.arrays[f][:] = r.normal(scale=1e-5, size=p.n)
    q, m, order, _ = orc.reorder(p, l, s.search_distance)
    a = orc.OracleSim(p, l, s)
    a.rebuild()
    a.force()
    b = orc.OracleSim(q, m, s)
    b.rebuild()
    b.force()
    for f in ("xForce", "yForce", "zForce", "uForce", "vForce", "wForce"):
        want = np.asarray(a.p[f])[order]
        got = np.asarray(b.p[f])
        assert np.max(np.abs(got - want)) <= 1e-12 * max(np.max(np.abs(want)), 1.0), f
