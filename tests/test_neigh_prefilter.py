"""Safety of the neighbour-list build's FP32 first test (k_neigh in
hidem_amd/csrc/kernels_grid.hip), restated in NumPy: wherever the test on
FP32 records decides, it decides as the FP64 criterion of
NeighAccel.h:124-131 (sqrt(dx^2 + dy^2 + dz^2) - (r_i + r_j) < interactDist)
does, so the list is the one the FP64 test builds.  The FP32 test compares
the squared distance with (T -+ E)^2, T = interactDist + r_i + r_j; the
squared distance is taken with two fused multiply-adds, emulated here in
FP64 (a double rounding may differ from the device by an ulp, so the check
also runs with it moved 2 ulps either way).  Pairs are drawn around the
threshold at slab-sized and extreme magnitudes, and a bound sixteen times
too small must be caught."""
import numpy as np

f32 = np.float32


def exact(P, Q, idist):
    d = P[:, :3] - Q[:, :3]
    with np.errstate(all="ignore"):
        return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) - (P[:, 3] + Q[:, 3]) < idist


def fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f32)


E_SCALE = f32(2.0 ** -22)


def first_test(P, Q, idist, s_ulps=0, e_scale=E_SCALE):
    """(decided, near) of the FP32 stage."""
    with np.errstate(all="ignore"):
        Pf, Qf = P.astype(f32), Q.astype(f32)
        idf = f32(idist)
        fx, fy, fz = Pf[:, 0] - Qf[:, 0], Pf[:, 1] - Qf[:, 1], Pf[:, 2] - Qf[:, 2]
        s2 = fma32(fx, fx, fma32(fy, fy, fz * fz))
        # an overflowed square stays overflowed: only finite values move
        for _ in range(abs(s_ulps)):
            s2 = np.where(np.isfinite(s2), np.nextafter(s2, f32(np.inf) if s_ulps > 0 else f32(-np.inf)), s2)
        s2 = np.maximum(s2, f32(0))
        Rp = idf + Pf[:, 3]
        Cp = np.abs(Pf[:, 0]) + np.abs(Pf[:, 1]) + np.abs(Pf[:, 2]) + f32(4) * (np.abs(idf) + np.abs(Pf[:, 3]))
        T = Rp + Qf[:, 3]
        acc = Cp + np.abs(Qf[:, 0]) + np.abs(Qf[:, 1]) + np.abs(Qf[:, 2]) + f32(4) * np.abs(Qf[:, 3]) + f32(8) * np.abs(T)
        E = fma32(acc, np.full_like(acc, f32(e_scale)), np.full_like(acc, f32(1e-20)))
        lo, hi = T - E, T + E
        yes = (lo > 0) & (s2 < lo * lo)
        no = (s2 >= hi * hi) & (hi * hi <= np.finfo(f32).max)
    assert not (yes & no).any()
    return yes | no, yes


def check(P, Q, idist, e_scale=E_SCALE):
    want = exact(P, Q, idist)
    decided = None
    for u in (-2, 0, 2):
        dec, near = first_test(P, Q, idist, u, e_scale)
        bad = dec & (near != want)
        assert not bad.any(), (u, np.nonzero(bad)[0][:5])
        decided = dec if decided is None else decided & dec
    return decided


def pairs(r, n, scale, rad, idist, width):
    """Particles at |x| ~ scale with radii ~ rad, partners at a surface gap
    of idist +- width."""
    P = np.empty((n, 4))
    P[:, :3] = r.uniform(-scale, scale, (n, 3))
    P[:, 3] = rad * r.uniform(0.9, 1.1, n)
    Q = np.empty((n, 4))
    Q[:, 3] = rad * r.uniform(0.9, 1.1, n)
    u = r.normal(0, 1, (n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    gap = idist + r.uniform(-width, width, n)
    Q[:, :3] = P[:, :3] + u * (P[:, 3] + Q[:, 3] + gap)[:, None]
    return P, Q


def test_first_test_agrees_with_fp64_criterion():
    r = np.random.default_rng(3)
    for scale, rad, idist, width in ((400.0, 0.33, 0.0141, 1e-2), (400.0, 0.33, 0.0141, 1e-6),
                                     (1.0, 0.5, 0.2, 1e-8), (2e4, 0.5, 0.2, 0.1), (400.0, 0.33, -0.01, 1e-3)):
        P, Q = pairs(r, 200_000, scale, rad, idist, width)
        check(P, Q, idist)
    # candidates of a stencil lie at all distances up to ~2 cells: nearly
    # all of them are decided in FP32 on the 400 m slab
    P, Q = pairs(r, 200_000, 400.0, 0.33, 0.0141, 0.8)
    assert check(P, Q, 0.0141).mean() > 0.99


def test_first_test_extremes():
    r = np.random.default_rng(4)
    for scale, rad, idist in ((1e-30, 1e-31, 1e-32), (1e-40, 1e-41, 1e-41), (1e30, 1e29, 1e28), (3e38, 1.0, 1.0),
                              (1e300, 1.0, 1.0), (1.0, 0.5, 1e39), (0.0, 0.0, 0.0)):
        P, Q = pairs(r, 20_000, scale, rad, idist, abs(idist) * 1e-3 + 1e-300)
        check(P, Q, idist)
    # non-finite records never decide in FP32
    P = np.array([[np.nan, 0, 0, 0.5], [np.inf, 0, 0, 0.5], [1e39, 0, 0, 0.5]])
    Q = np.zeros((3, 4))
    decided, _ = first_test(P, Q, 0.1)
    assert not decided.any()


def test_undersized_bound_is_caught():
    """The pair sets above are fine enough to fail a bound 16x below the
    kernel's (i.e. below what the rounding analysis requires)."""
    r = np.random.default_rng(3)
    P, Q = pairs(r, 200_000, 400.0, 0.33, 0.0141, 1e-4)
    try:
        check(P, Q, 0.0141, e_scale=E_SCALE / f32(16))
    except AssertionError:
        return
    raise AssertionError("a bound 16x too small went unnoticed")
