"""Safety of the fused step's break prefilter (certainly_intact in
hidem_amd/csrc/kernels_step.hip, 16-bit threshold codes from k_threshold_max
and k_side_codes in kernels_links.hip), restated in float32 NumPy: whenever
it declares a link intact, the exact FP64 HiDEMDefLinkBreakFunc
(HiDEMDefaultFunctions.h:217-241) does not break it.  Links are drawn around
the threshold (where the float estimate is least certain) and at extreme
magnitudes, with per-link and scalar maxLoad; the device sqrtf is allowed to
be off by 2 ulp either way."""
import numpy as np

f32 = np.float32


def float_up(v):
    f = v.astype(f32)
    return np.where(f.astype(np.float64) < v, np.nextafter(f, f32(np.inf)), f)


def side_threshold(L0, M):
    """k_threshold_max + k_side_codes + the step's decode: code = floor(T /
    scale) held to [0, 65535] with T = L0 (1 + maxLoad) per link (M None: L0)
    and scale = float(largest T rounded up) / 65535; the step reads
    float(code) * scale.  Code 0 where L0 < 1e-30 or not finite."""
    with np.errstate(all="ignore"):
        ok = (L0 >= 1e-30) & (L0 <= 1e300)
        T = L0 * (1.0 + M) if M is not None else L0
        okT = ok & (T > 0) & (T <= 1e300)
        tmax = float_up(T[okT]).max() if okT.any() else f32(0)
        scale = f32(tmax) / f32(65535)
        if not (scale > 0 and scale < 3e38):
            scale = f32(0)
        q = np.floor(T / np.float64(scale)) if scale > 0 else np.zeros_like(T)
        q = np.where(ok & (scale > 0), q, 0.0)
        q = np.where(q >= 0, q, 0.0)
        q = np.minimum(q, 65535.0)
        return q.astype(np.uint16).astype(f32) * scale


def l0_max(L0):
    """Largest L0 rounded up to float; inf when some L0 is NaN or huge."""
    with np.errstate(all="ignore"):
        ok = (L0 >= 1e-30) & (L0 <= 1e300)
        bad = ~ok & ~(L0 < 1e-30)
        if bad.any():
            return f32(np.inf)
        f = L0[ok].astype(f32)
        up = np.where(f.astype(np.float64) < L0[ok], np.nextafter(f, f32(np.inf)), f)
        return up.max() if up.size else f32(0)


def prefilter_intact(thr, scalarM, bbf, l0max, dx, dy, dz, rot, sqrt_ulps=0):
    fx, fy, fz = dx.astype(f32), dy.astype(f32), dz.astype(f32)
    with np.errstate(all="ignore"):
        s = np.sqrt(fx * fx + fy * fy + fz * fz)
        if sqrt_ulps:
            s = np.nextafter(s, np.where(sqrt_ulps > 0, f32(np.inf), f32(-np.inf)))
            if abs(sqrt_ulps) > 1:
                s = np.nextafter(s, np.where(sqrt_ulps > 0, f32(np.inf), f32(-np.inf)))
            s = np.maximum(s, f32(0))  # sqrtf of a non-negative value is never negative
        T = thr if scalarM is None else thr * (f32(1) + f32(scalarM))
        lhs = s + bbf * rot * l0max
        margin = f32(1e-5) * (lhs + np.abs(T))
        return lhs + margin < T


def exact_breaks(L0, M, bb, dx, dy, dz, tp, tq):
    with np.errstate(all="ignore"):
        S = np.sqrt(dx * dx + dy * dy + dz * dz)
        b = np.abs(tp - tq)
        inner = np.where(b[:, 1] < b[:, 2], b[:, 2], b[:, 1])
        mb = np.where(b[:, 0] < inner, inner, b[:, 0])
        load = (S - L0) / L0 + bb * mb
        return load >= M


def rot_bound(t):
    return (np.abs(t[:, 0]) + np.abs(t[:, 1]) + np.abs(t[:, 2])).astype(f32)


def check(L0, M, bb, dx, dy, dz, tp, tq):
    """Per-link maxLoad M (array) and, with the same values as one scalar
    each, the scalar-maxLoad form; returns the per-link verdicts."""
    brk = exact_breaks(L0, M, bb, dx, dy, dz, tp, tq)
    rot = rot_bound(tp) + rot_bound(tq)
    lm = l0_max(L0)
    ok = None
    for u in (-2, 0, 2):
        intact = prefilter_intact(side_threshold(L0, M), None, f32(bb), lm, dx, dy, dz, rot, u)
        bad = intact & brk
        assert not bad.any(), (u, np.nonzero(bad)[0][:5])
        ok = intact if ok is None else ok & intact
    for m in np.unique(M)[:: max(1, len(np.unique(M)) // 5)]:
        brk_s = exact_breaks(L0, np.full_like(L0, m), bb, dx, dy, dz, tp, tq)
        for u in (-2, 0, 2):
            intact = prefilter_intact(side_threshold(L0, None), m, f32(bb), lm, dx, dy, dz, rot, u)
            bad = intact & brk_s
            assert not bad.any(), ("scalar", m, u, np.nonzero(bad)[0][:5])
    return ok, brk


def test_prefilter_never_passes_a_breaking_link():
    r = np.random.default_rng(0)
    n = 400_000
    L0 = r.uniform(0.3, 2.0, n)
    M = r.uniform(1e-4, 6e-4, n)
    bb = 0.01
    tp = r.normal(0, 1e-3, (n, 3))
    tq = tp + r.normal(0, 1e-5, (n, 3))
    b = np.abs(tp - tq)
    mb = np.maximum(b[:, 0], np.maximum(b[:, 1], b[:, 2]))
    # strains within +-3e-6 (absolute) of the breaking point, and a few far away
    eps = M - bb * mb + r.uniform(-3e-6, 3e-6, n)
    eps[: n // 10] = r.uniform(-1e-3, 1e-3, n // 10)
    S = L0 * (1 + eps)
    d = r.normal(0, 1, (n, 3))
    d *= (S / np.linalg.norm(d, axis=1))[:, None]
    intact, brk = check(L0, M, bb, d[:, 0], d[:, 1], d[:, 2], tp, tq)
    # far from the threshold the prefilter decides almost every intact link
    # (the codes resolve T to 1.5e-5 of the largest T, here 1e-4 of the
    # shortest link's)
    far = np.arange(n) < n // 10
    assert intact[far & ~brk].mean() > 0.85


def test_prefilter_extremes():
    r = np.random.default_rng(1)
    cases = []
    for L0s, ds in ((1e-20, 1e-20), (1e-35, 1e-35), (1e25, 1e25), (1.0, 1e30), (1.0, 0.0)):
        n = 20000
        L0 = np.full(n, L0s) * r.uniform(0.5, 2, n)
        d = r.normal(0, ds, (n, 3))
        cases.append((L0, d))
    for L0, d in cases:
        n = len(L0)
        M = r.uniform(-1e-3, 1e-3, n)
        tp = r.normal(0, 1.0, (n, 3))
        tq = r.normal(0, 1.0, (n, 3))
        check(L0, M, 0.35, d[:, 0], d[:, 1], d[:, 2], tp, tq)
    # non-finite inputs always take the exact path
    n = 6
    L0 = np.array([np.nan, 1.0, 1.0, np.inf, 1.0, 1.0])
    d = np.array([[1.0, 0, 0], [np.nan, 0, 0], [np.inf, 0, 0], [1.0, 0, 0], [1.0, 0, 0], [1.0, 0, 0]])
    tp = np.zeros((n, 3))
    tp[4, 1] = np.nan
    tq = np.zeros((n, 3))
    M = np.array([1e-3, 1e-3, 1e-3, 1e-3, 1e-3, np.nan])
    rot = rot_bound(tp) + rot_bound(tq)
    intact = prefilter_intact(side_threshold(L0, M), None, f32(0.35), l0_max(L0), d[:, 0], d[:, 1], d[:, 2], rot)
    assert not intact.any()


def test_prefilter_tight_threshold():
    """No rotation and no bend term: the float estimate alone separates the
    links, so the margin is what keeps links within rounding of the
    threshold on the exact path."""
    r = np.random.default_rng(2)
    n = 400_000
    L0 = r.uniform(0.3, 2.0, n)
    M = r.uniform(1e-4, 6e-4, n)
    S = L0 * (1 + M) * (1 + r.uniform(-3e-7, 3e-7, n))
    S[: n // 10] = (L0 * (1 + M) * (1 - r.uniform(1e-4, 1e-2, n)))[: n // 10]
    d = r.normal(0, 1, (n, 3))
    d *= (S / np.linalg.norm(d, axis=1))[:, None]
    t = np.zeros((n, 3))
    intact, brk = check(L0, M, 0.0, d[:, 0], d[:, 1], d[:, 2], t, t)
    assert brk.sum() > n // 4
    # links 1e-4 (relative) below the threshold pass
    far = S < L0 * (1 + M) * (1 - 1e-4)
    assert far.sum() == n // 10 and intact[far].all()
