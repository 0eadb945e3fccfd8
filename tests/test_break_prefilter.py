"""Safety of the fused step's break prefilter (certainly_intact in
hidem_amd/csrc/kernels_step.hip), restated in float32 NumPy: whenever it
declares a link intact, the exact FP64 HiDEMDefLinkBreakFunc
(HiDEMDefaultFunctions.h:217-241) does not break it.  Links are drawn around
the threshold (where the float estimate is least certain) and at extreme
magnitudes; the device sqrtf is allowed to be off by 2 ulp either way."""
import numpy as np

f32 = np.float32


def prefilter_intact(L0f, Mf, bbf, dx, dy, dz, rot, sqrt_ulps=0):
    fx, fy, fz = dx.astype(f32), dy.astype(f32), dz.astype(f32)
    with np.errstate(all="ignore"):
        s = np.sqrt(fx * fx + fy * fy + fz * fz)
        if sqrt_ulps:
            s = np.nextafter(s, np.where(sqrt_ulps > 0, f32(np.inf), f32(-np.inf)))
            if abs(sqrt_ulps) > 1:
                s = np.nextafter(s, np.where(sqrt_ulps > 0, f32(np.inf), f32(-np.inf)))
        bend = bbf * rot * L0f
        lhs = (s - L0f) + bend
        rhs = Mf * L0f
        margin = f32(1e-5) * (s + L0f + bend + np.abs(rhs) + np.abs(lhs))
        return (L0f >= f32(1e-30)) & (lhs + margin < rhs)


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
    brk = exact_breaks(L0, M, bb, dx, dy, dz, tp, tq)
    rot = rot_bound(tp) + rot_bound(tq)
    ok = None
    for u in (-2, 0, 2):
        intact = prefilter_intact(L0.astype(f32), M.astype(f32), f32(bb), dx, dy, dz, rot, u)
        bad = intact & brk
        assert not bad.any(), (u, np.nonzero(bad)[0][:5])
        ok = intact if ok is None else ok & intact
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
    far = np.arange(n) < n // 10
    assert intact[far & ~brk].mean() > 0.9


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
    intact = prefilter_intact(L0.astype(f32), M.astype(f32), f32(0.35), d[:, 0], d[:, 1], d[:, 2], rot)
    assert not intact.any()
