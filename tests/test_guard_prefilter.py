"""Safety of the device guard's FP32 first stage (guard_trips in
hidem_amd/csrc/kernels_step.hip), restated in NumPy: wherever the test on
the rebuild snapshot rounded to FP32 decides, it decides as the exact FP64
test 2 |c - x0| >= threshold does (the rebuild test whose trip the learned
scheduler must not miss or invent).  Positions are drawn around the
threshold, where the FP32 snapshot is least certain, at slab-sized and
extreme magnitudes."""
import numpy as np

f32 = np.float32


def exact(c, x0, thr):
    d = c - x0
    with np.errstate(all="ignore"):
        return 2.0 * np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) >= thr


def first_stage(c, x0, thr):
    """(decided, verdict) of the FP32 stage: undecided entries read x0."""
    with np.errstate(all="ignore"):
        f = x0.astype(f32).astype(np.float64)
        a = c - f
        d = np.sqrt(a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1] + a[:, 2] * a[:, 2])
        e = (np.abs(f[:, 0]) + np.abs(f[:, 1]) + np.abs(f[:, 2])) * 2.0 ** -23 + 1e-30
        m = (d + e) * 1e-12
        no = 2.0 * (d + e + m) < thr
        yes = ((d - e - m) > 0.0) & (2.0 * (d - e - m) >= thr)
    assert not (no & yes).any()
    return no | yes, yes


def check(c, x0, thr):
    want = exact(c, x0, thr)
    decided, got = first_stage(c, x0, thr)
    bad = decided & (got != want)
    assert not bad.any(), np.nonzero(bad)[0][:5]
    return decided, want


def around(r, n, scale, thr, width):
    """Snapshots at |x0| ~ scale, new positions at distance thr/2 +- width."""
    x0 = r.uniform(-scale, scale, (n, 3))
    u = r.normal(0, 1, (n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    dist = thr / 2 + r.uniform(-width, width, n)
    return x0 + u * dist[:, None], x0


def test_first_stage_agrees_with_exact_test():
    r = np.random.default_rng(0)
    thr = 0.0141  # the slab's skin: 2 % of the 0.707 m spacing
    for scale, width in ((400.0, 1e-4), (400.0, 1e-7), (1.0, 1e-9), (5e4, 1e-2)):
        c, x0 = around(r, 200_000, scale, thr, width)
        check(c, x0, thr)
    # far from the threshold the FP32 stage decides nearly every particle
    c, x0 = around(r, 200_000, 400.0, thr, 0.5 * thr)
    decided, _ = check(c, x0, thr)
    assert decided.mean() > 0.99


def test_first_stage_extremes():
    r = np.random.default_rng(1)
    for scale, thr in ((1e-30, 1e-30), (1e-40, 1e-38), (1e30, 1e25), (3e38, 1.0), (1e300, 1.0), (0.0, 0.0)):
        c, x0 = around(r, 20_000, scale, thr, thr * 1e-3 + 1e-300)
        check(c, x0, thr)
    # non-finite values never decide in the FP32 stage
    x0 = np.array([[np.nan, 0, 0], [np.inf, 0, 0], [1e39, 0, 0], [0, 0, 0]])
    c = np.array([[0.0, 0, 0], [0.0, 0, 0], [0.0, 0, 0], [np.nan, 0, 0]])
    decided, _ = first_stage(c, x0, 0.1)
    assert not decided.any()
