"""Pins the scheduler network's CPU oracle: gradients of the float64 form by
central finite differences, the censored loss, and the bfloat16 rounding of
the device-emulating form against torch's bfloat16 conversion."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _data(n=48, seed=1):
    r = np.random.default_rng(seed)
    X = r.normal(0.5, 2.0, (n, M.NF))
    t = r.normal(3.0, 1.0, n)
    return X, t


def test_finite_difference_gradient():
    X, t = _data()
    w = M.init_weights(3, mean=np.full(M.NF, 0.5), std=np.full(M.NF, 2.0)).astype(np.float64)
    u = M.unpack(w)
    u["b1"][...] = 0.1
    u["b2"][...] = 0.05
    loss, g = M.loss_and_grad(w, X, t)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(M.N_PARAMS - 1, 60, replace=False), [M.N_PARAMS - 1]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (M.loss_and_grad(wp, X, t)[0] - M.loss_and_grad(wm, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])
    assert loss > 0


def test_gradient_closed_form_for_linear_output():
    # with W1 = W2 = 0 the hidden activations are relu(biases): the loss is a
    # quadratic in b3 whose gradient is 2 sum(y - t)
    X, t = _data(16)
    w = np.zeros(M.N_WEIGHTS)
    u = M.unpack(w)
    u["std"][...] = 1.0
    u["b3"] = 0.7
    w[M.N_PARAMS - 1] = 0.7
    loss, g = M.loss_and_grad(w, X, t)
    assert g[M.N_PARAMS - 1] == pytest.approx(2 * np.sum(0.7 - t))
    assert loss == pytest.approx(np.sum((0.7 - t) ** 2))
    assert np.all(g[:M.N_PARAMS - 1] == 0)


def test_censored_rows_only_pull_up():
    X, t = _data(32)
    w = M.init_weights(5).astype(np.float64)
    y = M.forward(w, X)
    cen = np.ones(32, np.uint8)
    t_low = y - 1.0          # predictions above every censored target: no loss, no gradient
    loss, g = M.loss_and_grad(w, X, t_low, cen)
    assert loss == 0.0 and np.all(g == 0)
    t_high = y + 1.0         # below the targets: same as uncensored
    l1, g1 = M.loss_and_grad(w, X, t_high, cen)
    l2, g2 = M.loss_and_grad(w, X, t_high, None)
    assert l1 == pytest.approx(l2) and np.allclose(g1, g2)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    r = np.random.default_rng(11)
    x = np.concatenate([r.normal(0, 1, 10000), r.normal(0, 1e-30, 100), r.normal(0, 1e30, 100),
                        [1.0, 1.00390625, 1.01171875, -2.5, 0.0, -0.0, np.inf, -np.inf]]).astype(np.float32)
    ours = M.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert M.bf16(np.float32(1.00390625)) == 1.0          # tie to even
    assert M.bf16(np.float32(1.01171875)) == 1.015625     # tie to even (upwards)
    assert np.isnan(M.bf16(np.float32(np.nan)))


def test_bf16_form_tracks_exact_form():
    X, t = _data(256, seed=4)
    w = M.init_weights(9, mean=np.full(M.NF, 0.5), std=np.full(M.NF, 2.0))
    ye = M.forward(w.astype(np.float64), X)
    yb = M.forward(w, X, emulate_bf16=True)
    assert np.max(np.abs(yb - ye)) <= 0.05 * (1 + np.max(np.abs(ye)))
    le, ge = M.loss_and_grad(w.astype(np.float64), X, t)
    lb, gb = M.loss_and_grad(w, X, t, emulate_bf16=True)
    assert np.linalg.norm(gb - ge) <= 0.05 * np.linalg.norm(ge)
    assert abs(lb - le) <= 0.05 * le
