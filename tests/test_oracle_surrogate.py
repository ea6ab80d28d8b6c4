"""Pins the surrogate oracle: direct-sum convolution, finite-difference
gradients, GELU closed form, BF16 rounding."""

import numpy as np

from oracle import surrogate_numpy as so


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1.0
    assert so.bf16(np.array([1.0 + 2**-8]))[0] == 1.0
    assert so.bf16(np.array([1.0 + 3 * 2**-8]))[0] == 1.0 + 2**-6
    assert so.bf16(np.array([-2.5]))[0] == -2.5
    x = np.random.randn(1000)
    r = so.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2**-8)


def test_gelu_closed_form():
    assert so.gelu(np.array([0.0]))[0] == 0.0
    # tanh form at z = 1: 0.5 (1 + tanh(sqrt(2/pi) 1.044715))
    exp = 0.5 * (1 + np.tanh(np.sqrt(2 / np.pi) * 1.044715))
    np.testing.assert_allclose(so.gelu(np.array([1.0]))[0], exp)
    z = np.linspace(-3, 3, 13)
    h = 1e-6
    np.testing.assert_allclose(so.gelu_grad(z), (so.gelu(z + h) - so.gelu(z - h)) / (2 * h), atol=1e-8)


def test_conv3x3_matches_direct_sum():
    rng = np.random.RandomState(0)
    x = rng.randn(2, 5, 6, 3)
    w = rng.randn(3, 3, 3, 4)
    b = rng.randn(4)
    out = so.conv3x3(x, w, b)
    ref = np.zeros_like(out)
    for n in range(2):
        for i in range(5):
            for j in range(6):
                for kh in range(3):
                    for kw in range(3):
                        ii, jj = i + kh - 1, j + kw - 1
                        if 0 <= ii < 5 and 0 <= jj < 6:
                            ref[n, i, j] += x[n, ii, jj] @ w[kh, kw]
    ref += b
    np.testing.assert_allclose(out, ref, atol=1e-12)


def test_conv_gradients_by_finite_differences():
    rng = np.random.RandomState(1)
    x = rng.randn(1, 4, 5, 2)
    w = rng.randn(3, 3, 2, 3)
    dy = rng.randn(1, 4, 5, 3)
    gx = so.conv3x3_input_grad(dy, w)
    gw = so.conv3x3_weight_grad(x, dy)
    h = 1e-6
    for idx in [(0, 1, 2, 0), (0, 0, 0, 1), (0, 3, 4, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (np.sum(so.conv3x3(xp, w) * dy) - np.sum(so.conv3x3(xm, w) * dy)) / (2 * h)
        np.testing.assert_allclose(gx[idx], fd, rtol=1e-6)
    for idx in [(0, 0, 0, 0), (1, 2, 1, 2), (2, 1, 0, 1)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        fd = (np.sum(so.conv3x3(x, wp) * dy) - np.sum(so.conv3x3(x, wm) * dy)) / (2 * h)
        np.testing.assert_allclose(gw[idx], fd, rtol=1e-6)


def test_network_gradients_by_finite_differences():
    rng = np.random.RandomState(2)
    params = so.init_params(1, 1, seed=3, hidden=4)
    # give the last layer non-zero weights so every gradient is exercised
    params[-1] = (rng.randn(*params[-1][0].shape) * 0.3, rng.randn(1) * 0.1)
    x = rng.randn(2, 5, 4, 1)
    t = rng.randn(2, 5, 4, 1)
    loss, grads = so.loss_and_grads(params, x, t, rounding=False)
    h = 1e-6
    for layer in (0, 2, 5):
        w, b = params[layer]
        idx = tuple(int(v) for v in np.unravel_index(7 % w.size, w.shape))
        wp = [(p[0].copy(), p[1].copy()) for p in params]
        wm = [(p[0].copy(), p[1].copy()) for p in params]
        wp[layer][0][idx] += h
        wm[layer][0][idx] -= h
        fd = (so.loss_and_grads(wp, x, t, False)[0] - so.loss_and_grads(wm, x, t, False)[0]) / (2 * h)
        np.testing.assert_allclose(grads[layer][0][idx], fd, rtol=1e-5, atol=1e-10)
        bp = [(p[0].copy(), p[1].copy()) for p in params]
        bm = [(p[0].copy(), p[1].copy()) for p in params]
        bp[layer][1][0] += h
        bm[layer][1][0] -= h
        fd = (so.loss_and_grads(bp, x, t, False)[0] - so.loss_and_grads(bm, x, t, False)[0]) / (2 * h)
        np.testing.assert_allclose(grads[layer][1][0], fd, rtol=1e-5, atol=1e-10)


def test_untrained_surrogate_is_identity():
    """The last layer starts at zero, so the increment is zero and the
    identity part is exact (it is never rounded to BF16)."""
    params = so.init_params(1, 1, seed=0, hidden=8)
    u = np.random.rand(2, 6, 7, 1)
    np.testing.assert_array_equal(so.predict(params, u), u)
