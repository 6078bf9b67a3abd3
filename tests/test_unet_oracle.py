"""CPU checks that pin the learned preconditioner's oracle (oracle/unet.py).

The network has no counterpart in the reference, so its oracle is pinned by
mathematics: central finite differences of its own loss for every layer's
weight gradient, the adjoint identities of each layer type, linearity of the
whole network, and the finite-difference gradient of the A-norm loss. A
transposed operand, a wrong tap flip, a dropped skip/pool term or a wrong 0.25
factor fails one of them.
"""
import numpy as np
import pytest

from oracle import unet


def _weights(seed=3):
    w = unet.init_weights(seed)
    return {k: v.astype(np.float64) for k, v in w.items()}


def test_param_count_matches_layer_table():
    assert sum(9 * ci * co for _, ci, co in unet.LAYERS) == 302112
    w = unet.init_weights(0)
    v = unet.flatten(w)
    assert v.size == 302112
    back = unet.unflatten(v)
    for name, _, _ in unet.LAYERS:
        np.testing.assert_array_equal(back[name], w[name])


@pytest.mark.parametrize("cin,cout", [(1, 3), (4, 2), (3, 5)])
def test_conv_adjoint_identity(cin, cout):
    rng = np.random.default_rng(cin * 10 + cout)
    x = rng.standard_normal((2, 6, 5, cin))
    g = rng.standard_normal((2, 6, 5, cout))
    w = rng.standard_normal((3, 3, cin, cout))
    lhs = np.sum(unet.conv3x3(x, w) * g)
    rhs = np.sum(x * unet.conv3x3_input_grad(g, w))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # weight gradient: d/dw sum(conv(x, w) * g)
    dw = unet.conv3x3_weight_grad(x, g)
    assert np.sum(dw * w) == pytest.approx(lhs, rel=1e-12)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((1, 4, 5, 2))
    w = rng.standard_normal((3, 3, 2, 3))
    y = unet.conv3x3(x, w)
    ref = np.zeros_like(y)
    for i in range(4):
        for j in range(5):
            for ky in range(3):
                for kx in range(3):
                    ii, jj = i + ky - 1, j + kx - 1
                    if 0 <= ii < 4 and 0 <= jj < 5:
                        ref[0, i, j] += x[0, ii, jj] @ w[ky, kx]
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 8, 6, 3))
    g = rng.standard_normal((2, 4, 3, 3))
    assert np.sum(unet.pool(x) * g) == pytest.approx(np.sum(x * unet.pool_grad(g)), rel=1e-12)
    xc = rng.standard_normal((2, 4, 3, 3))
    gf = rng.standard_normal((2, 8, 6, 3))
    assert np.sum(unet.up(xc) * gf) == pytest.approx(np.sum(xc * unet.up_grad(gf)), rel=1e-12)
    # pool of a constant is the constant; up then pool is the identity
    np.testing.assert_allclose(unet.pool(unet.up(xc)), xc, rtol=0, atol=0)


def test_network_is_linear():
    w = _weights()
    rng = np.random.default_rng(2)
    x1 = rng.standard_normal((1, 16, 24, 1))
    x2 = rng.standard_normal((1, 16, 24, 1))
    y = unet.forward(w, 2.0 * x1 - 3.0 * x2)
    np.testing.assert_allclose(y, 2.0 * unet.forward(w, x1) - 3.0 * unet.forward(w, x2),
                               rtol=1e-10, atol=1e-10)


def test_backward_matches_finite_differences():
    w = _weights(5)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, 16, 16, 1))
    gy = rng.standard_normal((2, 16, 16, 1))

    def loss(wd):
        return float(np.sum(unet.forward(wd, x) * gy))

    _, tape = unet.forward(w, x, keep=True)
    dw, gx = unet.backward(w, tape, gy)
    eps = 1e-5
    for name, cin, cout in unet.LAYERS:
        picks = [tuple(rng.integers(0, s) for s in (3, 3, cin, cout)) for _ in range(3)]
        for idx in picks:
            wp = {k: v.copy() for k, v in w.items()}
            wm = {k: v.copy() for k, v in w.items()}
            wp[name][idx] += eps
            wm[name][idx] -= eps
            fd = (loss(wp) - loss(wm)) / (2 * eps)
            assert dw[name][idx] == pytest.approx(fd, rel=1e-6, abs=1e-8), (name, idx)
    # input gradient (the network is linear: sum(N x * gy) = (x, N^T gy))
    assert np.sum(gx * x) == pytest.approx(loss(w), rel=1e-10)


def test_anorm_loss_gradient_and_value():
    rng = np.random.default_rng(4)
    n, h, wd = 2, 6, 5
    y = rng.standard_normal((n, h, wd))
    e = rng.standard_normal((n, h, wd))
    ce = rng.uniform(0.1, 0.3, (n, h, wd))
    cn = rng.uniform(0.1, 0.3, (n, h, wd))
    ce[:, :, -1] = 0
    cn[:, -1, :] = 0
    den = rng.uniform(1, 2, n)
    L, g = unet.anorm_loss(y, e, ce, cn, den)
    # value against the dense operator
    tot = 0.0
    for s in range(n):
        A = np.eye(h * wd)
        for i in range(h):
            for j in range(wd):
                c = i * wd + j
                if j + 1 < wd:
                    A[c, c + 1] = A[c + 1, c] = -ce[s, i, j]
                if i + 1 < h:
                    A[c, c + wd] = A[c + wd, c] = -cn[s, i, j]
        d = (y[s] - e[s]).ravel()
        tot += d @ A @ d / den[s]
    assert L == pytest.approx(tot / n, rel=1e-12)
    eps = 1e-6
    for _ in range(5):
        idx = tuple(rng.integers(0, s) for s in y.shape)
        yp, ym = y.copy(), y.copy()
        yp[idx] += eps
        ym[idx] -= eps
        fd = (unet.anorm_loss(yp, e, ce, cn, den)[0] - unet.anorm_loss(ym, e, ce, cn, den)[0]) / (2 * eps)
        assert g[idx] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0])
    r = unet.bf(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa -> 1
    np.testing.assert_array_equal(r, [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0])


def test_split2_keeps_sixteen_bits():
    rng = np.random.default_rng(3)
    a = rng.standard_normal(10000)
    rel = np.abs(unet.split2(a) - unet.f32(a)) / np.abs(unet.f32(a))
    assert rel.max() < 2.0 ** -16
    # a BF16-representable value splits exactly
    np.testing.assert_array_equal(unet.split2(unet.bf(a)), unet.bf(a))


def test_device_emulation_tracks_exact_network():
    """The device's rounding points move the output by BF16 weight rounding
    (a fixed, slightly different linear operator) and ~2^-16 activation
    rounding only."""
    w = _weights(6)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((1, 16, 32, 1))
    wb = {k: unet.bf(v) for k, v in w.items()}
    yb, tape = unet.forward_device(w, x)
    y_wb = unet.forward(wb, x)  # exact arithmetic with the BF16 weights
    assert np.linalg.norm(yb - y_wb) / np.linalg.norm(y_wb) < 1e-4
    gy = rng.standard_normal(yb.shape)
    _, t64 = unet.forward(wb, x, keep=True)
    dw64, _ = unet.backward(wb, t64, gy)
    dwb = unet.backward_device(w, tape, gy)
    for name, _, _ in unet.LAYERS:
        r = np.linalg.norm(dwb[name] - dw64[name]) / np.linalg.norm(dw64[name])
        assert r < 2e-2, name
