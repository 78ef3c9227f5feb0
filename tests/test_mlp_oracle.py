"""Pins of the FP64 warm-start MLP oracle: closed forms, a finite-difference gradient check
and the bfloat16 rounding rule."""
import numpy as np
import pytest

SHAPES = ((64, 256), (256, 256), (256, 256), (256, 32))


def offsets():
    w, b, off = [], [], 0
    for k, n in SHAPES:
        w.append(off)
        off += k * n
        b.append(off)
        off += n
    return w, b, off


def random_params(rng, scale=1.0):
    wo, bo, n = offsets()
    p = np.zeros(n)
    for l, (k, nn) in enumerate(SHAPES):
        a = scale * np.sqrt(6.0 / (k + nn))
        p[wo[l]:wo[l] + k * nn] = rng.uniform(-a, a, k * nn)
        p[bo[l]:bo[l] + nn] = rng.uniform(-0.1, 0.1, nn)
    return p


def zero_padding(p, n_in, n_out):
    """Zero the padded rows/columns (inputs >= n_in, outputs >= n_out)."""
    wo, bo, _ = offsets()
    W1 = p[wo[0]:wo[0] + 64 * 256].reshape(64, 256)
    W1[n_in:] = 0
    W4 = p[wo[3]:wo[3] + 256 * 32].reshape(256, 32)
    W4[:, n_out:] = 0
    p[bo[3] + n_out:bo[3] + 32] = 0
    return p


def test_param_count(orc):
    assert orc.mlp_n_params() == offsets()[2] == 64 * 256 + 2 * 256 * 256 + 256 * 32 + 3 * 256 + 32


def test_single_path_closed_form(orc):
    """One nonzero weight per layer along a path input 2 -> 5 -> 7 -> 1 -> output 3: the output
    is w4 tanh(w3 tanh(w2 tanh(w1 x2 + b1) + b2) + b3) + b4, all other outputs their bias.
    Asymmetric indices catch a transposed weight or a wrong padded stride."""
    wo, bo, n = offsets()
    p = np.zeros(n)
    w = (0.7, -1.3, 0.9, 2.1)
    b = (0.05, -0.2, 0.3, 0.4)
    p[wo[0] + 2 * 256 + 5] = w[0]
    p[bo[0] + 5] = b[0]
    p[wo[1] + 5 * 256 + 7] = w[1]
    p[bo[1] + 7] = b[1]
    p[wo[2] + 7 * 256 + 1] = w[2]
    p[bo[2] + 1] = b[2]
    p[wo[3] + 1 * 32 + 3] = w[3]
    p[bo[3] + 3] = b[3]
    p[bo[3] + 0] = -0.5
    x = np.array([[0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [-1.0, 0.0, 2.0, 0.0, 0.0, 0.0]])
    y = orc.mlp_forward(p, x, 5)
    for r in range(2):
        h1 = np.tanh(w[0] * x[r, 2] + b[0])
        h2 = np.tanh(w[1] * h1 + b[1])
        h3 = np.tanh(w[2] * h2 + b[2])
        # unused hidden units still carry tanh(bias): only unit 1 of layer 3 feeds output 3
        exp = np.zeros(5)
        exp[0] = -0.5
        exp[3] = w[3] * h3 + b[3]
        assert np.allclose(y[r], exp, rtol=0, atol=1e-15)


def test_hidden_bias_propagates(orc):
    """Zero weights except the last layer: y = tanh(b3) W4 + b4."""
    rng = np.random.default_rng(1)
    wo, bo, n = offsets()
    p = np.zeros(n)
    b3 = rng.uniform(-1, 1, 256)
    W4 = rng.uniform(-1, 1, (256, 32))
    b4 = rng.uniform(-1, 1, 32)
    p[bo[2]:bo[2] + 256] = b3
    p[wo[3]:wo[3] + 256 * 32] = W4.reshape(-1)
    p[bo[3]:bo[3] + 32] = b4
    y = orc.mlp_forward(p, rng.normal(size=(3, 4)), 7)
    exp = np.tanh(b3) @ W4 + b4
    assert np.allclose(y, np.broadcast_to(exp[:7], (3, 7)), rtol=1e-14, atol=1e-14)


def test_finite_difference_gradient(orc):
    rng = np.random.default_rng(7)
    n_in, n_out, M = 6, 5, 9
    p = zero_padding(random_params(rng, 1.5), n_in, n_out)
    x = rng.normal(size=(M, n_in))
    t = rng.normal(size=(M, n_out))
    loss, g = orc.mlp_loss_grad(p, x, t)
    y = orc.mlp_forward(p, x, n_out)
    assert loss == pytest.approx(np.mean((y - t) ** 2), rel=1e-14)
    wo, bo, _ = offsets()
    picks = []
    for l, (k, nn) in enumerate(SHAPES):
        kin = n_in if l == 0 else k
        nout = n_out if l == 3 else nn
        for _ in range(6):
            picks.append(wo[l] + rng.integers(kin) * nn + rng.integers(nout))
        for _ in range(3):
            picks.append(bo[l] + rng.integers(nout))
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.mlp_loss_grad(pp, x, t)[0] - orc.mlp_loss_grad(pm, x, t)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-9), i


def test_gradient_of_padding_is_zero(orc):
    rng = np.random.default_rng(3)
    n_in, n_out = 4, 3
    p = zero_padding(random_params(rng), n_in, n_out)
    _, g = orc.mlp_loss_grad(p, rng.normal(size=(5, n_in)), rng.normal(size=(5, n_out)))
    wo, bo, _ = offsets()
    assert not g[wo[0] + n_in * 256:wo[0] + 64 * 256].any()
    assert not g[wo[3]:wo[3] + 256 * 32].reshape(256, 32)[:, n_out:].any()
    assert not g[bo[3] + n_out:bo[3] + 32].any()


def test_bf16_rounding_rule(orc):
    r = orc.bf16_round
    assert r(1.0) == 1.0
    assert r(1.0 + 2.0 ** -8) == 1.0                       # tie -> even (down)
    assert r(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6        # tie -> even (up)
    assert r(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert r(-0.1) == -0.10009765625
    assert r(0.0) == 0.0
    assert np.isinf(r(np.inf))


def test_bf16_emulation_error_bound(orc):
    """Rounding operands to bfloat16 (8 significant bits) moves the output by O(2^-8)."""
    rng = np.random.default_rng(5)
    p = zero_padding(random_params(rng), 6, 5)
    x = rng.uniform(-1, 1, (50, 6))
    y = orc.mlp_forward(p, x, 5)
    ye = orc.mlp_forward(p, x, 5, emulate_bf16=True)
    rel = np.linalg.norm(ye - y) / np.linalg.norm(y)
    assert 1e-5 < rel < 2e-2
