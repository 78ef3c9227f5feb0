"""BF16-MFMA MLP kernels against the FP64 oracle (with and without the kernels' bfloat16
operand rounding)."""
import numpy as np
import pytest

from test_mlp_oracle import SHAPES, offsets, random_params, zero_padding

pytestmark = pytest.mark.gpu

N_IN, N_OUT = 6, 5


def rel(a, b):
    return np.linalg.norm(np.asarray(a, float) - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def net(gr):
    rng = np.random.default_rng(11)
    p = zero_padding(random_params(rng, 1.5), N_IN, N_OUT).astype(np.float32)
    n = gr.Mlp(N_IN, N_OUT, seed=1)
    n.set_params(p)
    return n, p.astype(np.float64)


def test_default_init_padding_is_zero(gr):
    n = gr.Mlp(3, 4, seed=5)
    p = n.params()
    wo, bo, total = offsets()
    assert p.size == total
    assert not p[wo[0] + 3 * 256:wo[0] + 64 * 256].any()
    assert not p[wo[3]:wo[3] + 256 * 32].reshape(256, 32)[:, 4:].any()
    assert not p[bo[0]:bo[0] + 256].any()
    assert np.abs(p[wo[1]:wo[1] + 256 * 256]).max() <= np.sqrt(6 / 512) + 1e-7


def test_forward_matches_oracle(gr, orc, net):
    n, p = net
    x = np.random.default_rng(2).uniform(-1, 1, (1000, N_IN)).astype(np.float32)
    y = n.forward(x)
    ye = orc.mlp_forward(p, x.astype(np.float64), N_OUT, emulate_bf16=True)
    y64 = orc.mlp_forward(p, x.astype(np.float64), N_OUT)
    print("forward rel err vs bf16-emulating oracle", rel(y, ye), "vs fp64", rel(y, y64))
    assert rel(y, ye) < 1e-4
    assert rel(y, y64) < 2e-2


def test_forward_rows_independent(gr, net):
    """Row r of the output depends on row r of the input only (also across the 2^18-row chunks)."""
    n, _ = net
    x = np.random.default_rng(3).uniform(-1, 1, ((1 << 18) + 300, N_IN)).astype(np.float32)
    y = n.forward(x)
    idx = np.array([0, 5, 127, 128, 4095, (1 << 18) - 1, 1 << 18, (1 << 18) + 299])
    assert np.array_equal(n.forward(x[idx]), y[idx])


def test_loss_gradient_matches_oracle(gr, orc, net):
    n, p = net
    rng = np.random.default_rng(4)
    M = 4096 + 77                                   # two row chunks of the weight-gradient kernel
    x = rng.uniform(-1, 1, (M, N_IN)).astype(np.float32)
    t = rng.normal(size=(M, N_OUT)).astype(np.float32)
    loss, g = n.loss_grad(x, t)
    le, ge = orc.mlp_loss_grad(p, x.astype(np.float64), t.astype(np.float64), emulate_bf16=True)
    l64, g64 = orc.mlp_loss_grad(p, x.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(le, rel=1e-5)
    wo, bo, _ = offsets()
    for l, (k, nn) in enumerate(SHAPES):
        for name, sl in (("W", slice(wo[l], wo[l] + k * nn)), ("b", slice(bo[l], bo[l] + nn))):
            e, e64 = rel(g[sl], ge[sl]), rel(g[sl], g64[sl])
            print(f"layer {l} {name}: rel err vs emulating oracle {e:.2e}, vs fp64 {e64:.2e}")
            assert e < 2e-3
            assert e64 < 5e-2
    # parameters unchanged by a gradient evaluation
    assert np.array_equal(n.params(), p.astype(np.float32))


def test_training_reduces_loss_and_is_deterministic(gr):
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (8192, 3)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] ** 2 - 0.3], axis=1).astype(np.float32)
    runs = []
    for _ in range(2):
        n = gr.Mlp(3, 2, seed=21)
        losses = [n.train_step(x[i * 1024:(i + 1) * 1024], t[i * 1024:(i + 1) * 1024], 1e-3)
                  for i in range(8) for _ in range(25)]
        runs.append((losses, n.params()))
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])
    losses = runs[0][0]
    print("loss first/last", losses[0], losses[-1])
    assert losses[-1] < 0.1 * losses[0]


def test_wide_network_matches_oracle_and_both_training_paths_agree(gr, orc):
    """40 inputs and 20 outputs: the first layer over all 64 padded input columns (the narrow
    networks take a 32-column product), five 8-column input groups in the resident-set gather."""
    n_in, n_out = 40, 20
    rng = np.random.default_rng(31)
    p = zero_padding(random_params(rng, 1.0), n_in, n_out).astype(np.float32)
    net = gr.Mlp(n_in, n_out, seed=3)
    net.set_params(p)
    p64 = p.astype(np.float64)
    x = rng.uniform(-1, 1, (777, n_in)).astype(np.float32)
    y = net.forward(x)
    assert rel(y, orc.mlp_forward(p64, x.astype(np.float64), n_out, emulate_bf16=True)) < 1e-4
    M = 4096
    xa = rng.uniform(-1, 1, (M, n_in)).astype(np.float32)
    ta = rng.normal(size=(M, n_out)).astype(np.float32)
    loss, g = net.loss_grad(xa, ta)
    le, ge = orc.mlp_loss_grad(p64, xa.astype(np.float64), ta.astype(np.float64), emulate_bf16=True)
    assert loss == pytest.approx(le, rel=1e-5)
    assert rel(g, ge) < 2e-3
    # the resident set's gather feeds the same rows (one batch = the whole set, permuted) as a
    # host batch: the same first-step loss, and after one Adam step the same second-step loss
    # up to the order of the row sums
    a = gr.Mlp(n_in, n_out, seed=3)
    a.set_params(p)
    a.set_data(xa, ta)
    la = a.train(M, 2, 1e-3, 1e-3, seed=5)
    b = gr.Mlp(n_in, n_out, seed=3)
    b.set_params(p)
    lb = [b.train_step(xa, ta, 1e-3) for _ in range(2)]
    assert la[0] == pytest.approx(lb[0], rel=1e-9)
    assert la[1] == pytest.approx(lb[1], rel=1e-4)
    assert la[1] < la[0]


def test_invalid_arguments(gr):
    with pytest.raises(gr.GraphiteError):
        gr.Mlp(0, 3)
    with pytest.raises(gr.GraphiteError):
        gr.Mlp(3, 33)
