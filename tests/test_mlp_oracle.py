"""Pins of the surrogate MLP oracle (oracle/mlp_oracle.py) on the CPU.

The backward pass and the loss gradient are checked against central finite
differences of the unquantised float64 network, so a transposed operand, a
dropped ReLU' factor or a wrong class column fails; bf16 rounding is checked
against exact bit patterns.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as O


def test_bf16_rounding_bits():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, -2.5, 3.0e38, 0.0],
                 dtype=np.float32)
    y = O.bf16(x)
    # ties go to even: 1 + 2^-8 -> 1;  1 + 3*2^-9 rounds up to 1 + 2^-7;  1 + 2^-7 + 2^-8 (tie) -> 1 + 2^-6
    assert y.tolist() == [1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, float(np.float32(O.bf16(np.float32(3.0e38)))), 0.0]
    assert (y.view(np.uint32) & 0xFFFF).max() == 0
    assert np.isnan(O.bf16(np.array([np.nan], dtype=np.float32)))[0]


def small_net(rng, width=None):
    shapes = O.layer_shapes()
    W = [rng.normal(0, 1.0 / np.sqrt(k), size=(k, n)) for k, n in shapes]
    b = [rng.normal(0, 0.1, size=n) for _, n in shapes]
    return W, b


def test_unquantised_forward_is_exact_forward():
    rng = np.random.default_rng(1)
    W, b = small_net(rng)
    x = rng.uniform(-1, 1, size=(7, O.IN_PAD))
    np.testing.assert_allclose(O.forward(W, b, x, quant=False), O.exact_forward(W, b, x), rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    W, b = small_net(rng)
    x = rng.uniform(-1, 1, size=(5, O.IN_PAD))
    dy = rng.normal(size=(5, O.OUT_PAD))
    y, hs, zs = O.forward(W, b, x, keep=True, quant=False)
    dW, db = O.backward(W, hs, zs, dy, quant=False)

    def L(Wp, bp):
        return float((O.exact_forward(Wp, bp, x) * dy).sum())

    eps = 1e-6
    for l in range(len(W)):
        for _ in range(6):
            i, j = rng.integers(W[l].shape[0]), rng.integers(W[l].shape[1])
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][i, j] += eps
            Wm[l][i, j] -= eps
            fd = (L(Wp, b) - L(Wm, b)) / (2 * eps)
            assert abs(fd - dW[l][i, j]) <= 1e-6 * (1 + abs(fd)), (l, i, j, fd, dW[l][i, j])
        j = rng.integers(b[l].shape[0])
        bp = [v.copy() for v in b]
        bm = [v.copy() for v in b]
        bp[l][j] += eps
        bm[l][j] -= eps
        fd = (L(W, bp) - L(W, bm)) / (2 * eps)
        assert abs(fd - db[l][j]) <= 1e-6 * (1 + abs(fd))


def test_quantised_backward_close_to_exact():
    """bf16 rounding moves a few pre-activations across zero, where ReLU'
    jumps; taken with the quantised forward's ReLU pattern, the exact
    float64 backward stays within the bf16 error of the quantised one."""
    rng = np.random.default_rng(3)
    W, b = small_net(rng)
    x = rng.uniform(-1, 1, size=(64, O.IN_PAD))
    x[:, 5:] = 0
    dy = rng.normal(size=(64, O.OUT_PAD))
    y, hs, zs = O.forward(W, b, x, keep=True)
    dW, db = O.backward(W, hs, zs, dy)
    y0, hs0, zs0 = O.forward(W, b, x, keep=True, quant=False)
    flips = np.mean([np.mean((z > 0) != (z0 > 0)) for z, z0 in zip(zs, zs0)])
    assert flips < 2e-2
    dW0, db0 = O.backward(W, hs0, zs, dy, quant=False)
    for l in range(len(W)):
        assert np.linalg.norm(dW[l] - dW0[l]) <= 2e-2 * np.linalg.norm(dW0[l]), l
    assert np.abs(y - y0).max() <= 5e-2 * (1 + np.abs(y0).max())


@pytest.mark.parametrize("cls", [0, 1, 2])
def test_loss_gradient_matches_finite_differences(cls):
    rng = np.random.default_rng(10 + cls)
    n = 6
    y = rng.normal(size=(n, O.OUT_PAD))
    t = np.zeros((n, 8))
    t[:, 0] = cls
    t[:, 1:5] = rng.normal(size=(n, 4))
    w_r, w_d = 3.0, 0.5
    sums, d = O.loss(y, t, w_r, w_d)

    def mean_loss(yy):
        s, _ = O.loss(yy, t, w_r, w_d)
        return (s[0] + s[1] + s[2]) / n

    eps = 1e-6
    for i in range(n):
        for c in range(8):
            yp, ym = y.copy(), y.copy()
            yp[i, c] += eps
            ym[i, c] -= eps
            fd = (mean_loss(yp) - mean_loss(ym)) / (2 * eps)
            assert abs(fd - d[i, c]) <= 1e-7 + 1e-6 * abs(fd), (i, c, fd, d[i, c])
    assert np.all(d[:, 8:] == 0)


def test_loss_cross_entropy_closed_form():
    # equal logits -> p = 1/3 for every class
    y = np.zeros((3, O.OUT_PAD))
    t = np.zeros((3, 8))
    t[:, 0] = [0, 1, 2]
    sums, d = O.loss(y, t, 1.0, 1.0)
    assert sums[0] == pytest.approx(3 * np.log(3.0))
    assert sums[1] == 0 and sums[2] == 0


def test_relu_closed_form():
    z = np.array([-3.0, -1e-30, 0.0, 1e-30, 2.5])
    np.testing.assert_array_equal(O.relu(z), [0.0, 0.0, 0.0, 1e-30, 2.5])
    np.testing.assert_array_equal(O.drelu(z), [0.0, 0.0, 0.0, 1.0, 1.0])


def test_hand_wired_network():
    """A network wired by hand: hidden unit 0 of layer 1 is relu(x0 - 0.5),
    units 0 of layers 2..4 pass it on (weight 1, or 2 into layer 3 with bias
    -1), the output y1 = 3 h - 0.25 and y0 = -h.  Every other weight is zero,
    so y1 = 3 relu(2 relu(x0 - 0.5) - 1) - 0.25, exactly representable in bf16
    for these inputs."""
    shapes = O.layer_shapes()
    W = [np.zeros(s, np.float32) for s in shapes]
    b = [np.zeros(s[1], np.float32) for s in shapes]
    W[0][0, 0], b[0][0] = 1.0, -0.5
    W[1][0, 0] = 1.0
    W[2][0, 0], b[2][0] = 2.0, -1.0
    W[3][0, 0] = 1.0
    W[4][0, 1], b[4][1] = 3.0, -0.25
    W[4][0, 0] = -1.0
    x = np.zeros((4, O.IN_PAD), np.float32)
    x[:, 0] = [0.25, 0.75, 1.5, 3.0]
    h = np.maximum(2 * np.maximum(x[:, 0] - 0.5, 0) - 1, 0)
    for quant in (True, False):
        y = O.forward(W, b, x, quant=quant)
        np.testing.assert_array_equal(y[:, 1], 3 * h - 0.25)
        np.testing.assert_array_equal(y[:, 0], -h)
        assert np.all(y[:, 2:] == 0)
