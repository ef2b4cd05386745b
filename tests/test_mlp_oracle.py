"""Pins of the surrogate MLP oracle (oracle/mlp_oracle.py) on the CPU.

The backward pass and the loss gradient are checked against central finite
differences of the unquantised float64 network, so a transposed operand, a
dropped SiLU' factor or a wrong class column fails; bf16 rounding is checked
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
    rng = np.random.default_rng(3)
    W, b = small_net(rng)
    x = rng.uniform(-1, 1, size=(64, O.IN_PAD))
    x[:, 5:] = 0
    dy = rng.normal(size=(64, O.OUT_PAD))
    y, hs, zs = O.forward(W, b, x, keep=True)
    dW, db = O.backward(W, hs, zs, dy)
    y0, hs0, zs0 = O.forward(W, b, x, keep=True, quant=False)
    dW0, db0 = O.backward(W, hs0, zs0, dy, quant=False)
    for l in range(len(W)):
        assert np.linalg.norm(dW[l] - dW0[l]) <= 2e-2 * np.linalg.norm(dW0[l])
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


def test_log2_domain_inference_is_the_same_network():
    # unquantised, the log2-domain restatement of inference equals the plain
    # float64 network: the factors -log2(e) and -ln 2 cancel exactly
    rng = np.random.default_rng(11)
    W, b = small_net(rng)
    x = rng.uniform(-1, 1, size=(9, O.IN_PAD))
    np.testing.assert_allclose(O.forward_infer(W, b, x, quant=False), O.exact_forward(W, b, x), rtol=1e-12, atol=1e-12)


def test_log2_domain_inference_quantised_close_to_exact():
    # with its bf16 rounding points it stays within the bf16 error of the
    # exact network, like forward(); and a dropped factor (the output layer's
    # -ln 2 left out) is far outside that
    rng = np.random.default_rng(12)
    W, b = small_net(rng)
    x = rng.uniform(-1, 1, size=(256, O.IN_PAD))
    x[:, 5:] = 0
    y0 = O.exact_forward(W, b, x)
    yi = O.forward_infer(W, b, x)
    yf = O.forward(W, b, x)
    tol = 5e-3 * (1 + np.abs(y0).max())
    assert np.abs(yi - y0).max() <= tol
    assert np.abs(yi - yf).max() <= tol
    W_bad = [w.copy() for w in W]
    W_bad[-1] = W_bad[-1] / np.float32(O.NEG_LN2)
    assert np.abs(O.forward_infer(W_bad, b, x) - y0).max() > 10 * tol


def test_log2_silu_closed_form():
    # u / (1 + 2^u) at u = -log2(e) z is -log2(e) silu(z): z = 0 -> 0,
    # z = 1 -> -log2(e) / (1 + e^-1), large |z| -> -log2(e) max(z, 0)
    z = np.array([0.0, 1.0, -40.0, 40.0])
    u = -z / np.log(2.0)
    h = u / (1.0 + np.exp2(u))
    want = -np.array([0.0, 1.0 / (1.0 + np.exp(-1.0)), 0.0, 40.0]) / np.log(2.0)
    np.testing.assert_allclose(h, want, rtol=1e-12, atol=1e-15)
