"""NumPy oracle of the surrogate MLP (TEST INFRASTRUCTURE ONLY).

Mirrors the quantisation points of the HIP kernels in
magik_amd/csrc/mgk_mlp.hip so that forward outputs and gradients can be
compared at tight tolerance.  Training (k_gemm), forward():

    h0 = bf16(x)                                (features padded to 16)
    z_l = h_{l-1} @ bf16(W_l) + b_l              (FP32 accumulation), l = 1..4
    h_l = bf16(relu(z_l))
    y   = h_4 @ bf16(W_5) + b_5                  (FP32)

Backward (the kernels' order):
    g_5 = bf16(dy);  dW_5 = h_4^T g_5;  db_5 = sum dy
    dz_l = (g_{l+1} @ bf16(W_{l+1})^T) * relu'(z_l);  g_l = bf16(dz_l)
    dW_l = h_{l-1}^T g_l;  db_l = sum dz_l

Inference (k_mlp_infer) has the same rounding points: bf16 features, bf16
weights, FP32 accumulators (started from the bias), bf16 ReLU outputs; only
its FP32 summation order differs, so forward() is its oracle too.

Accumulation here is float64; the kernels accumulate in FP32, so agreement is
to ~1e-5 relative, far inside the bf16 quantisation the two share.
"""
from __future__ import annotations

import numpy as np

IN_PAD = 16
HIDDEN = 128
N_HIDDEN = 4
OUT = 8
OUT_PAD = 32


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest-even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rnd = ((u >> 16) & 1) + 0x7FFF
    u = ((u + rnd) >> 16) << 16
    out = u.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    out = np.where(nan, a, out)
    return out.reshape(np.shape(x))


def relu(z):
    return np.maximum(z, 0.0)


def drelu(z):
    return (z > 0.0).astype(np.float64)


def layer_shapes():
    dims = [IN_PAD] + [HIDDEN] * N_HIDDEN + [OUT_PAD]
    return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]


def _q(x, quant):
    return bf16(x) if quant else x


def forward(W: list, b: list, x: np.ndarray, keep: bool = False, quant: bool = True):
    """W[l]: (in, out) float32 master weights (padded shapes); x: (n, 16).
    quant=False drops every bf16 rounding and float32 cast (float64 throughout)."""
    dt = np.float32 if quant else np.float64
    h = _q(x.astype(dt), quant)
    hs, zs = [h], []
    for l in range(len(W)):
        z = h.astype(np.float64) @ _q(W[l].astype(dt), quant).astype(np.float64) + b[l].astype(np.float64)
        z = z.astype(dt)
        if l < len(W) - 1:
            zs.append(z)
            h = _q(relu(z.astype(np.float64)).astype(dt), quant)
            hs.append(h)
        else:
            y = z
    if keep:
        return y, hs, zs
    return y


def backward(W: list, hs: list, zs: list, dy: np.ndarray, quant: bool = True):
    """Gradients (dW, db) for dL/dy = dy (n, 32)."""
    dt = np.float32 if quant else np.float64
    n_l = len(W)
    dW, db = [None] * n_l, [None] * n_l
    g = _q(dy.astype(dt), quant)
    dz = dy.astype(np.float64)
    for l in range(n_l - 1, -1, -1):
        dW[l] = (hs[l].astype(np.float64).T @ g.astype(np.float64)).astype(dt)
        db[l] = dz.sum(axis=0).astype(dt)
        if l > 0:
            dh = g.astype(np.float64) @ _q(W[l].astype(dt), quant).astype(np.float64).T
            dz = (dh * drelu(zs[l - 1].astype(np.float64))).astype(dt).astype(np.float64)
            g = _q(dz.astype(dt), quant)
    return dW, db


def loss(y: np.ndarray, t: np.ndarray, w_r: float, w_d: float):
    """The training loss of the kernels (k_loss): logits (y0 captured, y1 disk,
    0 escape) -> cross-entropy on the class t0; disk rows regress y2 -> t1;
    escape rows regress y3..y5 -> t2..t4.  Returns (sums [ce, r, dir, correct],
    dL/dy of the batch-mean loss, (n, 32))."""
    y = y.astype(np.float64)
    n = y.shape[0]
    cls = t[:, 0].astype(int)
    logits = np.stack([y[:, 0], y[:, 1], np.zeros(n)], axis=1)
    mx = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - mx)
    p = e / e.sum(axis=1, keepdims=True)
    pc = p[np.arange(n), cls]
    ce = -np.log(np.maximum(pc, 1e-30))
    correct = (p.argmax(axis=1) == cls).astype(np.float64)
    d = np.zeros((n, OUT_PAD))
    onehot = np.eye(3)[cls]
    d[:, 0] = (p[:, 0] - onehot[:, 0]) / n
    d[:, 1] = (p[:, 1] - onehot[:, 1]) / n
    disk = cls == 1
    er = y[:, 2] - t[:, 1]
    lr = np.where(disk, w_r * er * er, 0.0)
    d[:, 2] = np.where(disk, 2 * w_r * er / n, 0.0)
    esc = cls == 2
    ldir = np.zeros(n)
    for c in range(3):
        ec = y[:, 3 + c] - t[:, 2 + c]
        ldir += np.where(esc, w_d * ec * ec, 0.0)
        d[:, 3 + c] = np.where(esc, 2 * w_d * ec / n, 0.0)
    sums = np.array([ce.sum(), lr.sum(), ldir.sum(), correct.sum()])
    return sums, d


def exact_forward(W, b, x):
    """Float64 forward without any bf16 rounding (for finite-difference checks)."""
    h = x.astype(np.float64)
    for l in range(len(W)):
        z = h @ W[l].astype(np.float64) + b[l].astype(np.float64)
        h = relu(z) if l < len(W) - 1 else z
    return h
