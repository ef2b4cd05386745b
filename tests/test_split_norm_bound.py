"""Error bound of the mixed path's FP32 error norm (rms_norm_split in
magik_amd/csrc/mgk_kernels.hip).  The kernel forms the eight FP64 error
components and scales, takes their ratios in FP32 and re-takes the norm in
FP64 when the FP32 value lies within 2^-14 of 1, so its accept/reject
decision is the FP64 norm's whenever the FP32 form is within 2^-14 of the
FP64 one.  Here the FP32 form is restated in numpy float32 arithmetic
(correctly rounded reciprocal; the hardware one adds at most 0.5 u more) on
random components spread over many decades and on components whose norm
lies next to 1, and its relative distance from the FP64 norm must stay below
the 5.5 u = 3.3e-7 the kernel's comment derives, far inside the band."""
import numpy as np

U = 2.0 ** -24
BAND = 2.0 ** -14


def norm64(e, y0, y1, at, rt):
    sc = at + rt * np.maximum(np.abs(y0), np.abs(y1))
    return np.sqrt(np.mean((e / sc) ** 2, axis=-1))


def norm32(e, y0, y1, at, rt):
    sc = (at + rt * np.maximum(np.abs(y0), np.abs(y1))).astype(np.float32)
    q = e.astype(np.float32) * (np.float32(1.0) / sc)
    total = np.zeros(e.shape[0], np.float32)
    for k in range(8):   # left to right, as the kernel sums
        total = total + q[:, k] * q[:, k]
    return np.sqrt(total * np.float32(0.125)).astype(np.float64)


def cases(rng, n):
    at, rt = 1e-9, 1e-6
    y0 = rng.normal(size=(n, 8)) * 10.0 ** rng.uniform(-3, 3, (n, 8))
    y1 = y0 * (1 + 1e-3 * rng.normal(size=(n, 8)))
    e = rng.normal(size=(n, 8)) * (at + rt * np.abs(y0)) * 10.0 ** rng.uniform(-2, 1, (n, 1))
    return e, y0, y1, at, rt


def test_fp32_norm_is_within_its_bound_of_the_fp64_norm():
    rng = np.random.default_rng(3)
    e, y0, y1, at, rt = cases(rng, 400_000)
    n64, n32 = norm64(e, y0, y1, at, rt), norm32(e, y0, y1, at, rt)
    rel = np.abs(n32 / n64 - 1.0)
    assert rel.max() <= 5.5 * U, rel.max() / U


def test_trials_next_to_the_boundary_are_inside_the_band():
    """Components scaled so that the FP64 norm is 1 within 1e-6: the FP32
    form lands within the band, where the kernel takes the FP64 norm, and a
    decision made on the FP32 value outside the band agrees with FP64."""
    rng = np.random.default_rng(4)
    e, y0, y1, at, rt = cases(rng, 200_000)
    n64 = norm64(e, y0, y1, at, rt)
    target = 1.0 + rng.uniform(-1e-6, 1e-6, n64.shape)
    e = e * (target / n64)[:, None]
    n64, n32 = norm64(e, y0, y1, at, rt), norm32(e, y0, y1, at, rt)
    assert np.all(np.abs(n32 - 1.0) < BAND)
    # away from 1 the two decisions agree
    e2, y02, y12, _, _ = cases(rng, 200_000)
    a64, a32 = norm64(e2, y02, y12, at, rt), norm32(e2, y02, y12, at, rt)
    out = np.abs(a32 - 1.0) >= BAND
    assert out.mean() > 0.99
    np.testing.assert_array_equal((a32 < 1.0)[out], (a64 < 1.0)[out])
