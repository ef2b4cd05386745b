"""The reference's wavelength sampler (MGK_LAMBDA_CMF): the inverse CDF of
x+y+z with a floor of 5% of its mean (host_color_utilities.cuh:188-233,
color_pipeline.cuh:36), drawn by linear interpolation in 256 entries with
pdf = 1 / (255 (c[k+1] - c[k])) (device_sensor_utilities.cuh:25-33) and
throughput 1 / pdf (kernel_generate_thin_lens_ray.cu:73-74).

The table is pinned by what the mathematics fixes, not by a second copy of its
code: an independent float64 construction (trapezoid CDF, inverted by
interpolation), the floor as a lower bound of the pdf, the pdf's unit integral
(E[1/pdf] = range), the histogram of drawn wavelengths against the floored CMF
density, and an unbiased Monte Carlo estimate of a Planck integral.
"""
import numpy as np
import pytest

import magik_amd as M
from oracle import oracle as O

RANGES = [(380.0, 780.0), (360.0, 830.0), (400.0, 700.0)]


def table(lo, hi):
    return M.cmf_inverse_cdf(M.spectral(hero_min=lo, hero_max=hi, lambda_sampling=M.LAMBDA_CMF)).astype(np.float64)


def density(lo, hi, lam):
    """The floored CMF density (x+y+z + 0.05 mean) / integral, in float64 (the
    CMF linear between the 1 nm table points, as the sampler resamples it)."""
    cie = O.cie_table().astype(np.float64)
    grid = 360.0 + np.arange(cie.shape[0])
    s_of = lambda w: np.interp(w, grid, cie.sum(axis=1))
    g = np.linspace(lo, hi, 471)
    floor = 0.05 * s_of(g).mean()
    fine = np.linspace(lo, hi, 200001)
    z = np.trapezoid(s_of(fine) + floor, fine)
    return (s_of(lam) + floor) / z, floor / z


def draw(inv, u):
    x = u * 255.0
    k = np.minimum(np.floor(x), 254).astype(int)
    f = x - k
    lam = inv[k] * (1 - f) + inv[k + 1] * f
    pdf = 1.0 / (255.0 * (inv[k + 1] - inv[k]))
    return lam, pdf


@pytest.mark.parametrize("lo,hi", RANGES)
def test_endpoints_monotone_and_oracle(lo, hi):
    inv = table(lo, hi)
    assert inv[0] == lo and abs(inv[-1] - hi) <= 1e-4
    assert np.all(np.diff(inv) > 0)          # the floor keeps every bin open
    np.testing.assert_array_equal(inv.astype(np.float32), O.cmf_inverse_cdf(lo, hi, O.cie_table()))


@pytest.mark.parametrize("lo,hi", RANGES)
def test_matches_float64_construction(lo, hi):
    # trapezoid CDF of (s + floor) on the 471-point grid, inverted by
    # interpolation at u = k / 255: a dropped floor, a wrong half-weight or an
    # off-by-one bin moves entries by whole nanometres
    cie = O.cie_table().astype(np.float64)
    s = np.interp(np.linspace(lo, hi, 471), 360.0 + np.arange(cie.shape[0]), cie.sum(axis=1))
    h = (hi - lo) / 470.0
    cdf = np.concatenate([[0.0], np.cumsum((0.5 * (s[:-1] + s[1:]) + 0.05 * s.mean()) * h)])
    cdf /= cdf[-1]
    want = np.interp(np.arange(256) / 255.0, cdf, np.linspace(lo, hi, 471))
    # (float32 rounding of the CDF moves the wide tail bins by a few 1e-3 nm)
    assert np.abs(table(lo, hi) - want).max() <= 1e-2


@pytest.mark.parametrize("lo,hi", RANGES)
def test_floor_bounds_the_pdf(lo, hi):
    inv = table(lo, hi)
    pdf = 1.0 / (255.0 * np.diff(inv))
    _, p_floor = density(lo, hi, np.array([lo]))
    # nowhere below the floor's share (up to the bin's linear interpolation)
    assert pdf.min() >= 0.97 * p_floor
    # and at the red end of a range reaching 780 nm, where the CMF has all but
    # vanished, the floor is all that is left
    if hi >= 780.0:
        assert pdf[-1] <= 1.5 * p_floor


@pytest.mark.parametrize("lo,hi", RANGES)
def test_pdf_integrates_to_one(lo, hi):
    # E_u[1 / pdf] = integral of dlambda = hi - lo (each bin: probability
    # 1/255, inverse pdf 255 (c[k+1] - c[k]))
    inv = table(lo, hi)
    u = (np.arange(1_000_000) + 0.5) / 1_000_000
    _, pdf = draw(inv, u)
    assert abs(np.mean(1.0 / pdf) / (hi - lo) - 1.0) <= 1e-5


def test_table_bins_carry_equal_probability():
    # each table bin [c[k], c[k+1]] holds 1/255 of the floored CMF density
    lo, hi = 380.0, 780.0
    inv = table(lo, hi)
    fine = np.linspace(lo, hi, 400001)
    d, _ = density(lo, hi, fine)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (d[1:] + d[:-1]) * np.diff(fine))])
    p = np.diff(np.interp(inv, fine, cdf))
    assert np.max(np.abs(p * 255.0 - 1.0)) <= 1e-2


def test_drawn_wavelengths_follow_the_reported_pdf():
    # the histogram of drawn lambda against the integral of the pdf the draw
    # reports (an estimator weighting by 1/pdf is unbiased only if they agree)
    lo, hi = 380.0, 780.0
    inv = table(lo, hi)
    rng = np.random.default_rng(7)
    lam, _ = draw(inv, rng.uniform(size=2_000_000))
    edges = np.linspace(lo, hi, 41)
    hist, _ = np.histogram(lam, edges)
    # the reported pdf is 1 / (255 (c[k+1] - c[k])) on bin k: its integral up
    # to x is (k + (x - c[k]) / (c[k+1] - c[k])) / 255
    cdf_at = np.interp(edges, inv, np.arange(256) / 255.0)
    want = np.diff(cdf_at)
    got = hist / len(lam)
    assert np.max(np.abs(got - want) / want) <= 0.05


def test_unbiased_for_a_planck_spectrum():
    lo, hi = 380.0, 780.0
    inv = table(lo, hi)
    T = 6500.0
    f = lambda lam_nm: np.vectorize(lambda l: O.lib().orc_planck(O.F64, l, T))(lam_nm)
    fine = np.linspace(lo, hi, 4001)
    exact = np.trapezoid(f(fine), fine)
    u = (np.arange(20000) + 0.5) / 20000     # stratified, like the hero rotation
    lam, pdf = draw(inv, u)
    est = np.mean(f(lam) / pdf)
    assert abs(est / exact - 1.0) <= 2e-3
