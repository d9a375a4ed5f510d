"""The lambda bound of the count pass, checked on the CPU.

The count pass settles k = 0 for a lane of cells from one upper bound on the
reference's lambda: with b >= 0, every bias model (common.h:414-431) is
non-decreasing in delta and monotone in b, so over delta <= delta_max and
b in [b_lo, b_hi] it is bounded by the larger of its values at
(delta_max, b_lo) and (delta_max, b_hi).  These tests check that property of
the oracle's FP64 bias_model on random samples, show that it fails for b < 0
(where the pass gives no bound), and check the 16-bit threshold rule
u < (hi + 1) / 2^16 <= exp(-lambda_max) => k = 0 against the inversion's
definition k = min{k : u <= F(k)}.
"""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def bias():
    from oracle_lib import load
    lib = load()
    f = lib.orc_bias_model
    return np.vectorize(lambda d, b, m: f(float(d), float(b), int(m)))


def _samples(rng, n, bmin=0.0):
    blo = rng.uniform(bmin, 3.0, n)
    bhi = blo + rng.uniform(0.0, 2.0, n)
    dmax = np.concatenate([rng.uniform(-0.999, 1.0, n // 2), rng.uniform(1.0, 50.0, n - n // 2)])
    # deltas at and below the maximum, down to -1 and below
    d = dmax - rng.exponential(0.5, n) * (1 + np.abs(dmax))
    d[: n // 10] = -1.0 - rng.uniform(0, 1, n // 10)
    b = blo + rng.uniform(0, 1, n) * (bhi - blo)
    return d, b, dmax, blo, bhi


@pytest.mark.parametrize("model", [0, 1, 2])
def test_corner_bound_holds_for_nonnegative_bias(bias, model):
    rng = np.random.default_rng(100 + model)
    d, b, dmax, blo, bhi = _samples(rng, 4000)
    val = bias(d, b, model)
    bound = np.maximum(bias(dmax, blo, model), bias(dmax, bhi, model))
    assert np.all(val <= bound * (1 + 1e-12) + 1e-300), np.max(val - bound)


def test_corner_bound_fails_for_negative_bias(bias):
    """The control: with b < 0, (1+delta)^b grows toward voids and the corner
    value at delta_max is no bound -- the pass must not use it there."""
    rng = np.random.default_rng(7)
    d, b, dmax, blo, bhi = _samples(rng, 4000, bmin=-3.0)
    neg = bhi < 0
    val = bias(d[neg], b[neg], 0)
    bound = np.maximum(bias(dmax[neg], blo[neg], 0), bias(dmax[neg], bhi[neg], 0))
    assert np.any(val > bound * 1.01)


def test_high_bits_threshold_gives_zero():
    """hi < floor(2^16 exp(-lam) (1 - margin)) implies u < exp(-lam), i.e. the
    inversion returns k = 0 for every completion of the low 16 bits."""
    rng = np.random.default_rng(11)
    lam = np.concatenate([rng.uniform(0, 16, 20000), [0.0, 1e-9, 15.999]])
    e32 = np.exp(-lam.astype(np.float32)).astype(np.float64)   # an FP32 exp, as __expf
    hcrit = np.floor(65536.0 * e32 * (1 - 2e-6 - 4e-7 * lam)).astype(np.int64)
    hi = np.maximum(hcrit - 1, 0)
    u_max = (hi * 65536.0 + 65535.0) / 2.0 ** 32
    ok = hcrit > 0
    assert np.all(u_max[ok] < np.exp(-lam[ok]))
