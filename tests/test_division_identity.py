"""The FP32 line-wavefront U sweep divides by a diagonal entry d as
(float)((double)y * rd) with rd = 1 / (double)d rounded to FP64 (ilu.hip,
k_band_gather).  That equals the correctly rounded FP32 quotient y / d for
every pair of floats: the FP64 product is within 2^-51 relative of y / d,
while a quotient of two floats that is not a float lies at least about 2^-49
relative away from every FP32 rounding midpoint (it cannot be a midpoint
itself).  Checked here on the CPU (numpy's float32 division is IEEE) over
random bit patterns, quotients of integers (the closest cases to midpoints),
and the special values."""
import numpy as np


def _mismatches(x, d):
    x = x.astype(np.float32)
    d = d.astype(np.float32)
    with np.errstate(all="ignore"):
        ref = x / d
        got = (x.astype(np.float64) * (1.0 / d.astype(np.float64))).astype(np.float32)
    same = (ref == got) | (np.isnan(ref) & np.isnan(got))
    same &= (np.signbit(ref) == np.signbit(got)) | np.isnan(ref)
    return int(np.count_nonzero(~same))


def test_random_bit_patterns():
    rng = np.random.default_rng(0)
    n = 4_000_000
    for _ in range(2):
        x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
        d = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
        assert _mismatches(x, d) == 0


def test_integer_quotients_and_factor_values():
    rng = np.random.default_rng(1)
    n = 4_000_000
    xi = rng.integers(1, 2**24, n).astype(np.float32)
    di = rng.integers(1, 2**24, n).astype(np.float32)
    assert _mismatches(xi, di) == 0
    # the sweeps' range: residual-sized numerators over diagonal factors
    x = rng.standard_normal(n).astype(np.float32)
    d = (rng.random(n) * 30 + 1e-3).astype(np.float32)
    assert _mismatches(x, d) == 0


def test_special_values():
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1.17549435e-38,
                   3.4028235e38, -3.4028235e38, 1.0, 3.0, 7.0, 0.1], np.float32)
    X, D = np.meshgrid(sp, sp)
    assert _mismatches(X.ravel(), D.ravel()) == 0


def test_a_wrong_rule_is_caught():
    """The check has teeth: a single-precision reciprocal product
    (float)(y * (1 / d)) in FP32 is not the quotient."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal(100_000).astype(np.float32)
    d = (rng.random(100_000) * 3 + 0.5).astype(np.float32)
    ref = x / d
    bad = x * (np.float32(1) / d)
    assert np.count_nonzero(ref != bad) > 1000
