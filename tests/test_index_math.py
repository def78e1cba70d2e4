"""The contraction kernels divide their loader and store indices by block sizes
with one multiply-high by a per-product reciprocal (FDiv in
mbsymmetry_amd/csrc/contract.hip): q(x) = floor(x * m / 2^32) with
m = floor((2^32 - 1) / d) + 1 for d > 1.  The kernels rely on q(x) == x // d
for every index below 2^17 and every divisor below 2^13.

q is non-decreasing in x and x // d only steps at multiples of d, so checking
q(k d - 1) == k - 1 and q(k d) == k at every multiple k d in range covers every
x in range."""
import numpy as np

X_LIMIT = 1 << 17
D_LIMIT = 1 << 13


def _mulhi_div(x, d):
    m = np.uint64(0xFFFFFFFF // d + 1)
    return (x.astype(np.uint64) * m) >> np.uint64(32)


def test_reciprocal_division_exact_over_kernel_ranges():
    for d in range(2, D_LIMIT):
        k = np.arange(1, (X_LIMIT - 1) // d + 1, dtype=np.uint64)
        x_hi = k * np.uint64(d)
        x_lo = x_hi - np.uint64(1)
        assert np.array_equal(_mulhi_div(x_hi, d), k), d
        assert np.array_equal(_mulhi_div(x_lo, d), k - np.uint64(1)), d
        top = np.array([X_LIMIT - 1], dtype=np.uint64)
        assert _mulhi_div(top, d)[0] == (X_LIMIT - 1) // d, d


def test_reciprocal_division_fails_outside_the_ranges():
    """The range matters, so the check above is not vacuous: for d = 3,
    m d = 2^32 + 2 and q(x) = floor(x / 3 + 2 x / (3 * 2^32)), which rounds up
    at x = 2^31 (x = 2 mod 3, x d > 2^32)."""
    x = np.array([1 << 31], dtype=np.uint64)
    assert _mulhi_div(x, 3)[0] == (1 << 31) // 3 + 1
    # and the largest index the kernels use is still exact for d = 3
    assert _mulhi_div(np.array([X_LIMIT - 1], dtype=np.uint64), 3)[0] == (X_LIMIT - 1) // 3
