"""The kernels' correctly rounded square root and reciprocal, checked
exhaustively.

sph_sqrt (v_sqrt_f32 plus an FMA-residual choice among the three nearest
floats, sqrtf below 2^-96) must equal sqrtf(x) bit for bit for every float:
every pair term's bit-exactness rests on r = sqrt(r2).
"""
import pytest

import sph_amd as sa

pytestmark = pytest.mark.gpu

INF = 0x7F800000
SIGN = 0x80000000


def test_sqrt_equals_ieee_for_every_float():
    assert sa.selftest_sqrt(0, INF + 1) == 0                 # +0 .. +inf
    assert sa.selftest_sqrt(INF + 1, SIGN) == 0              # NaNs
    assert sa.selftest_sqrt(SIGN, 0xFFFFFFFF) == 0           # negative floats


def test_reciprocal_equals_ieee_division_for_every_float():
    """sph_rcp (v_rcp_f32, one FMA residual and one FMA correction) must
    equal 1.0f / x bit for bit over its domain (+-0, +-inf, NaN and
    2^-126 <= |x| <= 2^125, which holds every r = sqrt(r2) of a pair when
    the smoothing radius is at most 2^60): the pair terms' r_recip = 1/r
    rests on it."""
    assert sa.selftest_rcp(0, INF + 1) == 0                  # +0 .. +inf
    assert sa.selftest_rcp(INF + 1, SIGN) == 0               # NaNs
    assert sa.selftest_rcp(SIGN, 0xFFFFFFFF) == 0            # negative floats
