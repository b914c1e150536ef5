"""The kernels' correctly rounded square root, checked exhaustively.

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
