"""Error bound of the f32 closure's tanh formula, emulated on the host.

The closure's f32 network forms tanh(x) = sign(x) (1 - 2 / (e^(2|x|) + 1))
from the hardware exp and reciprocal, the last step as one fused
multiply-add (hg_ude.hpp, Mfma<float>::tanh_).  Here every f32 operation is
rounded as IEEE f32 (numpy), the exp additionally moved by up to 3 ulp (the
hardware exp works in base 2 from an f32 product with log2(e), and is not
correctly rounded) and the reciprocal by up to 1 ulp, and the result
compared with float64 tanh: the absolute error stays below 5e-7 everywhere
and below 1e-7 for correctly rounded steps away from the cancellation near
0.  Tanh's values lie in (-1, 1), where one f32 ulp
is up to 6e-8; the closure's outputs are checked against a float64 network to
f32 tolerance in the GPU tests.
"""
import numpy as np


def f32_tanh(x, exp_ulp=0, rcp_ulp=0):
    x = np.asarray(x, dtype=np.float32)
    e = np.exp(2.0 * np.abs(x.astype(np.float64))).astype(np.float32)
    e = (e * np.float32(1 + exp_ulp * 2.0**-23)).astype(np.float32)
    r = (np.float32(1) / (e + np.float32(1))).astype(np.float32)
    r = (r * np.float32(1 + rcp_ulp * 2.0**-23)).astype(np.float32)
    t = (1.0 - 2.0 * r.astype(np.float64)).astype(np.float32)  # fma(-2, r, 1): one rounding
    return np.copysign(t, x)


def test_f32_tanh_error_bound():
    x = np.linspace(-12.0, 12.0, 1_200_001).astype(np.float32)
    ref = np.tanh(x.astype(np.float64))
    worst = 0.0
    for pe in range(-3, 4):
        for pr in (-1, 0, 1):
            worst = max(worst, float(np.max(np.abs(f32_tanh(x, pe, pr).astype(np.float64) - ref))))
    assert worst < 5e-7
    far = np.abs(x) >= 0.5
    exact = np.abs(f32_tanh(x).astype(np.float64) - ref)
    assert float(np.max(exact[far])) < 1e-7
    # odd, saturating, exact at 0
    assert np.array_equal(f32_tanh(-x), -f32_tanh(x))
    assert f32_tanh(np.float32(0.0)) == 0.0
    assert f32_tanh(np.float32(12.0)) == 1.0


def test_f32_tanh_planted_mistake_is_caught():
    """The same formula without the factor 2 in the exponent (tanh(x/2) in
    effect) is off by far more than the bound."""
    x = np.linspace(-4.0, 4.0, 10_001).astype(np.float32)
    e = np.exp(np.abs(x.astype(np.float64))).astype(np.float32)
    wrong = np.copysign(1.0 - 2.0 / (e.astype(np.float64) + 1.0), x)
    assert np.max(np.abs(wrong - np.tanh(x.astype(np.float64)))) > 1e-2
