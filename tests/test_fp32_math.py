"""Accuracy of the FP32 arithmetic helpers of the step kernels (GPU).

The FP32 stencil kernels evaluate powers, exponentials and quotients with the
hardware reciprocal, base-2 logarithm and base-2 exponential instead of the
library routines (gpu-wm-next_amd/csrc/gwm_internal.hpp, q* functions).  Each
helper is evaluated through the diagnostic entry point gwm_fp32_math on
inputs spanning the ranges the kernels see, and compared with the float64
value of the same function of the same float32 input, in units of the last
place of the float32 result.  The bounds are the contract the kernels'
tolerance-based FP32 parity rests on.
"""
import numpy as np
import pytest

import gwm_amd as G

pytestmark = pytest.mark.gpu

KAPPA = np.float32(0.2854)
GAMMA = np.float32(1.0) / (np.float32(1.0) - KAPPA)


def ulp_err(got, ref):
    """|got - ref| in ulps of the float32 value nearest ref."""
    r32 = ref.astype(np.float32)
    ulp = np.abs(np.nextafter(r32, np.float32(np.inf)).astype(np.float64) - r32.astype(np.float64))
    return np.abs(got.astype(np.float64) - ref) / np.maximum(ulp, 1e-45)


@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0, G.F32)
    yield c
    c.close()


def signed_log_uniform(rng, n, lo, hi):
    mag = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    return (np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0) * mag).astype(np.float32)


CASES = [
    # fn, input sampler, float64 reference, max ulp, mean ulp
    (G.MATH_POW_EOS, lambda r, n: r.uniform(0.05, 2.0, n).astype(np.float32),
     lambda x: np.power(x, np.float64(GAMMA)), 3.0, 1.0),
    (G.MATH_POW_EXNER, lambda r, n: r.uniform(0.05, 1.2, n).astype(np.float32),
     lambda x: np.power(x, np.float64(KAPPA)), 3.0, 1.0),
    (G.MATH_EXP, lambda r, n: r.uniform(-20.0, 10.0, n).astype(np.float32), np.exp, 3.0, 1.0),
    (G.MATH_LOG1P, lambda r, n: signed_log_uniform(r, n, 1e-9, 0.5), np.log1p, 5.0, 1.0),
    (G.MATH_EXPM1, lambda r, n: signed_log_uniform(r, n, 1e-9, 0.7), np.expm1, 4.0, 1.0),
    (G.MATH_RCP, lambda r, n: np.exp(r.uniform(np.log(1e-6), np.log(1e6), n)).astype(np.float32),
     lambda x: 1.0 / x, 0.501, 0.26),
    # the pressure departure's (1 + x)^gamma - 1: binomial series below
    # |x| = 1/16, log1p then expm1 above (their errors compound)
    (G.MATH_POW1PM1, lambda r, n: signed_log_uniform(r, n, 1e-9, 0.0624),
     lambda x: np.expm1(np.float64(GAMMA) * np.log1p(x)), 2.0, 0.5),
    (G.MATH_POW1PM1, lambda r, n: signed_log_uniform(r, n, 0.0625, 0.3),
     lambda x: np.expm1(np.float64(GAMMA) * np.log1p(x)), 6.0, 1.5),
]


@pytest.mark.parametrize("fn,sample,ref,max_ulp,mean_ulp", CASES)
def test_fp32_helper_accuracy(ctx, fn, sample, ref, max_ulp, mean_ulp):
    rng = np.random.default_rng(fn + 1)
    x = sample(rng, 1 << 18)
    got = ctx.fp32_math(fn, x)
    e = ulp_err(got, ref(x.astype(np.float64)))
    print(f"fn {fn}: max {e.max():.2f} ulp, mean {e.mean():.3f} ulp")
    assert np.all(np.isfinite(got))
    assert e.max() <= max_ulp, (fn, e.max(), x[np.argmax(e)])
    assert e.mean() <= mean_ulp, (fn, e.mean())


def test_fp32_helpers_special_values(ctx):
    # log1p / expm1 are exact at 0 and keep tiny arguments (u == 1 branch)
    x = np.array([0.0, 1e-30, -1e-30, 1e-8, -1e-8], np.float32)
    assert np.array_equal(ctx.fp32_math(G.MATH_LOG1P, x), x)
    assert np.array_equal(ctx.fp32_math(G.MATH_EXPM1, x), x)
    # expm1 saturates at -1 for large negative arguments
    assert ctx.fp32_math(G.MATH_EXPM1, np.array([-40.0], np.float32))[0] == -1.0
    # x^gamma and x^kappa are 1 at 1
    one = np.ones(4, np.float32)
    assert np.all(ctx.fp32_math(G.MATH_POW_EOS, one) == 1.0)
    assert np.all(ctx.fp32_math(G.MATH_POW_EXNER, one) == 1.0)
    # (1 + x)^gamma - 1 is 0 at 0 and gamma x for tiny x, on both sides
    tiny = np.array([0.0, 1e-20, -1e-20], np.float32)
    assert np.array_equal(ctx.fp32_math(G.MATH_POW1PM1, tiny), GAMMA * tiny)
