"""Mixed-precision exchange against the FP64 path on identical walker states.

BF16x3 (hi*hi + hi*lo + lo*hi, FP32 accumulation) keeps ~2^-16 relative
accuracy per element and FP32 accumulation; the local energy must agree with
FP64 to 5e-6 of |E_L| per walker (walkers near the trial).  The single-pass BF16 variant (tens of mHa per
walker) is only bounded loosely: it is not used for the exact subset.
"""
import numpy as np
import pytest

from test_gpu_parity import build, perturbed_walkers
from vmc_amd.abi import AFQ_PREC_BF16, AFQ_PREC_BF16X3, AFQ_PREC_FP64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["rhf_rhf", "uhf_uhf", "uhf_rhf", "c0_rhf", "uhf_33_31", "uhf_n60", "c1_small"])
def test_exchange_bf16x3_vs_fp64(name):
    W = 7
    h, wrhf, ctx, orc = build(name, W)
    rng = np.random.default_rng(11)
    dets = perturbed_walkers(h, W, wrhf, rng, 0.05)
    ctx.set_walkers(dets)
    e64, _ = ctx.measure()
    ctx.set_precision(prec_exchange=AFQ_PREC_BF16X3)
    e3, _ = ctx.measure()
    ctx.set_precision(prec_exchange=AFQ_PREC_BF16)
    e1, _ = ctx.measure()
    ctx.set_precision(prec_exchange=AFQ_PREC_FP64)
    e64b, _ = ctx.measure()
    assert np.array_equal(e64, e64b)
    print(name, "max |dE| bf16x3 %.2e  bf16 %.2e" % (np.max(np.abs(e3 - e64)), np.max(np.abs(e1 - e64))))
    assert np.max(np.abs(e3 - e64) / np.abs(e64)) < 5e-6
    assert np.max(np.abs(e1 - e64)) < 0.5
