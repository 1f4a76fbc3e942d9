"""GPU LUTs vs the FP64 oracle (full Model::Init chain), through the C ABI.

Tolerance: max relative error per texel and channel <= 1e-3 for FP32 LUTs,
with the near-zero floor of parity.FLOOR (1e-6 of the texture's largest value
in that channel).
"""
import numpy as np
import pytest

import parity
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("orders", [1, 2, 3, 4])
def test_earth_small_per_order(orders):
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES)
    errs = parity.compare_init(cfg, orders)
    for name, e in errs.items():
        assert e <= parity.TOL_FP32, (name, e)


def test_earth_mid_four_orders():
    cfg = configs.earth(lut_sizes=parity.MID_SIZES)
    errs = parity.compare_init(cfg, 4)
    for name, e in errs.items():
        assert e <= parity.TOL_FP32, (name, e)


def test_random_atmospheres_small():
    rng = configs.SplitMix64(1234)
    for _ in range(3):
        cfg = configs.random_atmosphere(rng)
        cfg["lut_sizes"] = parity.SMALL_SIZES
        errs = parity.compare_init(cfg, min(cfg["orders"], 4))
        for name, e in errs.items():
            assert e <= parity.TOL_FP32, (name, e)
