"""Batched precompute (one launch per pass over many atmospheres, HIP graph)
gives the same LUTs as precomputing each atmosphere alone, and matches the
oracle. Atmospheres differ in radii, coefficients, LUT sizes and scattering
orders, as in the batched workload."""
import numpy as np
import pytest

import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu

NAMES = ("transmittance", "scattering", "single_mie_scattering", "irradiance")


def small_batch(seed, count):
    rng = configs.SplitMix64(seed)
    out = []
    for _ in range(count):
        cfg = configs.random_atmosphere(rng)
        # Keep the oracle fast: quarter-size LUTs, still uneven across models.
        cfg["lut_sizes"] = {k: max(2, v // 4 + (v // 4) % 2 if k == "scattering_mu" else v // 4)
                            for k, v in cfg["lut_sizes"].items()}
        out.append(cfg)
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_alone(use_graph):
    cfgs = small_batch(5, 6)
    alone, batched = [], []
    for cfg in cfgs:
        kw, orders = configs.model_kwargs(cfg)
        m = atmo.Model(**kw)
        m.init(orders)
        alone.append({n: m.lut(n) for n in NAMES})
        m.close()
    models = [atmo.Model(**configs.model_kwargs(c)[0]) for c in cfgs]
    orders = [c["orders"] for c in cfgs]
    b = atmo.Batch(models, orders, use_graph=use_graph)
    assert b.num_launches >= 3
    for _ in range(2):  # replays give the same result
        b.launch()
        atmo.synchronize()
        for m, ref in zip(models, alone):
            for n in NAMES:
                assert np.array_equal(m.lut(n), ref[n]), n
    b.close()
    for m in models:
        m.close()


def test_batch_matches_oracle():
    cfgs = small_batch(17, 3)
    models = [atmo.Model(**configs.model_kwargs(c)[0]) for c in cfgs]
    atmo.init_batch(models, [c["orders"] for c in cfgs])
    for cfg, m in zip(cfgs, models):
        ref = parity.oracle_init(cfg)
        for n, r in ref.items():
            g = m.lut(n).reshape(r.shape)
            e = parity.rel_err(g, r, channels=4 if n == "scattering" else 3)[0]
            assert e <= parity.TOL_FP32, (n, e)
        m.close()


def test_full_size_batch_equals_alone():
    # The first 6 atmospheres of the 64-atmosphere batched workload (seed
    # 2024) at their drawn LUT sizes (0.5x..2x of the reference per
    # dimension) and orders: batched graph replay == each alone, bit for bit.
    cfgs = configs.batch(2024, 6)
    alone = []
    for cfg in cfgs:
        kw, orders = configs.model_kwargs(cfg)
        m = atmo.Model(**kw)
        m.init(orders)
        alone.append({n: m.lut(n) for n in ("scattering", "irradiance")})
        m.close()
    models = [atmo.Model(**configs.model_kwargs(c)[0]) for c in cfgs]
    atmo.init_batch(models, [c["orders"] for c in cfgs])
    for m, ref in zip(models, alone):
        for n, r in ref.items():
            assert np.array_equal(m.lut(n), r), n
        m.close()


def test_full_workload_batch_equals_alone():
    # All 64 atmospheres of the batched workload (seed 2024) at their drawn
    # sizes and orders, batched in one graph as the benchmark runs them:
    # every LUT of every model equals the model initialised alone, bit for
    # bit (compared by digest, so that the 64 models' LUTs need not be held
    # on the host at once).
    import xxhash

    def digest(m):
        h = xxhash.xxh3_128()
        for n in NAMES:
            h.update(np.ascontiguousarray(m.lut(n)))
        return h.hexdigest()

    cfgs = configs.batch(2024, 64)
    ref = []
    for cfg in cfgs:
        kw, orders = configs.model_kwargs(cfg)
        m = atmo.Model(**kw)
        m.init(orders)
        ref.append(digest(m))
        m.close()
    models = [atmo.Model(**configs.model_kwargs(c)[0]) for c in cfgs]
    b = atmo.Batch(models, [c["orders"] for c in cfgs], use_graph=True)
    b.launch()
    atmo.synchronize()
    got = [digest(m) for m in models]
    b.close()
    for m in models:
        m.close()
    assert got == ref, [i for i, (a, r) in enumerate(zip(got, ref)) if a != r]


def test_batch_holds_its_models():
    """While a batch lives, its models cannot join another batch, drop their
    interleaved lookup copy or be destroyed, and Init keeps the temporaries
    the batch's launches use; after the batch is destroyed all of that works
    again. The batch's result is unchanged by an Init in between."""
    cfgs = small_batch(23, 2)
    orders = [c["orders"] for c in cfgs]
    models = [atmo.Model(**configs.model_kwargs(c)[0]) for c in cfgs]
    b = atmo.Batch(models, orders)
    b.launch()
    atmo.synchronize()
    ref = [{n: m.lut(n) for n in NAMES} for m in models]
    with pytest.raises(atmo.AtmoError, match="live batch"):
        atmo.Batch([models[0]], [orders[0]])
    with pytest.raises(atmo.AtmoError, match="live batch"):
        models[1].set_interleaved_lookups(False)
    with pytest.raises(atmo.AtmoError, match="live batch"):
        models[0].close()
    assert models[0].handle is not None
    # Init of a batched model (no retained intermediates) must not free the
    # temporaries: the replay below writes them.
    models[0].init(orders[0])
    b.launch()
    atmo.synchronize()
    for m, r in zip(models, ref):
        for n in NAMES:
            assert np.array_equal(m.lut(n), r[n]), n
    b.close()
    models[1].set_interleaved_lookups(False)
    b2 = atmo.Batch([models[0]], [orders[0]])
    b2.close()
    for m in models:
        m.close()
