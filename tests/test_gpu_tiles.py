"""Tile shapes of the scattering-density and multiple-scattering passes
(with the row order of the multiple-scattering blocks): every shape gives
bit-identical LUTs (so the autotuner's choice never
changes a result), the autotuner records a valid shape per size class, and
a batch whose jobs have different shapes (one launch per shape) equals each
atmosphere precomputed alone."""
import numpy as np
import pytest

import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu

NAMES = ("transmittance", "scattering", "single_mie_scattering", "irradiance")

# Size classes with half-empty default blocks (MU x NU = 64 x 4), the
# reference sizes, and rows of 512 texels (two texels per thread in
# 256-thread blocks by default).
SIZES = {
    "small": dict(transmittance_width=64, transmittance_height=16, scattering_r=8,
                  scattering_mu=64, scattering_mu_s=16, scattering_nu=4,
                  irradiance_width=32, irradiance_height=8),
    "reference": dict(configs.DEFAULT_SIZES),
    "wide": dict(transmittance_width=64, transmittance_height=16, scattering_r=8,
                 scattering_mu=32, scattering_mu_s=32, scattering_nu=16,
                 irradiance_width=32, irradiance_height=8),
}


@pytest.fixture(autouse=True)
def clean_table():
    atmo.reset_tiles()
    yield
    atmo.reset_tiles()


def run(cfg):
    kw, orders = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    m.init(orders)
    out = {n: m.lut(n) for n in NAMES}
    m.close()
    return out


def shapes_for(sizes):
    sd, ms = atmo.tile_candidates()
    d = atmo.get_tiles(sizes)
    dsd = (d["sd_nu_per_thread"], d["sd_block_threads"])
    dms = (d["ms_texels_per_thread"], d["ms_block_threads"])
    dmm = d["ms_row_order"]
    out = [(s, dms, dmm) for s in sd if s != dsd]
    w = sizes["scattering_nu"] * sizes["scattering_mu_s"]
    out += [(dsd, t, mm) for t in ms for mm in atmo.MS_ROW_ORDERS
            if (t, mm) != (dms, dmm) and w <= t[0] * t[1]]
    return out


@pytest.mark.parametrize("name", sorted(SIZES))
def test_every_shape_gives_identical_luts(name):
    sizes = SIZES[name]
    cfg = configs.earth(orders=4, lut_sizes=sizes)
    ref = run(cfg)
    for (chunk, block), (tpt, mblock), mm in shapes_for(sizes):
        print("case", name, chunk, block, tpt, mblock, mm, flush=True)
        atmo.set_tiles(sizes, chunk, block, tpt, mblock, ms_row_order=mm)
        got = run(cfg)
        for n in NAMES:
            assert np.array_equal(got[n], ref[n]), (name, chunk, block, tpt, mblock, mm, n)
        atmo.reset_tiles()


def test_every_shape_identical_illuminance_mode():
    # FP64 accumulators of the multiple-scattering epilogue (15 wavelengths).
    sizes = SIZES["small"]
    cfg = configs.spectral(orders=3, lut_sizes=sizes)
    ref = run(cfg)
    for (chunk, block), (tpt, mblock), mm in [((1, 256), (1, 256), "mu-major"),
                                               ((4, 64), (2, 256), "r-major"),
                                               ((2, 128), (2, 512), "ray")]:
        print("case spectral", chunk, block, tpt, mblock, mm, flush=True)
        atmo.set_tiles(sizes, chunk, block, tpt, mblock, ms_row_order=mm)
        got = run(cfg)
        for n in NAMES:
            assert np.array_equal(got[n], ref[n]), (chunk, block, tpt, n)


@pytest.mark.parametrize("use_graph", [True, False])
def test_autotune_then_mixed_shape_batch_equals_alone(use_graph):
    cfgs = configs.batch(23, 8)
    sizes = [c["lut_sizes"] for c in cfgs]
    alone = [run(c) for c in cfgs]  # default shapes
    atmo.autotune(sizes)
    sd, ms = atmo.tile_candidates()
    for z in sizes:
        t = atmo.get_tiles(z)
        assert t["sd_tuned"] and t["ms_tuned"]
        assert (t["sd_nu_per_thread"], t["sd_block_threads"]) in sd
        w = z["scattering_nu"] * z["scattering_mu_s"]
        assert (t["ms_texels_per_thread"], t["ms_block_threads"]) in ms
        assert w > 1024 or w <= t["ms_texels_per_thread"] * t["ms_block_threads"]
    # Force two shapes on the first two size classes so that the batch has
    # jobs of different shapes in one pass level whatever the tuner chose.
    atmo.set_tiles(sizes[0], 1, 256, 2, 512, ms_row_order="ray")
    atmo.set_tiles(sizes[1], 4, 64, 2, 512, ms_row_order="r-major")
    models = [atmo.Model(**configs.model_kwargs(c)[0]) for c in cfgs]
    b = atmo.Batch(models, [c["orders"] for c in cfgs], use_graph=use_graph)
    # The launches of one level that differ in tile shape run concurrently
    # on forked streams; replays give the same result.
    for _ in range(2):
        b.launch()
        atmo.synchronize()
        for i, (m, ref) in enumerate(zip(models, alone)):
            for n in NAMES:
                assert np.array_equal(m.lut(n), ref[n]), (i, n)
    b.close()
    for m in models:
        m.close()


def test_tuned_shape_used_by_init():
    sizes = SIZES["small"]
    atmo.autotune([sizes])
    t1 = atmo.get_tiles(sizes)
    atmo.autotune([sizes])  # already tuned: unchanged, no re-timing needed
    assert atmo.get_tiles(sizes) == t1
    cfg = configs.earth(orders=3, lut_sizes=sizes)
    got = run(cfg)
    atmo.reset_tiles()
    ref = run(cfg)
    for n in NAMES:
        assert np.array_equal(got[n], ref[n]), n
