"""FP16 LUT storage (half_precision): the scattering LUTs are accumulated in
FP32 and stored as FP16 with per-channel power-of-two scales; transmittance
and irradiance stay FP32 (as in the reference, atmosphere/model.cc:432-434).

Bars: FP16-stored textures <= 5e-3 max relative error vs the FP64 oracle,
FP32 textures <= 1e-3 (floor as in parity.FLOOR).
"""
import math

import numpy as np
import pytest

import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu

HALF = ("scattering", "single_mie_scattering")


def check(cfg, orders):
    cfg = dict(cfg)
    cfg["half_precision"] = True
    errs = parity.compare_init(cfg, orders)
    for name, e in errs.items():
        tol = parity.TOL_FP16 if name in HALF else parity.TOL_FP32
        assert e <= tol, (name, e)
    return errs


def test_earth_small_fp16():
    check(configs.earth(lut_sizes=parity.SMALL_SIZES), 4)


def test_spectral_small_fp16():
    check(configs.spectral(lut_sizes=parity.SMALL_SIZES), 3)


def test_fp16_views_and_scales():
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES, half_precision=True)
    kw, orders = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    m.init(orders)
    v = m.lut_view("scattering")
    assert v.dtype == 1  # ATMO_DTYPE_F16
    for c in range(4):
        sc = v.scale[c]
        assert sc > 0 and math.log2(sc) == int(math.log2(sc))
    res, scratch = m.footprint()
    m32 = atmo.Model(**configs.model_kwargs(configs.earth(lut_sizes=parity.SMALL_SIZES))[0])
    res32, _ = m32.footprint()
    assert res < res32
    # The device rule is the host one, on the FP32 sums it stores (the FP32
    # model computes the same bits).
    m32.init(orders)
    s32 = m32.lut("scattering")
    scale = atmo.lib().atmo_fp16_channel_scale
    for c in range(4):
        assert v.scale[c] == scale(float(np.abs(s32[..., c]).max())), c
    m.close()
    m32.close()


def test_fp16_batch_equals_alone():
    rng = configs.SplitMix64(99)
    cfgs = []
    for _ in range(3):
        c = configs.random_atmosphere(rng)
        c["lut_sizes"] = parity.SMALL_SIZES
        c["half_precision"] = True
        cfgs.append(c)
    alone = []
    for c in cfgs:
        kw, o = configs.model_kwargs(c)
        m = atmo.Model(**kw)
        m.init(o)
        alone.append(m.lut("scattering"))
        m.close()
    models = [atmo.Model(**configs.model_kwargs(c)[0]) for c in cfgs]
    atmo.init_batch(models, [c["orders"] for c in cfgs])
    for m, ref in zip(models, alone):
        assert np.array_equal(m.lut("scattering"), ref)
        m.close()


@pytest.mark.parametrize("spectral,half", [(True, False), (True, True), (False, True)])
def test_combined_textures_with_other_options(spectral, half):
    """COMBINED_SCATTERING_TEXTURES (the Mie red channel in the alpha of the
    scattering LUT, no single-Mie LUT) together with illuminance mode (FP64
    accumulators over 5 iterations) and with FP16 storage."""
    make = configs.spectral if spectral else configs.earth
    cfg = make(lut_sizes=parity.SMALL_SIZES, half_precision=half)
    cfg["combine_scattering_textures"] = True
    print("case", "spectral" if spectral else "earth", "fp16" if half else "fp32", flush=True)
    errs = parity.compare_init(cfg, 3)
    assert "single_mie_scattering" not in errs
    for name, e in errs.items():
        tol = parity.TOL_FP16 if half and name in HALF else parity.TOL_FP32
        assert e <= tol, (name, e)


def test_fp16_delta_j_21bit_pass_local():
    """FP16 models keep delta J as three 21-bit floats per texel (the FP32
    sign and exponent and 12 mantissa bits, ATMO_DTYPE_RGB21). Pass-local at
    the reference sizes (spectral atmosphere, last Precompute iteration): the
    density pass's output, decoded, against the oracle's density from the
    GPU's order-1 deltas at 2000 texels (rounding 2^-13 plus the FP32 pass
    error), and the multiple-scattering pass against the oracle's on that
    decoded delta J (FP32 bar: the decoded values enter the sums exactly)."""
    import ctypes as C
    import oracle_lib as ol
    cfg = configs.spectral(half_precision=True)
    spec, s, _ = parity.oracle_spec(cfg)
    kw, _ = configs.model_kwargs(cfg)
    m = atmo.Model(retain_intermediates=True, **kw)
    lambdas, _ = m.params.iteration(m.params.num_iterations() - 1)
    a = ol.lower(spec, lambdas)
    T = np.zeros(ol.transmittance_shape(s))
    ol.lib().or_pass_transmittance(C.byref(a), s, ol.ptr(T))
    n_s = s.scattering_r * s.scattering_mu * s.scattering_mu_s * s.scattering_nu
    rng = np.random.Generator(np.random.PCG64(17))
    idx = np.ascontiguousarray(rng.integers(0, n_s, 2000), dtype=np.int32)
    ip = idx.ctypes.data_as(C.POINTER(C.c_int))
    f64 = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    m.init(1)
    dR, dM, dE = (f64(m.lut(k)) for k in ("delta_rayleigh", "delta_mie", "delta_irradiance"))
    m.init(2)
    v = m.lut_view("delta_scattering_density")
    assert v.dtype == 2 and v.channels == 3  # ATMO_DTYPE_RGB21
    dJ = f64(m.lut("delta_scattering_density"))
    J = np.zeros((len(idx), 4))
    ol.lib().or_texels_scattering_density(C.byref(a), s, ol.ptr(T), ol.ptr(dR), ol.ptr(dM),
                                          ol.ptr(dR), ol.ptr(dE), 2, len(idx), ip, ol.ptr(J))
    e_sd = parity.rel_err(dJ.reshape(-1, 4)[idx], J)[0]
    ms = np.zeros((len(idx), 4))
    ol.lib().or_texels_multiple_scattering(C.byref(a), s, ol.ptr(T), ol.ptr(dJ), len(idx), ip,
                                           ol.ptr(ms))
    e_ms = parity.rel_err(m.lut("delta_rayleigh").reshape(-1, 4)[idx], ms)[0]
    print("FP16 model: 21-bit delta J vs oracle density %.2e, multiple "
          "scattering on it %.2e" % (e_sd, e_ms), flush=True)
    m.close()
    assert e_sd <= 2.0 ** -13 + 1e-4, e_sd
    assert e_ms <= parity.TOL_FP32, e_ms
