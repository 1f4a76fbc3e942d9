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


def test_spectral_illuminance_mode_small():
    """15 precomputed wavelengths: 5 Precompute iterations blended through the
    luminance_from_radiance matrices, then the transmittance recomputed at
    680/550/440 nm (atmosphere/model.cc:914-963)."""
    cfg = configs.spectral(lut_sizes=parity.SMALL_SIZES)
    errs = parity.compare_init(cfg, 3)
    for name, e in errs.items():
        assert e <= parity.TOL_FP32, (name, e)


def test_combined_scattering_textures_small():
    """COMBINED_SCATTERING_TEXTURES: no separate single-Mie LUT; the alpha of the
    scattering LUT carries the Mie red channel (atmosphere/model.cc:154-155)."""
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES)
    cfg["combine_scattering_textures"] = True
    errs = parity.compare_init(cfg, 3)
    assert "single_mie_scattering" not in errs
    for name, e in errs.items():
        assert e <= parity.TOL_FP32, (name, e)


def test_retained_intermediates_match_oracle_passes():
    """Pass-local check at the reference LUT sizes: after Init(n) the retained
    delta LUTs of order n are fed, in FP64, to the oracle's passes for a
    sample of texels (scattering density of order n from the GPU's order n-1
    inputs is checked indirectly through the multiple-scattering pass, which
    reads the GPU's delta J)."""
    import oracle_lib as ol
    import ctypes as C
    cfg = configs.earth()
    kw, _ = configs.model_kwargs(cfg)
    m = parity.gpu_model(cfg, retain_intermediates=True)
    spec, s, _ = parity.oracle_spec(cfg)
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    rng = np.random.Generator(np.random.PCG64(3))
    n_s = s.scattering_r * s.scattering_mu * s.scattering_mu_s * s.scattering_nu
    idx = np.ascontiguousarray(rng.integers(0, n_s, 1500), dtype=np.int32)
    ip = idx.ctypes.data_as(C.POINTER(C.c_int))
    f64 = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    # Order 1: single scattering from the GPU transmittance.
    m.init(1)
    T = f64(m.lut("transmittance"))
    ray, mie = np.zeros((len(idx), 4)), np.zeros((len(idx), 4))
    ol.lib().or_texels_single_scattering(C.byref(a), s, ol.ptr(T), len(idx), ip,
                                         ol.ptr(ray), ol.ptr(mie))
    g_r = m.lut("delta_rayleigh").reshape(-1, 4)[idx]
    g_m = m.lut("delta_mie").reshape(-1, 4)[idx]
    assert parity.rel_err(g_r, ray)[0] <= parity.TOL_FP32
    assert parity.rel_err(g_m, mie)[0] <= parity.TOL_FP32
    # Order 2: scattering density from the GPU's order-1 deltas, and the
    # multiple scattering from the GPU's delta J.
    dR, dM = f64(m.lut("delta_rayleigh")), f64(m.lut("delta_mie"))
    dE = f64(m.lut("delta_irradiance"))
    m.init(2)
    J = np.zeros((len(idx), 4))
    ol.lib().or_texels_scattering_density(C.byref(a), s, ol.ptr(T), ol.ptr(dR), ol.ptr(dM),
                                          ol.ptr(dR), ol.ptr(dE), 2, len(idx), ip, ol.ptr(J))
    g_j = m.lut("delta_scattering_density").reshape(-1, 4)[idx]
    assert parity.rel_err(g_j, J)[0] <= parity.TOL_FP32
    dJ = f64(m.lut("delta_scattering_density"))
    ms = np.zeros((len(idx), 4))
    ol.lib().or_texels_multiple_scattering(C.byref(a), s, ol.ptr(T), ol.ptr(dJ), len(idx), ip,
                                           ol.ptr(ms))
    g_ms = m.lut("delta_rayleigh").reshape(-1, 4)[idx]
    assert parity.rel_err(g_ms, ms)[0] <= parity.TOL_FP32
    m.close()


# Packed rows of NU x MU_S = 2048 texels: longer than the 1024 the LDS path
# of the multiple-scattering pass holds, so that pass takes its per-texel
# gather kernel.
LONG_ROW_SIZES = dict(transmittance_width=64, transmittance_height=16, scattering_r=4,
                      scattering_mu=16, scattering_mu_s=64, scattering_nu=32,
                      irradiance_width=16, irradiance_height=8)


def test_long_rows_multiple_scattering_every_row_order():
    """The gather kernel of the multiple-scattering pass, pass-local against
    the oracle on the GPU's own delta J, in every row order of its blocks
    (which must give bit-identical LUTs)."""
    import ctypes as C
    import oracle_lib as ol
    import precomputed_atmospheric_scattering_amd as atmo
    cfg = configs.earth(lut_sizes=LONG_ROW_SIZES)
    spec, s, _ = parity.oracle_spec(cfg)
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    n_s = s.scattering_r * s.scattering_mu * s.scattering_mu_s * s.scattering_nu
    rng = np.random.Generator(np.random.PCG64(5))
    idx = np.ascontiguousarray(rng.integers(0, n_s, 1500), dtype=np.int32)
    ip = idx.ctypes.data_as(C.POINTER(C.c_int))
    out = {}
    try:
        for mm in atmo.MS_ROW_ORDERS:
            print("case long rows, %s" % mm, flush=True)
            atmo.set_tiles(LONG_ROW_SIZES, 4, 256, 1, 256, ms_row_order=mm)
            m = parity.gpu_model(cfg, retain_intermediates=True)
            m.init(2)
            T = np.ascontiguousarray(m.lut("transmittance"), dtype=np.float64)
            dJ = np.ascontiguousarray(m.lut("delta_scattering_density"), dtype=np.float64)
            ms = np.zeros((len(idx), 4))
            ol.lib().or_texels_multiple_scattering(C.byref(a), s, ol.ptr(T), ol.ptr(dJ),
                                                   len(idx), ip, ol.ptr(ms))
            g = m.lut("delta_rayleigh").reshape(-1, 4)[idx]
            e = parity.rel_err(g, ms)[0]
            print("long rows, multiple scattering order 2: max rel err %.2e" % e, flush=True)
            assert e <= parity.TOL_FP32, (mm, e)
            m.init(4)
            out[mm] = {k: m.lut(k) for k in ("scattering", "single_mie_scattering",
                                             "irradiance")}
            m.close()
    finally:
        atmo.reset_tiles()
    for mm in atmo.MS_ROW_ORDERS[1:]:
        for k in out["r-major"]:
            assert np.array_equal(out["r-major"][k], out[mm][k]), (mm, k)


# The ends of the table sizes of the scattering-density and indirect-
# irradiance passes: 32 nu slices and a 512-texel irradiance row (the
# largest LDS tables), and every LUT dimension at its minimum (2; 4 for
# MU, two halves of 2).
EDGE_SIZES = {
    "largest tables": dict(transmittance_width=32, transmittance_height=8, scattering_r=2,
                           scattering_mu=8, scattering_mu_s=4, scattering_nu=32,
                           irradiance_width=512, irradiance_height=4),
    "smallest": dict(transmittance_width=2, transmittance_height=2, scattering_r=2,
                     scattering_mu=4, scattering_mu_s=2, scattering_nu=2,
                     irradiance_width=2, irradiance_height=2),
}


@pytest.mark.parametrize("name", sorted(EDGE_SIZES))
def test_table_size_limits(name):
    print("case", name, flush=True)
    cfg = configs.earth(lut_sizes=EDGE_SIZES[name])
    errs = parity.compare_init(cfg, 3)
    for tex, e in errs.items():
        assert e <= parity.TOL_FP32, (name, tex, e)
