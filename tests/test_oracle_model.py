"""Pins of the oracle's model-level code (atmosphere/model.cc restated in
oracle/oracle.cc): colour tables, luminance factors, illuminance iterations and
the full Init chain, by facts that hold independently of the implementation:
published values of the CIE and sRGB tables, identities between functions,
symmetries and linearity of the radiative transfer, and closed forms for
degenerate atmospheres. The reference holds no golden vectors for these
(its model tests need an OpenGL context); each check below is chosen so that a wrong term (a missing
blend, a wrong iteration wavelength, a transposed matrix, a dropped factor)
changes the result."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as ol
from precomputed_atmospheric_scattering_amd import configs

L = ol.lib
TINY = dict(transmittance_width=32, transmittance_height=8, scattering_r=4,
            scattering_mu=16, scattering_mu_s=4, scattering_nu=4,
            irradiance_width=8, irradiance_height=4)


def cie(l, col):
    return L().or_cie_table_value(l, col)


def test_cie_table_published_values():
    # CIE 1931 2-degree observer: y_bar(555 nm) = 1; the integrals of x_bar,
    # y_bar and z_bar are equal (equal-energy white has x = y = 1/3) and
    # close to 106.86 nm.
    assert cie(555.0, 2) == 1.0
    sums = [sum(cie(l, c) for l in range(361, 830)) for c in (1, 2, 3)]
    for s in sums:
        assert abs(s - 106.86) < 0.05, sums
    # Linear interpolation between the 5 nm samples.
    assert cie(557.5, 2) == pytest.approx(0.5 * (cie(555.0, 2) + cie(560.0, 2)))
    # Outside (360, 830) the reference returns 0 (model.cc:522-524).
    assert cie(360.0, 2) == 0.0 and cie(830.0, 1) == 0.0


# CIE 1931 2-degree colour-matching functions at single wavelengths, as the
# CIE publishes them (CIE 15:2004 tables, 5 nm steps): a check of the shared
# table against an outside source, value by value, so that a wrong or shifted
# entry fails even where the integrals above still hold.
CIE_1931_PUBLISHED = {
    380.0: (0.001368, 0.000039, 0.006450),
    400.0: (0.014310, 0.000396, 0.067850),
    420.0: (0.134380, 0.004000, 0.645600),
    440.0: (0.348280, 0.023000, 1.747060),
    450.0: (0.336200, 0.038000, 1.772110),
    460.0: (0.290800, 0.060000, 1.669200),
    480.0: (0.095640, 0.139020, 0.812950),
    500.0: (0.004900, 0.323000, 0.272000),
    520.0: (0.063270, 0.710000, 0.078250),
    540.0: (0.290400, 0.954000, 0.020300),
    560.0: (0.594500, 0.995000, 0.003900),
    580.0: (0.916300, 0.870000, 0.001650),
    600.0: (1.062200, 0.631000, 0.000800),
    620.0: (0.854450, 0.381000, 0.000190),
    640.0: (0.447900, 0.175000, 0.000020),
    650.0: (0.283500, 0.107000, 0.000000),
    700.0: (0.011359, 0.004102, 0.000000),
}


@pytest.mark.parametrize("lam", sorted(CIE_1931_PUBLISHED))
def test_cie_table_matches_published_samples(lam):
    for col, want in zip((1, 2, 3), CIE_1931_PUBLISHED[lam]):
        assert cie(lam, col) == pytest.approx(want, rel=1e-4, abs=1e-6), (lam, col)


def test_xyz_to_srgb_maps_d65_white_to_one():
    # sRGB (IEC 61966-2-1): D65 white XYZ = (0.9505, 1.0, 1.0890) -> (1, 1, 1).
    m = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415],
                  [0.0557, -0.2040, 1.0570]])
    assert np.allclose(m @ np.array([0.9505, 1.0, 1.0890]), 1.0, atol=2e-3)
    # The oracle's conversion uses that matrix on the 1 nm sums of the colour
    # matching functions: rgb(unit spectrum) = 683 * m @ sums.
    wl = np.array([400.0, 700.0])
    flat = np.ones(2)
    rgb = np.zeros(3)
    L().or_convert_spectrum_to_linear_srgb(2, ol.ptr(wl), ol.ptr(flat), ol.ptr(rgb))
    sums = np.array([sum(cie(l, c) for l in range(360, 830)) for c in (1, 2, 3)])
    assert np.allclose(rgb, 683.0 * m @ sums, rtol=1e-12)
    # Equal-energy white (X = Y = Z) in linear sRGB is m @ (1, 1, 1).
    assert np.allclose(rgb / rgb[1], (m @ np.ones(3)) / (m @ np.ones(3))[1], rtol=5e-3)


def test_luminance_factors_identity():
    # For a flat solar spectrum and lambda_power = 0 the factors of
    # ComputeSpectralRadianceToLuminanceFactors reduce to
    # ConvertSpectrumToLinearSrgb of a unit spectrum (model.cc:562-595 vs
    # :1020-1040).
    wl = np.arange(360.0, 831.0, 10.0)
    flat = np.full(len(wl), 1.7)
    k = np.zeros(3)
    L().or_luminance_factors(len(wl), ol.ptr(wl), ol.ptr(flat), 0.0, ol.ptr(k))
    rgb = np.zeros(3)
    one = np.ones(len(wl))
    L().or_convert_spectrum_to_linear_srgb(len(wl), ol.ptr(wl), ol.ptr(one), ol.ptr(rgb))
    assert np.allclose(k, rgb, rtol=1e-12)
    # lambda_power = -3 weights short wavelengths up for the blue channel.
    k3 = np.zeros(3)
    L().or_luminance_factors(len(wl), ol.ptr(wl), ol.ptr(flat), -3.0, ol.ptr(k3))
    assert k3[2] > k[2] * 0.9 and not np.allclose(k3, k)


def iterations(spec):
    n = L().or_num_iterations(C.byref(spec.c))
    out = []
    for i in range(n):
        lam, mat = (C.c_double * 3)(), (C.c_double * 9)()
        L().or_iteration(C.byref(spec.c), i, lam, mat)
        out.append((np.array(list(lam)), np.array(list(mat)).reshape(3, 3)))
    return out


def spectral_spec(**over):
    cfg = configs.spectral(lut_sizes=TINY)
    kw, orders = configs.model_kwargs(cfg)
    kw.pop("lut_sizes")
    kw.update(over)
    return ol.ModelSpec(**kw)


def test_illuminance_iterations():
    spec = spectral_spec()
    its = iterations(spec)
    assert len(its) == 5  # ceil(15 / 3) (model.cc:917)
    lams = np.concatenate([l for l, _ in its])
    # 15 midpoints of equal bins over [360, 830] (model.cc:918-924).
    assert np.allclose(lams, 360.0 + (np.arange(15) + 0.5) * 470.0 / 15.0)
    # The matrices integrate the colour matching functions: their sum over all
    # iterations is a 15-point midpoint rule of the 1 nm sums of
    # ConvertSpectrumToLinearSrgb / 683, within the rule's accuracy.
    total = sum(m.sum(axis=1) for _, m in its)
    wl = np.arange(360.0, 831.0, 10.0)
    rgb = np.zeros(3)
    L().or_convert_spectrum_to_linear_srgb(len(wl), ol.ptr(wl), ol.ptr(np.ones(len(wl))),
                                           ol.ptr(rgb))
    assert np.allclose(total, rgb / 683.0, rtol=0.03)
    # Each coefficient is a float value (model.cc:934 rounds them to float).
    for _, m in its:
        assert np.array_equal(m, m.astype(np.float32).astype(np.float64))
    # 3 wavelengths: identity matrix at 680/550/440 nm (model.cc:907-913).
    kw, _ = configs.model_kwargs(configs.earth())
    kw.pop("lut_sizes")
    (lam, m), = iterations(ol.ModelSpec(**kw))
    assert list(lam) == [680.0, 550.0, 440.0] and np.array_equal(m, np.eye(3))


def init(spec, orders, sizes=TINY):
    s = ol.sizes(**sizes)
    return ol.model_init(spec, s, orders)


def earth_spec(**over):
    kw, _ = configs.model_kwargs(configs.earth())
    kw.pop("lut_sizes")
    kw.update(over)
    return ol.ModelSpec(**kw)


def test_full_chain_ranges():
    out = init(earth_spec(), 4)
    T = out["transmittance"][..., :3]
    assert T.min() > 0.0 and T.max() <= 1.0
    for k in ("scattering", "single_mie_scattering", "irradiance"):
        assert out[k].min() >= 0.0 and np.isfinite(out[k]).all()
    # Vertical transmittance from the ground at 680 nm is a textbook value for
    # this atmosphere: exp(-(5.802e-6 * 8000 + 4.44e-6 * 1200)).
    assert T[0, 0, 0] == pytest.approx(math.exp(-(5.802e-6 * 8000 + 4.44e-6 * 1200)),
                                       rel=2e-3)


def test_linear_in_solar_irradiance():
    a = init(earth_spec(), 3)
    b = init(earth_spec(solar_irradiance=[3.0, 3.0, 3.0]), 3)
    assert np.array_equal(a["transmittance"], b["transmittance"])
    for k in ("scattering", "single_mie_scattering", "irradiance"):
        assert np.allclose(b[k], 2.0 * a[k], rtol=1e-12, atol=0)


def test_no_scattering_atmosphere():
    spec = earth_spec(rayleigh_scattering=[0.0] * 3, mie_scattering=[0.0] * 3,
                      mie_extinction=[0.0] * 3,
                      absorption_density=[(0.0, 0.0, 0.0, 0.0, 1.0)],
                      absorption_extinction=[2e-6, 4e-6, 8e-6])
    out = init(spec, 3)
    for k in ("scattering", "single_mie_scattering", "irradiance"):
        assert np.all(out[k] == 0.0), k
    # Uniform absorber: T = exp(-beta * distance to the top) at every texel.
    s = ol.sizes(**TINY)
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    T = out["transmittance"]
    for j in range(s.transmittance_height):
        for i in range(s.transmittance_width):
            r, mu = C.c_double(), C.c_double()
            L().or_rmu_from_transmittance_uv(C.byref(a), s, (i + 0.5) / s.transmittance_width,
                                             (j + 0.5) / s.transmittance_height,
                                             C.byref(r), C.byref(mu))
            d = L().or_distance_to_top(C.byref(a), r.value, mu.value)
            assert T[j, i, 0] == pytest.approx(math.exp(-a.absorption_extinction[0] * d),
                                               rel=1e-9)


def test_grey_atmosphere_channels_equal():
    spec = earth_spec(rayleigh_scattering=[1e-5] * 3)
    out = init(spec, 3)
    S = out["scattering"]
    assert np.array_equal(S[..., 0], S[..., 1]) and np.array_equal(S[..., 1], S[..., 2])
    # Alpha of the packed scattering texel is the Mie red channel
    # (model.cc:154-155): equal to the single-Mie LUT's red channel.
    assert np.array_equal(S[..., 3], out["single_mie_scattering"][..., 0])


def test_illuminance_blend_identity():
    # With wavelength-independent coefficients every iteration computes the
    # same radiance, so the blended luminance LUT is the radiance LUT times the
    # row sums of the summed luminance_from_radiance matrices.
    grey = dict(solar_irradiance=[1.5] * 47, rayleigh_scattering=[1e-5] * 47,
                mie_scattering=[4e-6] * 47, mie_extinction=[4.4e-6] * 47,
                absorption_extinction=[0.0] * 47)
    spec15 = spectral_spec(**grey)
    spec3 = spectral_spec(num_precomputed_wavelengths=3, **grey)
    lum = init(spec15, 2)
    rad = init(spec3, 2)
    rows = sum(m for _, m in iterations(spec15)).sum(axis=1)
    for k in ("scattering", "irradiance"):
        exp = rad[k][..., :3] * rows
        assert np.allclose(lum[k][..., :3], exp, rtol=1e-10, atol=1e-14), k
    assert np.allclose(lum["transmittance"], rad["transmittance"], rtol=0, atol=0)


def test_orders_add_decreasing_contributions():
    prev = None
    prev_inc = None
    for n in (1, 2, 3, 4):
        S = init(earth_spec(), n)["scattering"][..., :3]
        if prev is not None:
            inc = S - prev
            assert inc.min() >= -1e-15
            if prev_inc is not None:
                assert inc.sum() < prev_inc.sum()
            prev_inc = inc
        prev = S


def test_ground_albedo_adds_light():
    dark = init(earth_spec(ground_albedo=[0.0] * 3), 3)
    bright = init(earth_spec(ground_albedo=[0.3] * 3), 3)
    d = bright["scattering"][..., :3] - dark["scattering"][..., :3]
    assert d.min() >= -1e-15 and d.max() > 0
    # Single scattering does not see the ground.
    assert np.array_equal(bright["single_mie_scattering"], dark["single_mie_scattering"])


@pytest.mark.parametrize("make", [configs.earth, configs.spectral])
def test_per_order_snapshots_equal_separate_inits(make):
    # or_model_init_orders computes every iteration on its own and sums the
    # per-order states; that must equal Init(n) with blended iterations.
    cfg = make(lut_sizes=TINY)
    kw, _ = configs.model_kwargs(cfg)
    kw.pop("lut_sizes")
    spec = ol.ModelSpec(**kw)
    s = ol.sizes(**TINY)
    snaps = ol.model_init_orders(spec, s, 3)
    for n in (1, 2, 3):
        ref = ol.model_init(spec, s, n)
        for k, r in ref.items():
            scale = max(np.abs(r).max(), 1e-300)
            assert np.abs(snaps[n - 1][k] - r).max() <= 1e-12 * scale, (n, k)
