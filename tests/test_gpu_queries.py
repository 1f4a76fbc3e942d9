"""Sky-shading lookups on the GPU vs the FP64 oracle (functions.glsl:1633-1896).

Two comparisons per lookup:
  - lookup kernels alone: the oracle evaluates the lookup in FP64 on the GPU's
    own LUTs (decoded to float64), so only the lookup arithmetic differs;
  - end to end: the oracle's lookup on the oracle's LUTs.
Rays are seeded (camera altitude 0-100 km, view direction uniform on the
sphere, sun zenith 0-100 deg, hit distance 0-50 km) plus hand-picked edge cases
(camera in space looking away/at the atmosphere, horizon rays, light shafts).

Tolerance: 1e-3 relative error per ray and channel with a floor of 1e-4 of the
batch's largest value in that channel. GetSkyRadianceToPoint subtracts two
nearly equal scatterings for short segments (functions.glsl:1849-1851), which
amplifies LUT differences, so its end-to-end error is taken relative to the
size of the subtracted terms (the sky radiance along the same ray); its
lookup-kernel-alone comparison uses the common metric.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu

QFLOOR = 1e-4
TOL = 1e-3


def f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def edge_rays(rb, rt):
    """Hand-picked rays: in space looking away / through the atmosphere,
    grazing the horizon from the ground and from altitude, straight up/down,
    sun below the horizon."""
    cams, views, suns = [], [], []
    def add(cam, view, sun):
        v = np.array(view, float)
        s = np.array(sun, float)
        cams.append(cam)
        views.append(v / np.linalg.norm(v))
        suns.append(s / np.linalg.norm(s))
    add([0, 0, rt + 100.0], [0, 0, 1], [0, 0.3, 1])        # space, looking away
    add([0, 0, rt + 100.0], [0, 0, -1], [0, 0.3, 1])       # space, looking down
    add([0, 0, rt + 100.0], [1, 0, -0.2], [0, 0.5, 1])     # space, grazing
    add([0, 0, rb + 0.001], [1, 0, 0], [0, 0, 1])          # ground, horizontal
    add([0, 0, rb + 0.001], [1, 0, 1e-3], [1, 0, 0.01])    # ground, just above horizon
    h = rb + 10.0
    mu_h = -np.sqrt(1 - (rb / h) ** 2)
    add([0, 0, h], [np.sqrt(1 - (mu_h + 1e-3) ** 2), 0, mu_h + 1e-3], [0, 1, 0.2])
    add([0, 0, h], [np.sqrt(1 - (mu_h - 1e-3) ** 2), 0, mu_h - 1e-3], [0, 1, 0.2])
    add([0, 0, h], [0, 0, 1], [0, 0, 1])
    add([0, 0, h], [0, 0, -1], [0.5, 0, -0.3])             # sun below horizon
    return (np.array(cams, np.float32), np.array(views, np.float32),
            np.array(suns, np.float32))


@pytest.fixture(scope="module")
def setup():
    cfg = configs.earth(lut_sizes=parity.MID_SIZES)
    kw, orders = configs.model_kwargs(cfg)
    sizes = kw.pop("lut_sizes")
    m = atmo.Model(lut_sizes=sizes, **kw)
    m.init(orders)
    spec, s, _ = parity.oracle_spec(cfg)
    ref = ol.model_init(spec, s, orders)
    gpu = {k: f64(m.lut(k)) for k in ("transmittance", "scattering",
                                      "single_mie_scattering", "irradiance")}
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    rb, rt = kw["bottom_radius"] / 1000.0, kw["top_radius"] / 1000.0
    q = configs.sky_queries(99, 4096, rb, rt)
    cams, views, suns = edge_rays(rb, rt)
    n_e = len(cams)
    q["camera"] = np.concatenate([q["camera"], cams])
    q["view_ray"] = np.concatenate([q["view_ray"], views])
    q["sun_direction"] = np.concatenate([q["sun_direction"], suns])
    q["point"] = np.concatenate([q["point"], cams + views * 5.0])
    q["normal"] = np.concatenate([q["normal"], cams / np.linalg.norm(cams, axis=1,
                                                                     keepdims=True)])
    q["shadow_length"] = np.concatenate([q["shadow_length"], np.zeros(n_e, np.float32)])
    yield m, ref, gpu, a, s, q
    m.close()


def oracle_sky(a, s, luts, q, shadow=None):
    n = len(q["camera"])
    r, t = np.zeros((n, 3)), np.zeros((n, 3))
    sh = f64(q["shadow_length"] if shadow is None else shadow)
    ol.lib().or_sky_radiance(C.byref(a), s, ol.ptr(luts["transmittance"]),
                             ol.ptr(luts["scattering"]),
                             ol.ptr(luts["single_mie_scattering"]), 0, n,
                             ol.ptr(f64(q["camera"])), ol.ptr(f64(q["view_ray"])), ol.ptr(sh),
                             ol.ptr(f64(q["sun_direction"])), ol.ptr(r), ol.ptr(t))
    return r, t


def oracle_to_point(a, s, luts, q, shadow=None):
    n = len(q["camera"])
    r, t = np.zeros((n, 3)), np.zeros((n, 3))
    sh = f64(q["shadow_length"] if shadow is None else shadow)
    ol.lib().or_sky_radiance_to_point(C.byref(a), s, ol.ptr(luts["transmittance"]),
                                      ol.ptr(luts["scattering"]),
                                      ol.ptr(luts["single_mie_scattering"]), 0, n,
                                      ol.ptr(f64(q["camera"])), ol.ptr(f64(q["point"])),
                                      ol.ptr(sh), ol.ptr(f64(q["sun_direction"])),
                                      ol.ptr(r), ol.ptr(t))
    return r, t


def oracle_irradiance(a, s, luts, q):
    n = len(q["camera"])
    sun, sky = np.zeros((n, 3)), np.zeros((n, 3))
    ol.lib().or_sun_and_sky_irradiance(C.byref(a), s, ol.ptr(luts["transmittance"]),
                                       ol.ptr(luts["irradiance"]), n, ol.ptr(f64(q["point"])),
                                       ol.ptr(f64(q["normal"])),
                                       ol.ptr(f64(q["sun_direction"])), ol.ptr(sun),
                                       ol.ptr(sky))
    return sun, sky


def err(g, r):
    return parity.rel_err(g, r, floor=QFLOOR)[0]


def test_sky_radiance(setup):
    m, ref, gpu, a, s, q = setup
    rad, tr = m.sky_radiance(q["camera"], q["view_ray"], q["shadow_length"],
                             q["sun_direction"])
    for luts in (gpu, ref):
        r, t = oracle_sky(a, s, luts, q)
        assert err(rad, r) <= TOL and err(tr, t) <= TOL, (err(rad, r), err(tr, t))
    # Camera in space looking away: no scattering, full transmittance.
    i = len(q["camera"]) - 9
    assert np.all(rad[i] == 0.0) and np.all(tr[i] == 1.0)


def test_sky_radiance_light_shafts(setup):
    m, ref, gpu, a, s, q = setup
    shadow = np.linspace(0.0, 20.0, len(q["camera"])).astype(np.float32)
    rad, tr = m.sky_radiance(q["camera"], q["view_ray"], shadow, q["sun_direction"])
    r, t = oracle_sky(a, s, gpu, q, shadow)
    assert err(rad, r) <= TOL and err(tr, t) <= TOL, (err(rad, r), err(tr, t))


def test_sky_radiance_to_point(setup):
    m, ref, gpu, a, s, q = setup
    rad, tr = m.sky_radiance_to_point(q["camera"], q["point"], q["shadow_length"],
                                      q["sun_direction"])
    # Lookup kernel alone: same LUTs, strict metric.
    r, t = oracle_to_point(a, s, gpu, q)
    assert err(rad, r) <= TOL and err(tr, t) <= TOL, (err(rad, r), err(tr, t))
    # End to end: GetSkyRadianceToPoint is S(x) - T S(x_p) (functions.glsl:1849),
    # so LUT differences are measured against the size of the subtracted terms,
    # i.e. the sky radiance along the same ray (GetSkyRadiance of the ray).
    r, t = oracle_to_point(a, s, ref, q)
    view = f64(q["point"]) - f64(q["camera"])
    view /= np.linalg.norm(view, axis=1, keepdims=True)
    qv = dict(q, view_ray=view)
    sky, _ = oracle_sky(a, s, ref, qv)
    e2e = np.max(np.abs(rad - r) / (np.abs(r) + np.abs(sky) + 1e-30))
    assert e2e <= TOL and err(tr, t) <= TOL, (e2e, err(tr, t))
    shadow = np.linspace(0.0, 10.0, len(q["camera"])).astype(np.float32)
    rad, tr = m.sky_radiance_to_point(q["camera"], q["point"], shadow, q["sun_direction"])
    r, t = oracle_to_point(a, s, gpu, q, shadow)
    assert err(rad, r) <= TOL and err(tr, t) <= TOL, (err(rad, r), err(tr, t))


def test_sun_and_sky_irradiance(setup):
    m, ref, gpu, a, s, q = setup
    sun, sky = m.sun_and_sky_irradiance(q["point"], q["normal"], q["sun_direction"])
    for luts in (gpu, ref):
        rs, rk = oracle_irradiance(a, s, luts, q)
        assert err(sun, rs) <= TOL and err(sky, rk) <= TOL, (err(sun, rs), err(sky, rk))


def test_solar_radiance(setup):
    m = setup[0]
    # GetSolarRadiance of the GL API: solar / (pi alpha^2) (model.cc:227-230).
    expected = 1.5 / (np.pi * 0.00935 ** 2)
    assert np.allclose(m.solar_radiance(), expected, rtol=1e-6)


def test_lookup_before_init_is_an_error():
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES)
    kw, _ = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    with pytest.raises(atmo.AtmoError):
        m.sky_radiance(np.zeros((1, 3), np.float32) + [0, 0, 6400],
                       np.array([[0, 0, 1]], np.float32), np.zeros(1, np.float32),
                       np.array([[0, 0, 1]], np.float32))
    m.close()


def test_luminance_api(setup):
    """GetSkyLuminance & co. are the radiance lookups times the
    SKY/SUN_SPECTRAL_RADIANCE_TO_LUMINANCE factors (model.cc:252-281), whose
    values follow ComputeSpectralRadianceToLuminanceFactors (model.cc:562-595)."""
    m, ref, gpu, a, s, q = setup
    sky_k, sun_k = m.luminance_factors()
    p = m.params
    ok = np.zeros(3)
    ol.lib().or_luminance_factors(len(p.wavelengths), ol.ptr(p.wavelengths),
                                  ol.ptr(p.solar_irradiance), -3.0, ol.ptr(ok))
    assert np.allclose(sky_k, ok, rtol=1e-12)
    ol.lib().or_luminance_factors(len(p.wavelengths), ol.ptr(p.wavelengths),
                                  ol.ptr(p.solar_irradiance), 0.0, ol.ptr(ok))
    assert np.allclose(sun_k, ok, rtol=1e-12)
    rad, tr = m.sky_radiance(q["camera"], q["view_ray"], q["shadow_length"],
                             q["sun_direction"])
    lum, tr2 = m.sky_luminance(q["camera"], q["view_ray"], q["shadow_length"],
                               q["sun_direction"])
    assert np.allclose(lum, rad * sky_k, rtol=1e-6, atol=0) and np.array_equal(tr, tr2)
    sun, sky = m.sun_and_sky_irradiance(q["point"], q["normal"], q["sun_direction"])
    sun_l, sky_l = m.sun_and_sky_illuminance(q["point"], q["normal"], q["sun_direction"])
    assert np.allclose(sun_l, sun * sun_k, rtol=1e-6) and np.allclose(sky_l, sky * sky_k,
                                                                      rtol=1e-6)
    assert np.allclose(m.solar_luminance(), m.solar_radiance() * sun_k, rtol=1e-6)


def test_illuminance_mode_has_no_radiance_api():
    cfg = configs.spectral(lut_sizes=parity.SMALL_SIZES)
    kw, _ = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    m.init(2)
    one = np.array([[0, 0, 6400.0]], np.float32)
    up = np.array([[0, 0, 1.0]], np.float32)
    with pytest.raises(atmo.AtmoError):
        m.sky_radiance(one, up, np.zeros(1, np.float32), up)
    lum, tr = m.sky_luminance(one, up, np.zeros(1, np.float32), up)
    assert np.all(np.isfinite(lum)) and np.all(lum > 0)
    m.close()


def test_illuminance_mode_lookups_match_oracle():
    """The lookups of the spectral workload: GetSkyLuminance,
    GetSkyLuminanceToPoint and GetSunAndSkyIlluminance on a 15-wavelength
    model, whose LUTs hold luminance. Each equals the radiance lookup
    arithmetic on those LUTs at 680/550/440 nm times the luminance factors
    (MAX_LUMINOUS_EFFICACY for the sky in this mode, model.cc:735-742);
    the oracle evaluates that arithmetic in FP64 on the GPU's own LUTs."""
    cfg = configs.spectral(lut_sizes=parity.MID_SIZES)
    kw, _ = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    m.init(3)
    spec, s, _ = parity.oracle_spec(cfg)
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    luts = {k: f64(m.lut(k)) for k in ("transmittance", "scattering",
                                       "single_mie_scattering", "irradiance")}
    rb, rt = kw["bottom_radius"] / 1000.0, kw["top_radius"] / 1000.0
    q = configs.sky_queries(17, 4096, rb, rt)
    sky_k, sun_k = (np.asarray(k) for k in m.luminance_factors())
    lum, tr = m.sky_luminance(q["camera"], q["view_ray"], q["shadow_length"],
                              q["sun_direction"])
    r, t = oracle_sky(a, s, luts, q)
    e1 = (err(lum, r * sky_k), err(tr, t))
    lum, tr = m.sky_luminance_to_point(q["camera"], q["point"], q["shadow_length"],
                                       q["sun_direction"])
    r, t = oracle_to_point(a, s, luts, q)
    e2 = (err(lum, r * sky_k), err(tr, t))
    sun, sky = m.sun_and_sky_illuminance(q["point"], q["normal"], q["sun_direction"])
    rs, rk = oracle_irradiance(a, s, luts, q)
    e3 = (err(sun, rs * sun_k), err(sky, rk * sky_k))
    m.close()
    print("illuminance-mode lookups: sky %.2e/%.2e, to point %.2e/%.2e, sun and sky "
          "%.2e/%.2e" % (e1 + e2 + e3), flush=True)
    for e in e1 + e2 + e3:
        assert e <= TOL


def test_combined_textures_lookups():
    """COMBINED_SCATTERING_TEXTURES lookups extrapolate the single Mie
    scattering from the alpha channel (functions.glsl:1633-1645)."""
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES)
    cfg["combine_scattering_textures"] = True
    kw, orders = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    m.init(orders)
    spec, s, _ = parity.oracle_spec(cfg)
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    T = f64(m.lut("transmittance"))
    S = f64(m.lut("scattering"))
    q = configs.sky_queries(5, 1024, 6360.0, 6420.0)
    rad, tr = m.sky_radiance(q["camera"], q["view_ray"], q["shadow_length"],
                             q["sun_direction"])
    n = len(q["camera"])
    r, t = np.zeros((n, 3)), np.zeros((n, 3))
    ol.lib().or_sky_radiance(C.byref(a), s, ol.ptr(T), ol.ptr(S), None, 1, n,
                             ol.ptr(f64(q["camera"])), ol.ptr(f64(q["view_ray"])),
                             ol.ptr(f64(q["shadow_length"])), ol.ptr(f64(q["sun_direction"])),
                             ol.ptr(r), ol.ptr(t))
    assert err(rad, r) <= TOL and err(tr, t) <= TOL
    rad, tr = m.sky_radiance_to_point(q["camera"], q["point"], q["shadow_length"],
                                      q["sun_direction"])
    ol.lib().or_sky_radiance_to_point(C.byref(a), s, ol.ptr(T), ol.ptr(S), None, 1, n,
                                      ol.ptr(f64(q["camera"])), ol.ptr(f64(q["point"])),
                                      ol.ptr(f64(q["shadow_length"])),
                                      ol.ptr(f64(q["sun_direction"])), ol.ptr(r), ol.ptr(t))
    assert err(rad, r) <= TOL and err(tr, t) <= TOL
    m.close()


def test_query_generator_matches_host_twin():
    n = 50000
    dev = atmo.generate_sky_queries(31, n, 6360.0, 6420.0)
    host = configs.hashed_sky_queries(31, n, 6360.0, 6420.0)
    for k, v in dev.items():
        g = v.to_host().reshape(host[k].shape)
        # FP64 on both sides, rounded to FP32; sin/cos/sqrt may differ in the
        # last bit of FP64 before that rounding.
        assert np.allclose(g, host[k], rtol=2e-7, atol=2e-7), k


@pytest.mark.parametrize("half", [False, True])
def test_interleaved_lookups_equal_separate(half):
    """The interleaved lookup copy of S and S_Mie (one fetch per tap) gives
    bit-identical lookups to reading the two LUTs, in FP32 and FP16 storage,
    for every lookup; turning the option off and on again on an initialized
    model rebuilds the copy."""
    cfg = configs.earth(half_precision=half, lut_sizes=dict(
        transmittance_width=64, transmittance_height=16, scattering_r=8, scattering_mu=32,
        scattering_mu_s=8, scattering_nu=8, irradiance_width=16, irradiance_height=4))
    kw, orders = configs.model_kwargs(cfg)
    rb, rt = kw["bottom_radius"] / 1000.0, kw["top_radius"] / 1000.0
    q = configs.sky_queries(11, 4096, rb, rt)
    shadow = np.where(np.arange(4096) % 3 == 0, 2.0, 0.0).astype(np.float32)
    def lookups(m):
        return (m.sky_radiance(q["camera"], q["view_ray"], shadow, q["sun_direction"]) +
                m.sky_radiance_to_point(q["camera"], q["point"], shadow, q["sun_direction"]) +
                m.sun_and_sky_irradiance(q["point"], q["normal"], q["sun_direction"]))
    a = atmo.Model(**kw)  # interleaved copy (the default)
    a.init(orders)
    b = atmo.Model(interleaved_lookups=False, **kw)
    b.init(orders)
    ra, rb_ = lookups(a), lookups(b)
    for x, y in zip(ra, rb_):
        assert np.isfinite(x).all() and np.array_equal(x, y)
    assert a.footprint()[0] > b.footprint()[0]
    a.set_interleaved_lookups(False)
    a.set_interleaved_lookups(True)
    for x, y in zip(lookups(a), rb_):
        assert np.array_equal(x, y)
    a.close()
    b.close()
