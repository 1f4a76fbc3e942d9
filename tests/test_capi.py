"""CPU checks of the C ABI boundary: the library loads, exports every entry
point include/atmo.h declares, validates parameters before touching the
device, and its host-side lowering agrees bit for bit with the oracle's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as ol
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import _capi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "atmo.h")).read()
    return sorted(set(re.findall(r"\b(atmo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    # The ctypes layer declares exactly the header's entry points.
    assert sorted(_capi.SIGNATURES) == names


def test_peaks_tool_loads():
    # The measurement library of bench.py (on-device rate ceilings) is built
    # beside libatmo.so and exports its two entry points; it rejects a NULL
    # output array before touching the device.
    lib = C.CDLL(os.path.join(ROOT, "precomputed_atmospheric_scattering_amd",
                              "libatmo_peaks.so"))
    lib.atmo_peaks_measure.restype = C.c_int
    lib.atmo_peaks_measure.argtypes = [C.POINTER(C.c_double), C.c_int]
    assert lib.atmo_peaks_measure(None, 7) != 0
    assert hasattr(lib, "atmo_peaks_flush")


def params(**over):
    kw, _ = configs.model_kwargs(configs.earth())
    kw.pop("lut_sizes")
    kw.update(over)
    return kw


def create_status(kw, sizes=None):
    p = atmo.ModelParameters(**kw)
    z = _capi.LutSizes(**(sizes or atmo.DEFAULT_SIZES))
    h = C.c_void_p()
    st = _capi.lib().atmo_model_create(C.byref(p.c), C.byref(z), C.byref(h))
    if st == 0:
        _capi.lib().atmo_model_destroy(h)
    return st, _capi.lib().atmo_last_error().decode()


@pytest.mark.parametrize("bad", [
    dict(bottom_radius=6420e3, top_radius=6360e3),
    dict(sun_angular_radius=0.0),
    dict(length_unit_in_meters=0.0),
    dict(mie_phase_function_g=1.0),
    dict(max_sun_zenith_angle=0.0),
    dict(wavelengths=[550.0, 440.0, 680.0]),
    dict(rayleigh_scattering=[1e-5, float("nan"), 1e-5]),
    dict(rayleigh_density=[(0, 1, -1 / 8e3, 0, 0)] * 3),
])
def test_invalid_parameters_rejected(bad):
    if "rayleigh_density" in bad:
        with pytest.raises(ValueError):
            atmo.ModelParameters(**params(**bad))
        return
    st, msg = create_status(params(**bad))
    assert st == 1, (st, msg)   # ATMO_ERROR_INVALID_ARGUMENT, before any device call
    assert msg


@pytest.mark.parametrize("sizes", [
    dict(scattering_mu=127), dict(scattering_mu=2), dict(scattering_nu=1),
    dict(scattering_nu=64),
    dict(irradiance_height=0), dict(irradiance_width=513),
])
def test_invalid_sizes_rejected(sizes):
    z = dict(atmo.DEFAULT_SIZES)
    z.update(sizes)
    st, msg = create_status(params(), z)
    assert st == 1, (st, msg)


def test_null_arguments():
    lib = _capi.lib()
    h = C.c_void_p()
    assert lib.atmo_model_create(None, None, C.byref(h)) == 1
    assert lib.atmo_model_init(None, 4, None) == 1
    assert lib.atmo_get_sky_radiance(None, 1, None, None, None, None, None, None, None) == 1
    lib.atmo_model_destroy(None)  # no-op


def test_server_null_arguments():
    # Argument checks run before any device call: no GPU needed.
    lib = _capi.lib()
    rid = C.c_uint64()
    n = C.c_int()
    assert lib.atmo_server_create(None, None) == 1
    assert lib.atmo_server_submit_precompute(None, None, None, 4, C.byref(rid)) == 1
    assert lib.atmo_server_submit_lookup(None, 1, 0, 0, 1, None, None, None, None, None,
                                         None, C.byref(rid)) == 1
    assert lib.atmo_server_poll(None, None, 0, C.byref(n), 0) == 1
    assert lib.atmo_server_release(None, 1) == 1
    assert lib.atmo_server_model(None, 1, None) == 1
    assert lib.atmo_server_get_stats(None, None) == 1
    lib.atmo_server_destroy(None)  # no-op


ATMO_FIELDS = ("solar_irradiance", "rayleigh_scattering", "mie_scattering",
               "mie_extinction", "absorption_extinction", "ground_albedo")


@pytest.mark.parametrize("cfg_name", ["earth", "spectral", "random"])
def test_lowering_matches_oracle_bit_exact(cfg_name):
    if cfg_name == "earth":
        cfg = configs.earth()
    elif cfg_name == "spectral":
        cfg = configs.spectral()
    else:
        cfg = configs.random_atmosphere(configs.SplitMix64(7))
    kw, _ = configs.model_kwargs(cfg)
    kw.pop("lut_sizes")
    mine = atmo.ModelParameters(**kw)
    spec = ol.ModelSpec(**kw)
    n_it = mine.num_iterations()
    assert n_it == ol.lib().or_num_iterations(C.byref(spec.c))
    for it in range(n_it):
        lam, mat = mine.iteration(it)
        lam_o = (C.c_double * 3)()
        mat_o = (C.c_double * 9)()
        ol.lib().or_iteration(C.byref(spec.c), it, lam_o, mat_o)
        assert lam == list(lam_o)
        # The GL program receives luminance_from_radiance as float (model.cc:934);
        # the oracle's coefficients are those float values.
        assert np.array_equal(mat.reshape(-1).astype(np.float64), np.array(list(mat_o)))
        a = mine.lower(lam)
        b = ol.lower(spec, lam)
        for f in ATMO_FIELDS:
            assert list(getattr(a, f)) == list(getattr(b, f)), f
        for f in ("sun_angular_radius", "bottom_radius", "top_radius",
                  "mie_phase_function_g", "mu_s_min"):
            assert getattr(a, f) == getattr(b, f), f
        for f in ("rayleigh_density", "mie_density", "absorption_density"):
            for k in range(2):
                la, lb = getattr(a, f)[k], getattr(b, f)[k]
                for g in ("width", "exp_term", "exp_scale", "linear_term", "constant_term"):
                    assert getattr(la, g) == getattr(lb, g), (f, k, g)


def test_profile_padding_at_front():
    kw = params(absorption_density=[(25000.0, 0, 0, 1 / 15000.0, -2 / 3.0)])
    a = atmo.ModelParameters(**kw).lower([680.0, 550.0, 440.0])
    assert a.absorption_density[0].width == 0.0 and a.absorption_density[0].linear_term == 0.0
    assert a.absorption_density[1].width == 25.0  # km: divided by the length unit
    assert a.absorption_density[1].linear_term == pytest.approx(1000.0 / 15000.0)


def test_convert_spectrum_matches_oracle():
    wl = np.arange(360.0, 831.0, 10.0)
    spectrum = 1.0 + 0.5 * np.sin(wl / 50.0)
    r, g, b = atmo.Model.ConvertSpectrumToLinearSrgb(wl, spectrum)
    o = (C.c_double * 3)()
    ol.lib().or_convert_spectrum_to_linear_srgb(len(wl), ol.ptr(wl), ol.ptr(spectrum), o)
    assert (r, g, b) == tuple(o)


def test_product_fails_loudly_without_library(tmp_path, monkeypatch):
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_capi, "_lib", None)
    with pytest.raises(ImportError):
        _capi.lib()


def test_counter_based_query_stream_matches_splitmix_sequence():
    # Ray i of the query generator takes outputs 6i+1 .. 6i+6 of the
    # sequential SplitMix64 stream: the counter form must agree with it.
    rng = configs.SplitMix64(77)
    seq = [rng.uniform() for _ in range(60)]
    assert list(configs.splitmix_uniforms(77, 1, 60)) == seq
    q = configs.hashed_sky_queries(77, 10, 6360.0, 6420.0)
    u = np.array(seq).reshape(10, 6)
    assert np.allclose(q["camera"][:, 2], 6360.0 + 100.0 * u[:, 0], rtol=1e-7)
    assert np.allclose(q["view_ray"][:, 2], 2.0 * u[:, 1] - 1.0, atol=1e-7)
    assert np.allclose(np.linalg.norm(q["view_ray"], axis=1), 1.0, atol=1e-6)
    r = np.linalg.norm(q["point"].astype(np.float64), axis=1)
    assert (r >= 6360.0 - 1e-3).all() and (r <= 6420.0).all()


def test_lut_file_null_arguments():
    lib = _capi.lib()
    assert lib.atmo_model_export_luts(None, b"/tmp") == 1
    assert lib.atmo_model_import_luts(None, b"/tmp") == 1
    assert lib.atmo_model_init_cached(None, 4, b"/tmp", None) == 1


def test_fp16_channel_scale_stays_finite():
    """The FP16 storage scale 2^min(14 - e, 126) (max = f 2^e, f in
    [0.5, 1)) maps a channel maximum into [2^13, 2^14) and stays finite, with
    a normal inverse, for tiny and denormal maxima."""
    import math
    f = _capi.lib().atmo_fp16_channel_scale
    for mx in (1.0, 0.75, 3.0e4, 1.5e-3, 7.0e-20, 2.0 ** -100):
        sc = f(mx)
        assert sc == 2.0 ** round(math.log2(sc))
        assert 2.0 ** 13 <= np.float32(mx) * sc < 2.0 ** 14, (mx, sc)
    for mx in (2.0 ** -113, 1e-38, 1e-45, np.finfo(np.float32).tiny):
        sc = f(mx)
        assert math.isfinite(sc) and sc == 2.0 ** 126, (mx, sc)
        assert np.float32(mx) * sc < 2.0 ** 14
        assert 1.0 / sc >= np.finfo(np.float32).tiny
    for mx in (0.0, -1.0, float("inf"), float("nan")):
        assert f(mx) == 1.0


def test_static_scratch_size_matches_header():
    text = open(os.path.join(ROOT, "include", "atmo.h")).read()
    m = re.search(r"#define ATMO_STATIC_SCRATCH_BYTES \((\d+)u << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == _capi.STATIC_SCRATCH_BYTES
