"""MI355X precomputation of Bruneton-style atmospheric scattering LUTs.

Python host-side mirror of the reference's atmosphere::Model
(atmosphere/model.h:160-337) over the C ABI of libatmo.so (include/atmo.h):
the same constructor parameters, init(num_scattering_orders) (alias Init),
ConvertSpectrumToLinearSrgb and kLambdaR/G/B. The precomputed LUTs live in
device memory; lut() copies one to the host, and the sky lookups run on
device buffers (DeviceArray) or on host arrays for convenience.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import AtmoError, check, lib

kLambdaR = 680.0
kLambdaG = 550.0
kLambdaB = 440.0

DEFAULT_SIZES = dict(transmittance_width=256, transmittance_height=64,
                     scattering_r=32, scattering_mu=128, scattering_mu_s=32,
                     scattering_nu=8, irradiance_width=64, irradiance_height=16)


class DensityProfileLayer(tuple):
    """(width, exp_term, exp_scale, linear_term, constant_term), in m and 1/m
    (atmosphere/model.h:165-178)."""

    def __new__(cls, width=0.0, exp_term=0.0, exp_scale=0.0, linear_term=0.0,
                constant_term=0.0):
        return super().__new__(cls, (float(width), float(exp_term), float(exp_scale),
                                     float(linear_term), float(constant_term)))


def _f64(v):
    return np.ascontiguousarray(v, dtype=np.float64)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Stream:
    """A HIP stream owned by libatmo (for callers without a HIP runtime)."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().atmo_stream_create(C.byref(h)), "atmo_stream_create")
        self.handle = h

    def synchronize(self):
        check(lib().atmo_stream_synchronize(self.handle), "atmo_stream_synchronize")

    def close(self):
        if self.handle:
            check(lib().atmo_stream_destroy(self.handle), "atmo_stream_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """A HIP event (timing work on a stream)."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().atmo_event_create(C.byref(h)), "atmo_event_create")
        self.handle = h

    def record(self, stream=None):
        check(lib().atmo_event_record(self.handle, _stream_handle(stream)),
              "atmo_event_record")

    def elapsed_ms(self, stop):
        """Milliseconds from this event to 'stop' (waits for 'stop')."""
        ms = C.c_float()
        check(lib().atmo_event_elapsed_ms(self.handle, stop.handle, C.byref(ms)),
              "atmo_event_elapsed_ms")
        return ms.value

    def close(self):
        if self.handle:
            lib().atmo_event_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_handle(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


# Lookup argument layouts: 3 = n x 3 floats, 1 = n floats (shadow_length).
_SKY_INPUTS = (3, 3, 1, 3)        # camera, view_ray/point, shadow_length, sun
_SUN_SKY_INPUTS = (3, 3, 3)       # point, normal, sun_direction


def _check_device_arrays(n, arrays, widths, what):
    """Raises ValueError unless n >= 0 and every device array is float32 and
    holds at least n x width floats. The kernels index the arrays with n
    only, so a short array would be read or written out of bounds."""
    if isinstance(n, bool) or int(n) != n or n < 0:
        raise ValueError("%s: n must be a non-negative integer" % what)
    n = int(n)
    for i, (a, w) in enumerate(zip(arrays, widths)):
        if not isinstance(a, DeviceArray) or a.ptr is None:
            raise ValueError("%s: argument %d is not a live DeviceArray" % (what, i))
        if a.dtype != np.float32:
            raise ValueError("%s: argument %d must be float32" % (what, i))
        if a.nbytes < 4 * n * w:
            raise ValueError("%s: argument %d holds %d bytes, %d x %d floats need %d"
                             % (what, i, a.nbytes, n, w, 4 * n * w))
    return n


def _check_host_arrays(arrays, widths, what):
    """Host inputs of a lookup as float32 arrays of one leading length n
    (n x 3, or n for shadow_length); returns (n, arrays)."""
    out = [np.ascontiguousarray(a, dtype=np.float32) for a in arrays]
    n = len(out[0])
    for i, (a, w) in enumerate(zip(out, widths)):
        want = (n, 3) if w == 3 else (n,)
        if a.shape != want:
            raise ValueError("%s: argument %d has shape %s, expected %s"
                             % (what, i, a.shape, want))
    return n, out


class DeviceArray:
    """A float32 device buffer allocated through libatmo."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().atmo_malloc(C.byref(p), max(self.nbytes, 4)), "atmo_malloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().atmo_memcpy_to_device(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes,
                                          _stream_handle(stream)), "atmo_memcpy_to_device")
        if stream is None:
            check(lib().atmo_stream_synchronize(None), "atmo_stream_synchronize")
        return d

    def to_host(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        check(lib().atmo_memcpy_to_host(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes,
                                        _stream_handle(stream)), "atmo_memcpy_to_host")
        return out

    def free(self):
        if self.ptr:
            check(lib().atmo_free(self.ptr), "atmo_free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ModelParameters:
    """The Model constructor parameters (atmosphere/model.h:182-281), checked
    and packed into the C struct atmo_model_params. Host-only: also gives the
    lowered AtmosphereParameters and the illuminance-mode iterations."""

    def __init__(self, wavelengths, solar_irradiance, sun_angular_radius, bottom_radius,
                 top_radius, rayleigh_density, rayleigh_scattering, mie_density,
                 mie_scattering, mie_extinction, mie_phase_function_g, absorption_density,
                 absorption_extinction, ground_albedo, max_sun_zenith_angle,
                 length_unit_in_meters, num_precomputed_wavelengths,
                 combine_scattering_textures, half_precision):
        self.wavelengths = _f64(wavelengths)
        n = len(self.wavelengths)
        spectra = dict(solar_irradiance=solar_irradiance,
                       rayleigh_scattering=rayleigh_scattering,
                       mie_scattering=mie_scattering, mie_extinction=mie_extinction,
                       absorption_extinction=absorption_extinction,
                       ground_albedo=ground_albedo)
        for k, v in spectra.items():
            v = _f64(v)
            if v.shape != (n,):
                raise ValueError("%s must have one value per wavelength" % k)
            setattr(self, k, v)
        mk = lambda ls: [DensityProfileLayer(*l) for l in ls]
        self.rayleigh_density = mk(rayleigh_density)
        self.mie_density = mk(mie_density)
        self.absorption_density = mk(absorption_density)
        for name in ("rayleigh_density", "mie_density", "absorption_density"):
            if len(getattr(self, name)) > 2:
                raise ValueError("%s: at most 2 layers" % name)
        self.sun_angular_radius = float(sun_angular_radius)
        self.bottom_radius = float(bottom_radius)
        self.top_radius = float(top_radius)
        self.mie_phase_function_g = float(mie_phase_function_g)
        self.max_sun_zenith_angle = float(max_sun_zenith_angle)
        self.length_unit_in_meters = float(length_unit_in_meters)
        self.num_precomputed_wavelengths = int(num_precomputed_wavelengths)
        self.combine_scattering_textures = bool(combine_scattering_textures)
        self.half_precision = bool(half_precision)
        carr = lambda ls: (_capi.DensityLayer * max(1, len(ls)))(
            *[_capi.DensityLayer(*l) for l in ls])
        self._layers = (carr(self.rayleigh_density), carr(self.mie_density),
                        carr(self.absorption_density))
        self.c = _capi.ModelParams(
            n, _dptr(self.wavelengths), _dptr(self.solar_irradiance),
            self.sun_angular_radius, self.bottom_radius, self.top_radius,
            len(self.rayleigh_density), self._layers[0], _dptr(self.rayleigh_scattering),
            len(self.mie_density), self._layers[1], _dptr(self.mie_scattering),
            _dptr(self.mie_extinction), self.mie_phase_function_g,
            len(self.absorption_density), self._layers[2],
            _dptr(self.absorption_extinction), _dptr(self.ground_albedo),
            self.max_sun_zenith_angle, self.length_unit_in_meters,
            self.num_precomputed_wavelengths, int(self.combine_scattering_textures),
            int(self.half_precision))

    def lower(self, lambdas):
        """AtmosphereParameters at the 3 wavelengths (atmo_lower_parameters)."""
        out = _capi.Atmosphere()
        check(lib().atmo_lower_parameters(C.byref(self.c), (C.c_double * 3)(*lambdas),
                                          C.byref(out)), "lower_parameters")
        return out

    def num_iterations(self):
        return lib().atmo_num_iterations(C.byref(self.c))

    def iteration(self, i):
        """(3 wavelengths, 3x3 luminance_from_radiance) of iteration i."""
        lam = (C.c_double * 3)()
        mat = (C.c_float * 9)()
        check(lib().atmo_iteration(C.byref(self.c), i, lam, mat), "iteration")
        return list(lam), np.array(list(mat), dtype=np.float32).reshape(3, 3)


class Model:
    """atmosphere::Model (atmosphere/model.h:182-281) on the GPU."""

    kLambdaR = kLambdaR
    kLambdaG = kLambdaG
    kLambdaB = kLambdaB

    def __init__(self, wavelengths, solar_irradiance, sun_angular_radius, bottom_radius,
                 top_radius, rayleigh_density, rayleigh_scattering, mie_density,
                 mie_scattering, mie_extinction, mie_phase_function_g, absorption_density,
                 absorption_extinction, ground_albedo, max_sun_zenith_angle,
                 length_unit_in_meters, num_precomputed_wavelengths,
                 combine_scattering_textures, half_precision, lut_sizes=None,
                 retain_intermediates=False, pass_timing=False, interleaved_lookups=True):
        self._handle = None
        self.params = ModelParameters(
            wavelengths, solar_irradiance, sun_angular_radius, bottom_radius, top_radius,
            rayleigh_density, rayleigh_scattering, mie_density, mie_scattering,
            mie_extinction, mie_phase_function_g, absorption_density, absorption_extinction,
            ground_albedo, max_sun_zenith_angle, length_unit_in_meters,
            num_precomputed_wavelengths, combine_scattering_textures, half_precision)
        self.combine_scattering_textures = self.params.combine_scattering_textures
        self.half_precision = self.params.half_precision
        self.lut_sizes = dict(DEFAULT_SIZES)
        if lut_sizes:
            self.lut_sizes.update(lut_sizes)
        z = _capi.LutSizes(**self.lut_sizes)
        h = C.c_void_p()
        check(lib().atmo_model_create(C.byref(self.params.c), C.byref(z), C.byref(h)),
              "Model")
        self._handle = h
        if retain_intermediates:
            check(lib().atmo_model_set_option(h, _capi.OPTION_RETAIN_INTERMEDIATES, 1),
                  "set_option")
        self.pass_timing = bool(pass_timing)
        if pass_timing:
            check(lib().atmo_model_set_option(h, _capi.OPTION_PASS_TIMING, 1),
                  "set_option")
        if not interleaved_lookups:
            self.set_interleaved_lookups(False)

    def set_interleaved_lookups(self, on):
        """ATMO_OPTION_INTERLEAVED_LOOKUPS: keep S and S_Mie interleaved for
        the lookups (separate textures only; synchronous)."""
        check(lib().atmo_model_set_option(self._handle, _capi.OPTION_INTERLEAVED_LOOKUPS,
                                          int(bool(on))), "set_option")

    def pass_timings(self):
        """[(pass name, order, ms)] of the launches of the last Init (needs
        pass_timing=True)."""
        n = C.c_int()
        check(lib().atmo_model_pass_timings(self._handle, None, 0, C.byref(n)),
              "pass_timings")
        buf = (_capi.PassTiming * max(1, n.value))()
        check(lib().atmo_model_pass_timings(self._handle, buf, n.value, C.byref(n)),
              "pass_timings")
        return [(_capi.PASS_NAMES[t.pass_], t.order, t.ms) for t in buf[:n.value]]

    @property
    def handle(self):
        return self._handle

    def close(self):
        """Frees the model (~Model). Raises AtmoError, keeping the model, if
        a live Batch holds it."""
        if self._handle:
            if not getattr(self, "_borrowed", False):
                check(lib().atmo_model_destroy(self._handle), "Model.close")
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # atmosphere/model.h:285
    def init(self, num_scattering_orders=4, stream=None):
        check(lib().atmo_model_init(self._handle, int(num_scattering_orders),
                                    _stream_handle(stream)), "Init")

    Init = init

    def lut_view(self, name):
        v = _capi.LutView()
        check(lib().atmo_model_get_lut(self._handle, _capi.LUT_IDS[name], C.byref(v)),
              "lut_view")
        return v

    def lut(self, name):
        """Copy of a LUT as float32 (depth, height, width, 4) (2D LUTs:
        (height, width, 4)), decoded to physical values."""
        v = self.lut_view(name)
        out = np.empty((v.depth, v.height, v.width, 4), dtype=np.float32)
        check(lib().atmo_model_copy_lut(self._handle, _capi.LUT_IDS[name],
                                        out.ctypes.data_as(C.POINTER(C.c_float))), "lut")
        return out[0] if v.depth == 1 and name in (
            "transmittance", "irradiance", "delta_irradiance") else out

    def export_luts(self, directory):
        """Write transmittance.dat, scattering.dat, single_mie_scattering.dat
        (separate Mie only) and irradiance.dat (raw float32 RGBA)."""
        check(lib().atmo_model_export_luts(self._handle, str(directory).encode()),
              "export_luts")

    def import_luts(self, directory):
        """Load LUT files written by export_luts (or the reference's WebGL
        exporter); the model is then initialised."""
        check(lib().atmo_model_import_luts(self._handle, str(directory).encode()),
              "import_luts")

    def init_cached(self, num_scattering_orders, directory, stream=None):
        """Init, or load the LUTs from 'directory' when it holds them (the
        reference CPU model's cache)."""
        check(lib().atmo_model_init_cached(self._handle, int(num_scattering_orders),
                                           str(directory).encode(), _stream_handle(stream)),
              "init_cached")

    def footprint(self):
        a, b = C.c_size_t(), C.c_size_t()
        check(lib().atmo_model_footprint(self._handle, C.byref(a), C.byref(b)), "footprint")
        return a.value, b.value

    def solar_radiance(self):
        out = (C.c_float * 3)()
        check(lib().atmo_get_solar_radiance(self._handle, out), "GetSolarRadiance")
        return np.array(list(out), dtype=np.float32)

    # ---- lookups on device arrays ----
    def sky_radiance_device(self, n, camera, view_ray, shadow_length, sun_direction,
                            radiance, transmittance, stream=None):
        n = _check_device_arrays(
            n, (camera, view_ray, shadow_length, sun_direction, radiance, transmittance),
            _SKY_INPUTS + (3, 3), "GetSkyRadiance")
        check(lib().atmo_get_sky_radiance(self._handle, n, camera.ptr, view_ray.ptr,
                                          shadow_length.ptr, sun_direction.ptr,
                                          radiance.ptr, transmittance.ptr,
                                          _stream_handle(stream)), "GetSkyRadiance")

    def sky_radiance_to_point_device(self, n, camera, point, shadow_length, sun_direction,
                                     radiance, transmittance, stream=None):
        n = _check_device_arrays(
            n, (camera, point, shadow_length, sun_direction, radiance, transmittance),
            _SKY_INPUTS + (3, 3), "GetSkyRadianceToPoint")
        check(lib().atmo_get_sky_radiance_to_point(self._handle, n, camera.ptr, point.ptr,
                                                   shadow_length.ptr, sun_direction.ptr,
                                                   radiance.ptr, transmittance.ptr,
                                                   _stream_handle(stream)),
              "GetSkyRadianceToPoint")

    def sun_and_sky_irradiance_device(self, n, point, normal, sun_direction,
                                      sun_irradiance, sky_irradiance, stream=None):
        n = _check_device_arrays(
            n, (point, normal, sun_direction, sun_irradiance, sky_irradiance),
            _SUN_SKY_INPUTS + (3, 3), "GetSunAndSkyIrradiance")
        check(lib().atmo_get_sun_and_sky_irradiance(self._handle, n, point.ptr, normal.ptr,
                                                    sun_direction.ptr, sun_irradiance.ptr,
                                                    sky_irradiance.ptr,
                                                    _stream_handle(stream)),
              "GetSunAndSkyIrradiance")

    # ---- luminance API (atmosphere/model.cc:252-281) on device arrays ----
    def sky_luminance_device(self, n, camera, view_ray, shadow_length, sun_direction,
                             luminance, transmittance, stream=None):
        n = _check_device_arrays(
            n, (camera, view_ray, shadow_length, sun_direction, luminance, transmittance),
            _SKY_INPUTS + (3, 3), "GetSkyLuminance")
        check(lib().atmo_get_sky_luminance(self._handle, n, camera.ptr, view_ray.ptr,
                                           shadow_length.ptr, sun_direction.ptr,
                                           luminance.ptr, transmittance.ptr,
                                           _stream_handle(stream)), "GetSkyLuminance")

    def sky_luminance_to_point_device(self, n, camera, point, shadow_length,
                                      sun_direction, luminance, transmittance, stream=None):
        n = _check_device_arrays(
            n, (camera, point, shadow_length, sun_direction, luminance, transmittance),
            _SKY_INPUTS + (3, 3), "GetSkyLuminanceToPoint")
        check(lib().atmo_get_sky_luminance_to_point(self._handle, n, camera.ptr, point.ptr,
                                                    shadow_length.ptr, sun_direction.ptr,
                                                    luminance.ptr, transmittance.ptr,
                                                    _stream_handle(stream)),
              "GetSkyLuminanceToPoint")

    def sun_and_sky_illuminance_device(self, n, point, normal, sun_direction,
                                       sun_illuminance, sky_illuminance, stream=None):
        n = _check_device_arrays(
            n, (point, normal, sun_direction, sun_illuminance, sky_illuminance),
            _SUN_SKY_INPUTS + (3, 3), "GetSunAndSkyIlluminance")
        check(lib().atmo_get_sun_and_sky_illuminance(self._handle, n, point.ptr,
                                                     normal.ptr, sun_direction.ptr,
                                                     sun_illuminance.ptr,
                                                     sky_illuminance.ptr,
                                                     _stream_handle(stream)),
              "GetSunAndSkyIlluminance")

    def solar_luminance(self):
        out = (C.c_float * 3)()
        check(lib().atmo_get_solar_luminance(self._handle, out), "GetSolarLuminance")
        return np.array(list(out), dtype=np.float32)

    def luminance_factors(self):
        """(SKY_SPECTRAL_RADIANCE_TO_LUMINANCE, SUN_SPECTRAL_RADIANCE_TO_LUMINANCE)."""
        sky, sun = (C.c_double * 3)(), (C.c_double * 3)()
        check(lib().atmo_get_luminance_factors(self._handle, sky, sun), "luminance_factors")
        return np.array(list(sky)), np.array(list(sun))

    # ---- host-array conveniences (copy in, run, copy out) ----
    def _host_call(self, fn, arrays, widths, what):
        n, host = _check_host_arrays(arrays, widths, what)
        d = [DeviceArray.from_host(a) for a in host]
        outs = [DeviceArray((n, 3)) for _ in range(2)]
        fn(n, *d, *outs)
        return tuple(o.to_host() for o in outs)

    def sky_luminance(self, camera, view_ray, shadow_length, sun_direction):
        return self._host_call(self.sky_luminance_device,
                               (camera, view_ray, shadow_length, sun_direction),
                               _SKY_INPUTS, "GetSkyLuminance")

    def sky_luminance_to_point(self, camera, point, shadow_length, sun_direction):
        return self._host_call(self.sky_luminance_to_point_device,
                               (camera, point, shadow_length, sun_direction),
                               _SKY_INPUTS, "GetSkyLuminanceToPoint")

    def sun_and_sky_illuminance(self, point, normal, sun_direction):
        return self._host_call(self.sun_and_sky_illuminance_device,
                               (point, normal, sun_direction), _SUN_SKY_INPUTS,
                               "GetSunAndSkyIlluminance")

    def sky_radiance(self, camera, view_ray, shadow_length, sun_direction):
        return self._host_call(self.sky_radiance_device,
                               (camera, view_ray, shadow_length, sun_direction),
                               _SKY_INPUTS, "GetSkyRadiance")

    def sky_radiance_to_point(self, camera, point, shadow_length, sun_direction):
        return self._host_call(self.sky_radiance_to_point_device,
                               (camera, point, shadow_length, sun_direction),
                               _SKY_INPUTS, "GetSkyRadianceToPoint")

    def sun_and_sky_irradiance(self, point, normal, sun_direction):
        return self._host_call(self.sun_and_sky_irradiance_device,
                               (point, normal, sun_direction), _SUN_SKY_INPUTS,
                               "GetSunAndSkyIrradiance")

    # atmosphere/model.h:301-304
    @staticmethod
    def ConvertSpectrumToLinearSrgb(wavelengths, spectrum):
        w, s = _f64(wavelengths), _f64(spectrum)
        if w.shape != s.shape:
            raise ValueError("spectrum must have one value per wavelength")
        r, g, b = C.c_double(), C.c_double(), C.c_double()
        check(lib().atmo_convert_spectrum_to_linear_srgb(len(w), _dptr(w), _dptr(s),
                                                         C.byref(r), C.byref(g),
                                                         C.byref(b)),
              "ConvertSpectrumToLinearSrgb")
        return r.value, g.value, b.value


class Batch:
    """Reusable batched precompute of several models (atmo_batch_*)."""

    def __init__(self, models, orders, use_graph=True):
        self.models = list(models)
        n = len(self.models)
        hs = (C.c_void_p * n)(*[m.handle for m in self.models])
        os_ = (C.c_uint * n)(*[int(o) for o in orders])
        h = C.c_void_p()
        check(lib().atmo_batch_create(hs, os_, n, int(bool(use_graph)), C.byref(h)),
              "atmo_batch_create")
        self.handle = h

    @property
    def num_launches(self):
        return lib().atmo_batch_num_launches(self.handle)

    def launch(self, stream=None):
        check(lib().atmo_batch_launch(self.handle, _stream_handle(stream)),
              "atmo_batch_launch")

    def close(self):
        if self.handle:
            lib().atmo_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over one GPU (atmo_server_*): precompute requests are
    admitted by device-memory footprint into concurrent batches (continuous
    batching); lookups run on resident models. Requests complete
    asynchronously; poll() returns their completion records."""

    def __init__(self, memory_budget=0, max_inflight_batches=0, max_batch_models=0,
                 lookup_streams=0):
        cfg = _capi.ServerConfig(int(memory_budget), int(max_inflight_batches),
                                 int(max_batch_models), int(lookup_streams))
        h = C.c_void_p()
        check(lib().atmo_server_create(C.byref(cfg), C.byref(h)), "atmo_server_create")
        self.handle = h
        self._keep = {}   # request id -> objects that must outlive the request

    def submit_precompute(self, num_scattering_orders=4, lut_sizes=None, **params):
        """Model constructor parameters as keywords; returns the request id,
        which is also the id of the resulting model."""
        p = ModelParameters(**params)
        sizes = dict(DEFAULT_SIZES)
        if lut_sizes:
            sizes.update(lut_sizes)
        z = _capi.LutSizes(**sizes)
        rid = C.c_uint64()
        check(lib().atmo_server_submit_precompute(self.handle, C.byref(p.c), C.byref(z),
                                                  int(num_scattering_orders),
                                                  C.byref(rid)), "submit_precompute")
        return rid.value

    def submit_lookup(self, model_id, lookup, n, inputs, outputs, luminance=False):
        """lookup: 'sky', 'sky_to_point' or 'sun_and_sky'. inputs/outputs are
        DeviceArrays in the order of the corresponding Model *_device method;
        they are kept alive until the request completes."""
        kind = {"sky": _capi.LOOKUP_SKY, "sky_to_point": _capi.LOOKUP_SKY_TO_POINT,
                "sun_and_sky": _capi.LOOKUP_SUN_AND_SKY}[lookup]
        widths = _SUN_SKY_INPUTS if lookup == "sun_and_sky" else _SKY_INPUTS
        if len(inputs) != len(widths) or len(outputs) != 2:
            raise ValueError("submit_lookup: %s takes %d inputs and 2 outputs"
                             % (lookup, len(widths)))
        n = _check_device_arrays(n, list(inputs) + list(outputs), widths + (3, 3),
                                 "submit_lookup")
        ins = [a.ptr for a in inputs] + [None] * (4 - len(inputs))
        outs = [a.ptr for a in outputs]
        rid = C.c_uint64()
        check(lib().atmo_server_submit_lookup(self.handle, int(model_id), kind,
                                              int(bool(luminance)), int(n), *ins, *outs,
                                              C.byref(rid)), "submit_lookup")
        self._keep[rid.value] = (inputs, outputs)
        return rid.value

    def poll(self, capacity=256, timeout_ms=0):
        buf = (_capi.Completion * capacity)()
        n = C.c_int()
        check(lib().atmo_server_poll(self.handle, buf, capacity, C.byref(n),
                                     int(timeout_ms)), "poll")
        out = []
        for c in buf[:n.value]:
            self._keep.pop(c.request_id, None)
            out.append(dict(request_id=c.request_id, kind=c.kind, status=c.status,
                            model_id=c.model_id, batch_models=c.batch_models,
                            submit_ms=c.submit_ms, start_ms=c.start_ms, end_ms=c.end_ms))
        return out

    def wait(self, request_ids, timeout_s=60.0):
        """Poll until every id in request_ids has completed; returns their
        completion records by id."""
        import time
        want = set(request_ids)
        got = {}
        deadline = time.monotonic() + timeout_s
        while want - set(got):
            if time.monotonic() > deadline:
                raise TimeoutError("requests did not complete: %s" % sorted(want - set(got)))
            for c in self.poll(timeout_ms=20):
                got[c["request_id"]] = c
        return got

    def model(self, model_id):
        """A borrowed Model view of a resident model (lut_view, lut, lookups);
        valid until the model is released."""
        h = C.c_void_p()
        check(lib().atmo_server_model(self.handle, int(model_id), C.byref(h)),
              "atmo_server_model")
        m = Model.__new__(Model)
        m._handle = h
        m._borrowed = True
        return m

    def release(self, model_id):
        check(lib().atmo_server_release(self.handle, int(model_id)), "release")

    def stats(self):
        st = _capi.ServerStats()
        check(lib().atmo_server_get_stats(self.handle, C.byref(st)), "stats")
        return {k: getattr(st, k) for k, _ in _capi.ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().atmo_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


QUERY_FIELDS = ("camera", "view_ray", "sun_direction", "point", "normal", "shadow_length")


def generate_sky_queries(seed, n, bottom_radius, top_radius, max_altitude=100.0,
                         max_distance=50.0, max_sun_zenith=None, out=None, stream=None):
    """Seeded synthetic query batch made on the device
    (atmo_generate_sky_queries), as {field: DeviceArray}; lengths in the
    model's unit. Defaults: altitude U[0, 100], hit distance U[0, 50], sun
    zenith U[0, 100 deg]. out: existing {field: DeviceArray} to fill."""
    import math
    if max_sun_zenith is None:
        max_sun_zenith = math.radians(100.0)
    if out is None:
        out = {k: DeviceArray((n,) if k == "shadow_length" else (n, 3))
               for k in QUERY_FIELDS}
    for k in QUERY_FIELDS:
        if out[k].nbytes < 4 * n * (1 if k == "shadow_length" else 3):
            raise ValueError("output array %s too small" % k)
    check(lib().atmo_generate_sky_queries(
        int(seed), int(n), float(bottom_radius), float(top_radius), float(max_altitude),
        float(max_distance), float(max_sun_zenith), *[out[k].ptr for k in QUERY_FIELDS],
        _stream_handle(stream)), "generate_sky_queries")
    return out


def init_batch(models, orders, stream=None):
    n = len(models)
    hs = (C.c_void_p * n)(*[m.handle for m in models])
    os_ = (C.c_uint * n)(*[int(o) for o in orders])
    check(lib().atmo_model_init_batch(hs, os_, n, _stream_handle(stream)), "init_batch")


def synchronize():
    check(lib().atmo_device_synchronize(), "atmo_device_synchronize")


# ---------------------------------------------------------------------------
# Tile shapes of the scattering-density and multiple-scattering passes, and
# their autotuner (atmo_autotune and friends). A shape is
# {"sd_nu_per_thread", "sd_block_threads", "ms_texels_per_thread",
# "ms_block_threads", "ms_row_order"} (row order: MS_ROW_ORDERS); sizes are
# LUT-size dicts
# (DEFAULT_SIZES keys; missing keys take the defaults).
# ---------------------------------------------------------------------------

# Multiple-scattering row orders (atmo_tiles.ms_row_order).
MS_ROW_ORDERS = ("r-major", "mu-major", "ray")


def _sizes(sizes):
    z = dict(DEFAULT_SIZES)
    if sizes:
        z.update(sizes)
    return _capi.LutSizes(**z)


def tile_candidates():
    """(scattering-density shapes [(nu per thread, block threads)],
    multiple-scattering shapes [(texels per thread, block threads)])."""
    n = lib().atmo_sd_tile_candidates(None, None, 0)
    a, b = (C.c_int * n)(), (C.c_int * n)()
    lib().atmo_sd_tile_candidates(a, b, n)
    m = lib().atmo_ms_tile_candidates(None, None, 0)
    c, d = (C.c_int * m)(), (C.c_int * m)()
    lib().atmo_ms_tile_candidates(c, d, m)
    return list(zip(a, b)), list(zip(c, d))


def autotune(sizes_list, retune=False, stream=None):
    """Times every tile shape for the size classes of sizes_list and records
    the fastest (synchronous)."""
    n = len(sizes_list)
    arr = (_capi.LutSizes * max(1, n))(*[_sizes(z) for z in sizes_list])
    check(lib().atmo_autotune(arr, n, int(bool(retune)), _stream_handle(stream)),
          "atmo_autotune")


def get_tiles(sizes=None):
    t = _capi.Tiles()
    z = _sizes(sizes)
    check(lib().atmo_get_tiles(C.byref(z), C.byref(t)), "atmo_get_tiles")
    return dict(sd_nu_per_thread=t.sd_nu_per_thread, sd_block_threads=t.sd_block_threads,
                ms_texels_per_thread=t.ms_texels_per_thread,
                ms_block_threads=t.ms_block_threads, ms_row_order=MS_ROW_ORDERS[t.ms_row_order], sd_tuned=bool(t.tuned & 1),
                ms_tuned=bool(t.tuned & 2))


def set_tiles(sizes, sd_nu_per_thread, sd_block_threads, ms_texels_per_thread,
              ms_block_threads, ms_row_order=None):
    """Pins the shapes of the size classes of 'sizes'; ms_row_order (a name
    of MS_ROW_ORDERS) None keeps the class's current row order."""
    if ms_row_order is None:
        ms_row_order = get_tiles(sizes)["ms_row_order"]
    t = _capi.Tiles(sd_nu_per_thread, sd_block_threads, ms_texels_per_thread,
                    ms_block_threads, MS_ROW_ORDERS.index(ms_row_order), 0)
    z = _sizes(sizes)
    check(lib().atmo_set_tiles(C.byref(z), C.byref(t)), "atmo_set_tiles")


def save_tiles(path):
    check(lib().atmo_tune_save(str(path).encode()), "atmo_tune_save")


def load_tiles(path):
    check(lib().atmo_tune_load(str(path).encode()), "atmo_tune_load")


def reset_tiles():
    lib().atmo_tune_reset()
