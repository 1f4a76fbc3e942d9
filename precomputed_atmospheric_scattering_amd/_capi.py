"""ctypes declarations of libatmo.so (include/atmo.h).

The library is the product: there is no fallback. Loading fails with a clear
error when libatmo.so has not been built.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libatmo.so")

ATMO_OK = 0
STATUS = {0: "ATMO_OK", 1: "ATMO_ERROR_INVALID_ARGUMENT", 2: "ATMO_ERROR_OUT_OF_MEMORY",
          3: "ATMO_ERROR_DEVICE", 4: "ATMO_ERROR_NOT_INITIALIZED", 5: "ATMO_ERROR_UNSUPPORTED"}

LUT_IDS = {
    "transmittance": 0, "scattering": 1, "single_mie_scattering": 2, "irradiance": 3,
    "delta_irradiance": 4, "delta_rayleigh": 5, "delta_mie": 6,
    "delta_scattering_density": 7,
}
OPTION_RETAIN_INTERMEDIATES = 0
OPTION_PASS_TIMING = 1
OPTION_INTERLEAVED_LOOKUPS = 2
# ATMO_STATIC_SCRATCH_BYTES: the static scratch of atmo_model_init_static_scratch.
STATIC_SCRATCH_BYTES = 64 << 20
PASS_NAMES = ("transmittance", "direct_irradiance", "single_scattering",
              "scattering_density", "indirect_irradiance", "multiple_scattering")


class AtmoError(RuntimeError):
    def __init__(self, status, what, message):
        super().__init__("%s failed (%s): %s" % (what, STATUS.get(status, status), message))
        self.status = status


class DensityLayer(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("width", "exp_term", "exp_scale", "linear_term", "constant_term")]


DP = C.POINTER(C.c_double)


class ModelParams(C.Structure):
    _fields_ = [
        ("num_wavelengths", C.c_int), ("wavelengths", DP), ("solar_irradiance", DP),
        ("sun_angular_radius", C.c_double), ("bottom_radius", C.c_double),
        ("top_radius", C.c_double),
        ("num_rayleigh_layers", C.c_int), ("rayleigh_density", C.POINTER(DensityLayer)),
        ("rayleigh_scattering", DP),
        ("num_mie_layers", C.c_int), ("mie_density", C.POINTER(DensityLayer)),
        ("mie_scattering", DP), ("mie_extinction", DP),
        ("mie_phase_function_g", C.c_double),
        ("num_absorption_layers", C.c_int), ("absorption_density", C.POINTER(DensityLayer)),
        ("absorption_extinction", DP), ("ground_albedo", DP),
        ("max_sun_zenith_angle", C.c_double), ("length_unit_in_meters", C.c_double),
        ("num_precomputed_wavelengths", C.c_uint),
        ("combine_scattering_textures", C.c_int), ("half_precision", C.c_int),
    ]


class LutSizes(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "transmittance_width", "transmittance_height", "scattering_r", "scattering_mu",
        "scattering_mu_s", "scattering_nu", "irradiance_width", "irradiance_height")]


class LutView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int), ("height", C.c_int),
                ("depth", C.c_int), ("channels", C.c_int), ("dtype", C.c_int),
                ("scale", C.c_float * 4)]


D3 = C.c_double * 3


class PassTiming(C.Structure):
    _fields_ = [("pass_", C.c_int), ("order", C.c_int), ("ms", C.c_float)]


class Atmosphere(C.Structure):
    """atmo_atmosphere: AtmosphereParameters at three wavelengths."""
    _fields_ = [
        ("solar_irradiance", D3), ("sun_angular_radius", C.c_double),
        ("bottom_radius", C.c_double), ("top_radius", C.c_double),
        ("rayleigh_density", DensityLayer * 2), ("rayleigh_scattering", D3),
        ("mie_density", DensityLayer * 2), ("mie_scattering", D3), ("mie_extinction", D3),
        ("mie_phase_function_g", C.c_double),
        ("absorption_density", DensityLayer * 2), ("absorption_extinction", D3),
        ("ground_albedo", D3), ("mu_s_min", C.c_double),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_size_t), ("max_inflight_batches", C.c_int),
                ("max_batch_models", C.c_int), ("lookup_streams", C.c_int)]


class Completion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("kind", C.c_int), ("status", C.c_int),
                ("model_id", C.c_uint64), ("batch_models", C.c_int),
                ("submit_ms", C.c_double), ("start_ms", C.c_double),
                ("end_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("budget_bytes", C.c_size_t), ("resident_bytes", C.c_size_t),
                ("scratch_bytes", C.c_size_t), ("resident_models", C.c_int),
                ("pending_precomputes", C.c_int), ("pending_lookups", C.c_int),
                ("inflight_batches", C.c_int), ("inflight_lookups", C.c_int),
                ("completed_precomputes", C.c_uint64), ("completed_lookups", C.c_uint64),
                ("failed", C.c_uint64)]


class Tiles(C.Structure):
    _fields_ = [("sd_nu_per_thread", C.c_int), ("sd_block_threads", C.c_int),
                ("ms_texels_per_thread", C.c_int), ("ms_block_threads", C.c_int),
                ("ms_row_order", C.c_int), ("tuned", C.c_int)]


REQUEST_PRECOMPUTE, REQUEST_LOOKUP = 0, 1
LOOKUP_SKY, LOOKUP_SKY_TO_POINT, LOOKUP_SUN_AND_SKY = 0, 1, 2


# Every symbol declared in include/atmo.h, with its ctypes signature.
VP, FP, I64, SZ = C.c_void_p, C.POINTER(C.c_float), C.c_int64, C.c_size_t
ST = C.c_int
SIGNATURES = {
    "atmo_model_create": (ST, [C.POINTER(ModelParams), C.POINTER(LutSizes), C.POINTER(VP)]),
    "atmo_model_destroy": (ST, [VP]),
    "atmo_fp16_channel_scale": (C.c_float, [C.c_float]),
    "atmo_model_set_option": (ST, [VP, C.c_int, C.c_int]),
    "atmo_model_pass_timings": (ST, [VP, C.POINTER(PassTiming), C.c_int,
                                     C.POINTER(C.c_int)]),
    "atmo_model_init": (ST, [VP, C.c_uint, VP]),
    "atmo_model_init_static_scratch": (ST, [VP, C.c_uint, VP]),
    "atmo_model_init_batch": (ST, [C.POINTER(VP), C.POINTER(C.c_uint), C.c_int, VP]),
    "atmo_batch_create": (ST, [C.POINTER(VP), C.POINTER(C.c_uint), C.c_int, C.c_int,
                               C.POINTER(VP)]),
    "atmo_batch_launch": (ST, [VP, VP]),
    "atmo_batch_num_launches": (C.c_int, [VP]),
    "atmo_batch_destroy": (None, [VP]),
    "atmo_sd_tile_candidates": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
    "atmo_ms_tile_candidates": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
    "atmo_autotune": (ST, [C.POINTER(LutSizes), C.c_int, C.c_int, VP]),
    "atmo_get_tiles": (ST, [C.POINTER(LutSizes), C.POINTER(Tiles)]),
    "atmo_set_tiles": (ST, [C.POINTER(LutSizes), C.POINTER(Tiles)]),
    "atmo_tune_save": (ST, [C.c_char_p]),
    "atmo_tune_load": (ST, [C.c_char_p]),
    "atmo_tune_reset": (None, []),
    "atmo_model_get_lut": (ST, [VP, C.c_int, C.POINTER(LutView)]),
    "atmo_model_copy_lut": (ST, [VP, C.c_int, FP]),
    "atmo_model_footprint": (ST, [VP, C.POINTER(SZ), C.POINTER(SZ)]),
    "atmo_model_export_luts": (ST, [VP, C.c_char_p]),
    "atmo_model_import_luts": (ST, [VP, C.c_char_p]),
    "atmo_model_init_cached": (ST, [VP, C.c_uint, C.c_char_p, VP]),
    "atmo_get_solar_radiance": (ST, [VP, FP]),
    "atmo_get_sky_radiance": (ST, [VP, I64, VP, VP, VP, VP, VP, VP, VP]),
    "atmo_get_sky_radiance_to_point": (ST, [VP, I64, VP, VP, VP, VP, VP, VP, VP]),
    "atmo_get_sun_and_sky_irradiance": (ST, [VP, I64, VP, VP, VP, VP, VP, VP]),
    "atmo_get_solar_luminance": (ST, [VP, FP]),
    "atmo_get_sky_luminance": (ST, [VP, I64, VP, VP, VP, VP, VP, VP, VP]),
    "atmo_get_sky_luminance_to_point": (ST, [VP, I64, VP, VP, VP, VP, VP, VP, VP]),
    "atmo_get_sun_and_sky_illuminance": (ST, [VP, I64, VP, VP, VP, VP, VP, VP]),
    "atmo_get_luminance_factors": (ST, [VP, DP, DP]),
    "atmo_convert_spectrum_to_linear_srgb": (ST, [C.c_int, DP, DP, DP, DP, DP]),
    "atmo_lower_parameters": (ST, [C.POINTER(ModelParams), DP, C.POINTER(Atmosphere)]),
    "atmo_num_iterations": (C.c_int, [C.POINTER(ModelParams)]),
    "atmo_iteration": (ST, [C.POINTER(ModelParams), C.c_int, DP, FP]),
    "atmo_server_create": (ST, [C.POINTER(ServerConfig), C.POINTER(VP)]),
    "atmo_server_destroy": (None, [VP]),
    "atmo_server_submit_precompute": (ST, [VP, C.POINTER(ModelParams), C.POINTER(LutSizes),
                                           C.c_uint, C.POINTER(C.c_uint64)]),
    "atmo_server_submit_lookup": (ST, [VP, C.c_uint64, C.c_int, C.c_int, I64, VP, VP, VP,
                                       VP, VP, VP, C.POINTER(C.c_uint64)]),
    "atmo_server_poll": (ST, [VP, C.POINTER(Completion), C.c_int, C.POINTER(C.c_int),
                              C.c_int]),
    "atmo_server_release": (ST, [VP, C.c_uint64]),
    "atmo_server_model": (ST, [VP, C.c_uint64, C.POINTER(VP)]),
    "atmo_server_get_stats": (ST, [VP, C.POINTER(ServerStats)]),
    "atmo_generate_sky_queries": (ST, [C.c_uint64, I64, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, VP, VP, VP, VP, VP, VP, VP]),
    "atmo_last_error": (C.c_char_p, []),
    "atmo_malloc": (ST, [C.POINTER(VP), SZ]),
    "atmo_free": (ST, [VP]),
    "atmo_memcpy_to_device": (ST, [VP, VP, SZ, VP]),
    "atmo_memcpy_to_host": (ST, [VP, VP, SZ, VP]),
    "atmo_stream_create": (ST, [C.POINTER(VP)]),
    "atmo_stream_destroy": (ST, [VP]),
    "atmo_stream_synchronize": (ST, [VP]),
    "atmo_device_synchronize": (ST, []),
    "atmo_event_create": (ST, [C.POINTER(VP)]),
    "atmo_event_destroy": (ST, [VP]),
    "atmo_event_record": (ST, [VP, VP]),
    "atmo_event_elapsed_ms": (ST, [VP, VP, C.POINTER(C.c_float)]),
}

_lib = None


def lib():
    """The loaded libatmo.so. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libatmo.so is not built (expected at %s); run "
                "`python -m precomputed_atmospheric_scattering_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what):
    if status != ATMO_OK:
        msg = lib().atmo_last_error()
        raise AtmoError(status, what, msg.decode() if msg else "")
