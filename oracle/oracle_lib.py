"""ctypes bindings of the FP64 oracle (oracle/liboracle.so).

Test infrastructure only: the oracle gives the reference results the
product is compared against. Only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py load it; the
product (precomputed_atmospheric_scattering_amd) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")


class Layer(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("width", "exp_term", "exp_scale", "linear_term", "constant_term")]


D3 = C.c_double * 3


class Atmosphere(C.Structure):
    _fields_ = [
        ("solar_irradiance", D3), ("sun_angular_radius", C.c_double),
        ("bottom_radius", C.c_double), ("top_radius", C.c_double),
        ("rayleigh_density", Layer * 2), ("rayleigh_scattering", D3),
        ("mie_density", Layer * 2), ("mie_scattering", D3), ("mie_extinction", D3),
        ("mie_phase_function_g", C.c_double),
        ("absorption_density", Layer * 2), ("absorption_extinction", D3),
        ("ground_albedo", D3), ("mu_s_min", C.c_double),
    ]


class Sizes(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "transmittance_width", "transmittance_height", "scattering_r",
        "scattering_mu", "scattering_mu_s", "scattering_nu", "irradiance_width",
        "irradiance_height")]


DP = C.POINTER(C.c_double)


class ModelParams(C.Structure):
    _fields_ = [
        ("num_wavelengths", C.c_int), ("wavelengths", DP), ("solar_irradiance", DP),
        ("sun_angular_radius", C.c_double), ("bottom_radius", C.c_double),
        ("top_radius", C.c_double),
        ("num_rayleigh_layers", C.c_int), ("rayleigh_density", C.POINTER(Layer)),
        ("rayleigh_scattering", DP),
        ("num_mie_layers", C.c_int), ("mie_density", C.POINTER(Layer)),
        ("mie_scattering", DP), ("mie_extinction", DP),
        ("mie_phase_function_g", C.c_double),
        ("num_absorption_layers", C.c_int), ("absorption_density", C.POINTER(Layer)),
        ("absorption_extinction", DP), ("ground_albedo", DP),
        ("max_sun_zenith_angle", C.c_double), ("length_unit_in_meters", C.c_double),
        ("num_precomputed_wavelengths", C.c_uint),
        ("combine_scattering_textures", C.c_int), ("half_precision", C.c_int),
    ]


DEFAULT_SIZES = dict(transmittance_width=256, transmittance_height=64,
                     scattering_r=32, scattering_mu=128, scattering_mu_s=32,
                     scattering_nu=8, irradiance_width=64, irradiance_height=16)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    A, S, LP = C.POINTER(Atmosphere), C.POINTER(Sizes), C.POINTER(Layer)
    d, i = C.c_double, C.c_int
    sig = {
        "or_set_threads": (None, [i]), "or_get_threads": (i, []),
        "or_distance_to_top": (d, [A, d, d]), "or_distance_to_bottom": (d, [A, d, d]),
        "or_ray_intersects_ground": (i, [A, d, d]),
        "or_distance_to_nearest": (d, [A, d, d, i]),
        "or_profile_density": (d, [LP, d]), "or_optical_length": (d, [A, LP, d, d]),
        "or_transmittance_to_top": (None, [A, d, d, DP]),
        "or_tex_coord_from_unit_range": (d, [d, i]),
        "or_unit_range_from_tex_coord": (d, [d, i]),
        "or_transmittance_uv_from_rmu": (None, [A, S, d, d, DP]),
        "or_rmu_from_transmittance_uv": (None, [A, S, d, d, DP, DP]),
        "or_get_transmittance_to_top": (None, [A, S, DP, d, d, DP]),
        "or_get_transmittance": (None, [A, S, DP, d, d, d, i, DP]),
        "or_get_transmittance_to_sun": (None, [A, S, DP, d, d, DP]),
        "or_single_scattering_integrand": (None, [A, S, DP, d, d, d, d, d, i, DP, DP]),
        "or_single_scattering": (None, [A, S, DP, d, d, d, d, i, DP, DP]),
        "or_rayleigh_phase": (d, [d]), "or_mie_phase": (d, [d, d]),
        "or_scattering_uvwz_from_rmumusnu": (None, [A, S, d, d, d, d, i, DP]),
        "or_rmumusnu_from_scattering_uvwz": (None, [A, S, DP, DP, C.POINTER(i)]),
        "or_rmumusnu_from_scattering_texel": (None, [A, S, i, i, i, DP, C.POINTER(i)]),
        "or_get_scattering": (None, [A, S, DP, d, d, d, d, i, DP]),
        "or_scattering_density": (None, [A, S, DP, DP, DP, DP, DP, d, d, d, d, i, DP]),
        "or_multiple_scattering": (None, [A, S, DP, DP, d, d, d, d, i, DP]),
        "or_direct_irradiance": (None, [A, S, DP, d, d, DP]),
        "or_indirect_irradiance": (None, [A, S, DP, DP, DP, d, d, i, DP]),
        "or_irradiance_uv_from_rmus": (None, [A, S, d, d, DP]),
        "or_rmus_from_irradiance_uv": (None, [A, S, d, d, DP, DP]),
        "or_get_irradiance": (None, [A, S, DP, d, d, DP]),
        "or_pass_transmittance": (None, [A, S, DP]),
        "or_pass_direct_irradiance": (None, [A, S, DP, DP]),
        "or_pass_single_scattering": (None, [A, S, DP, DP, DP]),
        "or_pass_scattering_density": (None, [A, S, DP, DP, DP, DP, DP, i, DP]),
        "or_pass_indirect_irradiance": (None, [A, S, DP, DP, DP, i, DP]),
        "or_pass_multiple_scattering": (None, [A, S, DP, DP, DP, DP]),
        "or_texels_single_scattering": (None, [A, S, DP, i, C.POINTER(i), DP, DP]),
        "or_texels_scattering_density": (None, [A, S, DP, DP, DP, DP, DP, i, i,
                                                C.POINTER(i), DP]),
        "or_texels_multiple_scattering": (None, [A, S, DP, DP, i, C.POINTER(i), DP]),
        "or_interpolate": (d, [i, DP, DP, d]), "or_cie_table_value": (d, [d, i]),
        "or_luminance_factors": (None, [i, DP, DP, d, DP]),
        "or_convert_spectrum_to_linear_srgb": (None, [i, DP, DP, DP]),
        "or_lower": (None, [C.POINTER(ModelParams), DP, A]),
        "or_num_iterations": (i, [C.POINTER(ModelParams)]),
        "or_iteration": (None, [C.POINTER(ModelParams), i, DP, DP]),
        "or_model_init": (i, [C.POINTER(ModelParams), S, C.c_uint, DP, DP, DP, DP]),
        "or_model_init_orders": (i, [C.POINTER(ModelParams), S, C.c_uint, DP,
                                     C.POINTER(DP), C.POINTER(DP), C.POINTER(DP)]),
        "or_precompute": (None, [A, S, DP, i, C.c_uint, i, DP, DP, DP, DP]),
        "or_sky_radiance": (None, [A, S, DP, DP, DP, i, i, DP, DP, DP, DP, DP, DP]),
        "or_sky_radiance_to_point": (None, [A, S, DP, DP, DP, i, i, DP, DP, DP, DP,
                                            DP, DP]),
        "or_sun_and_sky_irradiance": (None, [A, S, DP, DP, i, DP, DP, DP, DP, DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


# ---------------------------------------------------------------------------
# Helpers.
# ---------------------------------------------------------------------------

def ptr(a):
    """double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def layer(width=0.0, exp_term=0.0, exp_scale=0.0, linear_term=0.0, constant_term=0.0):
    return Layer(width, exp_term, exp_scale, linear_term, constant_term)


def profile(*layers):
    """Two-layer profile array; fewer layers are padded at the front."""
    ls = list(layers)
    while len(ls) < 2:
        ls.insert(0, Layer())
    return (Layer * 2)(*ls)


def sizes(**kw):
    s = dict(DEFAULT_SIZES)
    s.update(kw)
    return Sizes(**s)


def lut(w, h, d=1):
    return np.zeros((d, h, w, 4), dtype=np.float64)


def scattering_shape(s):
    return (s.scattering_r, s.scattering_mu, s.scattering_nu * s.scattering_mu_s, 4)


def transmittance_shape(s):
    return (s.transmittance_height, s.transmittance_width, 4)


def irradiance_shape(s):
    return (s.irradiance_height, s.irradiance_width, 4)


def v3(x):
    return (C.c_double * 3)(*x)


class ModelSpec:
    """Constructor parameters of the Model (atmosphere/model.h:182-281), kept
    alive for ctypes."""

    def __init__(self, wavelengths, solar_irradiance, sun_angular_radius,
                 bottom_radius, top_radius, rayleigh_density, rayleigh_scattering,
                 mie_density, mie_scattering, mie_extinction, mie_phase_function_g,
                 absorption_density, absorption_extinction, ground_albedo,
                 max_sun_zenith_angle, length_unit_in_meters,
                 num_precomputed_wavelengths, combine_scattering_textures,
                 half_precision):
        f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
        self.wavelengths = f(wavelengths)
        self.solar_irradiance = f(solar_irradiance)
        self.rayleigh_scattering = f(rayleigh_scattering)
        self.mie_scattering = f(mie_scattering)
        self.mie_extinction = f(mie_extinction)
        self.absorption_extinction = f(absorption_extinction)
        self.ground_albedo = f(ground_albedo)
        n = len(self.wavelengths)
        for v in (self.solar_irradiance, self.rayleigh_scattering, self.mie_scattering,
                  self.mie_extinction, self.absorption_extinction, self.ground_albedo):
            assert len(v) == n
        mk = lambda ls: (Layer * max(1, len(ls)))(*[Layer(*l) for l in ls])
        self.rayleigh_density = [tuple(l) for l in rayleigh_density]
        self.mie_density = [tuple(l) for l in mie_density]
        self.absorption_density = [tuple(l) for l in absorption_density]
        self._rd, self._md, self._ad = (mk(self.rayleigh_density), mk(self.mie_density),
                                        mk(self.absorption_density))
        self.sun_angular_radius = float(sun_angular_radius)
        self.bottom_radius = float(bottom_radius)
        self.top_radius = float(top_radius)
        self.mie_phase_function_g = float(mie_phase_function_g)
        self.max_sun_zenith_angle = float(max_sun_zenith_angle)
        self.length_unit_in_meters = float(length_unit_in_meters)
        self.num_precomputed_wavelengths = int(num_precomputed_wavelengths)
        self.combine_scattering_textures = bool(combine_scattering_textures)
        self.half_precision = bool(half_precision)
        self.c = ModelParams(
            n, ptr(self.wavelengths), ptr(self.solar_irradiance), self.sun_angular_radius,
            self.bottom_radius, self.top_radius,
            len(self.rayleigh_density), self._rd, ptr(self.rayleigh_scattering),
            len(self.mie_density), self._md, ptr(self.mie_scattering),
            ptr(self.mie_extinction), self.mie_phase_function_g,
            len(self.absorption_density), self._ad, ptr(self.absorption_extinction),
            ptr(self.ground_albedo), self.max_sun_zenith_angle, self.length_unit_in_meters,
            self.num_precomputed_wavelengths, int(self.combine_scattering_textures),
            int(self.half_precision))

    def kwargs(self):
        return dict(
            wavelengths=self.wavelengths, solar_irradiance=self.solar_irradiance,
            sun_angular_radius=self.sun_angular_radius, bottom_radius=self.bottom_radius,
            top_radius=self.top_radius, rayleigh_density=self.rayleigh_density,
            rayleigh_scattering=self.rayleigh_scattering, mie_density=self.mie_density,
            mie_scattering=self.mie_scattering, mie_extinction=self.mie_extinction,
            mie_phase_function_g=self.mie_phase_function_g,
            absorption_density=self.absorption_density,
            absorption_extinction=self.absorption_extinction,
            ground_albedo=self.ground_albedo,
            max_sun_zenith_angle=self.max_sun_zenith_angle,
            length_unit_in_meters=self.length_unit_in_meters,
            num_precomputed_wavelengths=self.num_precomputed_wavelengths,
            combine_scattering_textures=self.combine_scattering_textures,
            half_precision=self.half_precision)


def lower(spec, lambdas):
    a = Atmosphere()
    lib().or_lower(C.byref(spec.c), (C.c_double * 3)(*lambdas), C.byref(a))
    return a


def model_init(spec, s, orders):
    """Full FP64 Model::Init. Returns dict of 4-channel LUT arrays."""
    T = np.zeros(transmittance_shape(s))
    S = np.zeros(scattering_shape(s))
    M = np.zeros(scattering_shape(s))
    E = np.zeros(irradiance_shape(s))
    rc = lib().or_model_init(C.byref(spec.c), s, orders, ptr(T), ptr(S), ptr(M), ptr(E))
    assert rc == 0
    out = dict(transmittance=T, scattering=S, irradiance=E)
    if not spec.combine_scattering_textures:
        out["single_mie_scattering"] = M
    return out


def model_init_orders(spec, s, orders):
    """FP64 Model::Init(n) for n = 1..orders in one run: a list whose entry
    n - 1 is the dict of LUTs of Init(n)."""
    T = np.zeros(transmittance_shape(s))
    Ss = [np.zeros(scattering_shape(s)) for _ in range(orders)]
    Ms = [np.zeros(scattering_shape(s)) for _ in range(orders)]
    Es = [np.zeros(irradiance_shape(s)) for _ in range(orders)]
    arr = lambda xs: (DP * orders)(*[ptr(x) for x in xs])
    rc = lib().or_model_init_orders(C.byref(spec.c), s, orders, ptr(T), arr(Ss), arr(Ms),
                                    arr(Es))
    assert rc == 0
    out = []
    for n in range(orders):
        d = dict(transmittance=T, scattering=Ss[n], irradiance=Es[n])
        if not spec.combine_scattering_textures:
            d["single_mie_scattering"] = Ms[n]
        out.append(d)
    return out


def precompute(a, s, matrix9, blend, orders, combined, T, S, M, E):
    lib().or_precompute(C.byref(a), s, ptr(np.ascontiguousarray(matrix9, dtype=np.float64)),
                        int(blend), orders, int(combined), ptr(T), ptr(S), ptr(M), ptr(E))
