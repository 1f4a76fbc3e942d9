"""Seeded synthetic Model parameters for the benchmark workloads.

Each builder returns the keyword arguments of Model (the reference's
constructor parameters, atmosphere/model.h:182-281) plus 'lut_sizes' and
'orders'. Randomness comes from a counter-based splitmix64 stream so results
do not depend on numpy's generator versions.

  earth()              Earth, 3 wavelengths, constant solar spectrum, reference
                       LUT sizes, 4 orders (the plain Init workload).
  spectral(half)       47-band spectrum (5778 K blackbody, Rayleigh lambda^-4,
                       Angstrom Mie, ozone tent), 15 precomputed wavelengths.
  random_atmosphere()  one draw of the batched workload: radii, scale heights,
                       coefficients, albedo, per-dimension LUT size factors in
                       {0.5, 1, 2}, 2..8 orders.
  high_resolution()    64x256x64x16 scattering LUT, 512x128 / 128x32, 15
                       wavelengths, 8 orders, FP16 storage.
"""
import math

import numpy as np

DEFAULT_SIZES = dict(transmittance_width=256, transmittance_height=64,
                     scattering_r=32, scattering_mu=128, scattering_mu_s=32,
                     scattering_nu=8, irradiance_width=64, irradiance_height=16)

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): x += 0x9e3779b97f4a7c15, then
    two xor-shift-multiply rounds."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def choice(self, seq):
        return seq[self.next_u64() % len(seq)]

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)


def earth(sun_angular_radius=0.00935, orders=4, lut_sizes=None, half_precision=False):
    """Plain Earth-like atmosphere at 3 wavelengths (680, 550, 440 nm):
    Rayleigh 8 km scale height, beta (5.802, 13.558, 33.1)e-6 /m; Mie 1.2 km,
    scattering 3.996e-6 /m, extinction 4.44e-6 /m, g 0.8; no ozone; albedo 0.1;
    sun zenith <= 102 deg; length unit 1 km; constant solar 1.5 W/m^2/nm."""
    wl = [440.0, 550.0, 680.0]
    return dict(
        wavelengths=wl, solar_irradiance=[1.5] * 3, sun_angular_radius=sun_angular_radius,
        bottom_radius=6360e3, top_radius=6420e3,
        rayleigh_density=[(0.0, 1.0, -1.0 / 8000.0, 0.0, 0.0)],
        rayleigh_scattering=[33.1e-6, 13.558e-6, 5.802e-6],
        mie_density=[(0.0, 1.0, -1.0 / 1200.0, 0.0, 0.0)],
        mie_scattering=[3.996e-6] * 3, mie_extinction=[4.44e-6] * 3,
        mie_phase_function_g=0.8, absorption_density=[], absorption_extinction=[0.0] * 3,
        ground_albedo=[0.1] * 3, max_sun_zenith_angle=102.0 / 180.0 * math.pi,
        length_unit_in_meters=1000.0, num_precomputed_wavelengths=3,
        combine_scattering_textures=False, half_precision=half_precision,
        lut_sizes=dict(lut_sizes or DEFAULT_SIZES), orders=orders)


def planck(lambda_nm, temperature):
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    lam = lambda_nm * 1e-9
    return 2.0 * h * c * c / lam ** 5 / (math.exp(h * c / (lam * k * temperature)) - 1.0)


def ozone_cross_section(lambda_nm):
    """Synthetic smooth ozone absorption cross-section (m^2): a Chappuis-like
    band (peak 4.8e-25 m^2 at 602 nm, width 80 nm) plus a Huggins-like UV rise
    below 360 nm. Not measured data."""
    chappuis = 4.8e-25 * math.exp(-((lambda_nm - 602.0) / 80.0) ** 2)
    huggins = 1.0e-24 * math.exp(-(lambda_nm - 310.0) / 12.0)
    return chappuis + huggins


def spectral(half_precision=False, orders=4, lut_sizes=None,
             num_precomputed_wavelengths=15):
    """47 bands of 10 nm centred at 365..825 nm; 5778 K blackbody solar
    spectrum scaled to a 1.5 W/m^2/nm peak; Rayleigh 1.24062e-6 lambda_um^-4
    /m; Mie Angstrom alpha 0, beta 0.04 over a 1.2 km scale height with single
    scattering albedo 0.9; ozone tent 10-25-40 km at 300 Dobson units."""
    wl = [365.0 + 10.0 * i for i in range(47)]
    bb = [planck(l, 5778.0) for l in wl]
    peak = max(bb)
    solar = [1.5 * b / peak for b in bb]
    ray = [1.24062e-6 * (l * 1e-3) ** -4 for l in wl]
    mie_ext = [0.04 / 1200.0 * (l * 1e-3) ** -0.0 for l in wl]
    mie_sca = [0.9 * m for m in mie_ext]
    dobson = 2.687e20
    max_ozone = 300.0 * dobson / 15000.0
    ozone = [max_ozone * ozone_cross_section(l) for l in wl]
    return dict(
        wavelengths=wl, solar_irradiance=solar, sun_angular_radius=0.00935,
        bottom_radius=6360e3, top_radius=6420e3,
        rayleigh_density=[(0.0, 1.0, -1.0 / 8000.0, 0.0, 0.0)], rayleigh_scattering=ray,
        mie_density=[(0.0, 1.0, -1.0 / 1200.0, 0.0, 0.0)], mie_scattering=mie_sca,
        mie_extinction=mie_ext, mie_phase_function_g=0.8,
        absorption_density=[(25000.0, 0.0, 0.0, 1.0 / 15000.0, -2.0 / 3.0),
                            (0.0, 0.0, 0.0, -1.0 / 15000.0, 8.0 / 3.0)],
        absorption_extinction=ozone, ground_albedo=[0.1] * 47,
        max_sun_zenith_angle=102.0 / 180.0 * math.pi, length_unit_in_meters=1000.0,
        num_precomputed_wavelengths=num_precomputed_wavelengths,
        combine_scattering_textures=False, half_precision=half_precision,
        lut_sizes=dict(lut_sizes or DEFAULT_SIZES), orders=orders)


def scaled_sizes(factors):
    """Reference sizes scaled per dimension (factors: dict name -> factor)."""
    s = {}
    for k, v in DEFAULT_SIZES.items():
        n = int(round(v * factors.get(k, 1.0)))
        if k == "scattering_mu":
            n += n % 2
        s[k] = max(2, n)
    return s


def random_atmosphere(rng):
    """One atmosphere of the batched workload (seeded by the caller)."""
    rb = rng.uniform(3000e3, 7000e3)
    thickness = rng.uniform(40e3, 200e3)
    h_r = rng.uniform(5e3, 15e3)
    beta_scale = rng.uniform(0.3, 3.0)
    h_m = rng.uniform(0.5e3, 3e3)
    g = rng.uniform(0.6, 0.9)
    albedo = rng.uniform(0.0, 0.5)
    factors = {k: rng.choice((0.5, 1.0, 2.0)) for k in DEFAULT_SIZES}
    orders = rng.integer(2, 8)
    kw = earth()
    kw.update(
        bottom_radius=rb, top_radius=rb + thickness,
        rayleigh_density=[(0.0, 1.0, -1.0 / h_r, 0.0, 0.0)],
        rayleigh_scattering=[v * beta_scale for v in kw["rayleigh_scattering"]],
        mie_density=[(0.0, 1.0, -1.0 / h_m, 0.0, 0.0)], mie_phase_function_g=g,
        ground_albedo=[albedo] * 3, lut_sizes=scaled_sizes(factors), orders=orders)
    return kw


def batch_size_classes():
    """One LUT size per tile-shape class of the batched workload: scattering
    MU, MU_S and NU each at x{0.5, 1, 2} (27 classes; the tuner's classes
    are (MU, NU) and (MU, MU_S, NU))."""
    out = []
    for fmu in (0.5, 1.0, 2.0):
        for fmus in (0.5, 1.0, 2.0):
            for fnu in (0.5, 1.0, 2.0):
                out.append(scaled_sizes({"scattering_mu": fmu, "scattering_mu_s": fmus,
                                         "scattering_nu": fnu}))
    return out


def batch(seed, count=64):
    rng = SplitMix64(seed)
    return [random_atmosphere(rng) for _ in range(count)]


HIGH_RES_SIZES = dict(transmittance_width=512, transmittance_height=128,
                      scattering_r=64, scattering_mu=256, scattering_mu_s=64,
                      scattering_nu=16, irradiance_width=128, irradiance_height=32)


def high_resolution(half_precision=True):
    return spectral(half_precision=half_precision, orders=8, lut_sizes=HIGH_RES_SIZES)


def model_kwargs(cfg):
    """Split a config dict into (Model kwargs, orders)."""
    kw = dict(cfg)
    orders = kw.pop("orders")
    return kw, orders


def sky_queries(seed, n, bottom_radius_units, top_radius_units):
    """Seeded query batch in length units: camera altitude U[0, 100] km, view
    direction uniform on the sphere, sun zenith U[0, 100] deg, hit distance
    U[0, 50] km (point clamped inside the atmosphere), normal = local zenith,
    shadow length 0. Returns float32 arrays."""
    rs = np.random.Generator(np.random.PCG64(seed))
    alt = rs.uniform(0.0, 100.0, n)
    cam = np.zeros((n, 3))
    cam[:, 2] = bottom_radius_units + alt
    v = rs.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    zen = np.radians(rs.uniform(0.0, 100.0, n))
    az = rs.uniform(0.0, 2 * np.pi, n)
    sun = np.stack([np.sin(zen) * np.cos(az), np.sin(zen) * np.sin(az), np.cos(zen)], 1)
    dist = rs.uniform(0.0, 50.0, n)
    pt = cam + v * dist[:, None]
    r = np.linalg.norm(pt, axis=1)
    lo, hi = bottom_radius_units + 1e-3, top_radius_units - 1e-3
    pt *= (np.clip(r, lo, hi) / r)[:, None]
    normal = pt / np.linalg.norm(pt, axis=1, keepdims=True)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(camera=f(cam), view_ray=f(v), sun_direction=f(sun), point=f(pt),
                normal=f(normal), shadow_length=np.zeros(n, np.float32))


def splitmix_uniforms(seed, first, count):
    """Outputs first .. first+count-1 (1-based) of SplitMix64(seed) as uniform
    doubles in [0, 1): the n-th output is the finalizer of seed + n * gamma."""
    n = np.arange(first, first + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + n * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def hashed_sky_queries(seed, n, bottom_radius_units, top_radius_units, max_altitude=100.0,
                       max_distance=50.0, max_sun_zenith=math.radians(100.0)):
    """The query batch that atmo_generate_sky_queries makes on the device,
    computed on the host in float64 (for checking the generator and for host
    use): ray i takes SplitMix64 outputs 6i+1 .. 6i+6 as altitude, view z,
    view azimuth, sun zenith, sun azimuth, hit distance."""
    u = splitmix_uniforms(seed, 1, 6 * n).reshape(n, 6)
    alt = max_altitude * u[:, 0]
    vz = 2.0 * u[:, 1] - 1.0
    vphi = 2.0 * math.pi * u[:, 2]
    zen = max_sun_zenith * u[:, 3]
    az = 2.0 * math.pi * u[:, 4]
    dist = max_distance * u[:, 5]
    vs = np.sqrt(np.maximum(0.0, 1.0 - vz * vz))
    v = np.stack([vs * np.cos(vphi), vs * np.sin(vphi), vz], 1)
    cam = np.zeros((n, 3))
    cam[:, 2] = bottom_radius_units + alt
    pt = cam + v * dist[:, None]
    r = np.linalg.norm(pt, axis=1)
    lo, hi = bottom_radius_units + 1e-3, top_radius_units - 1e-3
    pt *= (np.clip(r, lo, hi) / r)[:, None]
    normal = pt / np.linalg.norm(pt, axis=1, keepdims=True)
    sun = np.stack([np.sin(zen) * np.cos(az), np.sin(zen) * np.sin(az), np.cos(zen)], 1)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(camera=f(cam), view_ray=f(v), sun_direction=f(sun), point=f(pt),
                normal=f(normal), shadow_length=np.zeros(n, np.float32))
