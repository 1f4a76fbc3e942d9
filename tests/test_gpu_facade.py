"""The C++ facade (include/atmo_model.h) as a C++ caller uses it: the client
tests/cpp/facade_client.cc builds its model with the constructor call of the
reference's demo (atmosphere/demo/demo.cc:278-284), runs Init(4) and one
batch of each lookup, and writes what it got. Its LUTs, lookups, solar
radiance and spectrum conversion must equal, bit for bit, those of the same
model built and run through the Python path (the C ABI by ctypes) with the
parameters the client reports."""
import os
import subprocess

import numpy as np
import pytest

import facade_build
import precomputed_atmospheric_scattering_amd as atmo

pytestmark = pytest.mark.gpu

ORDERS, N_RAYS, SEED = 4, 4096, 77
LUTS = ("transmittance", "scattering", "single_mie_scattering", "irradiance")


def read_params(path):
    rows = {}
    for line in open(path):
        name, *vals = line.split()
        rows[name] = [float(v) for v in vals]
    (sun_r, rb, rt, g, max_sz, unit, r_scale, m_scale, oz_a, oz_b, oz_c,
     oz_d) = rows["scalars"]
    return dict(
        wavelengths=rows["wavelengths"], solar_irradiance=rows["solar_irradiance"],
        sun_angular_radius=sun_r, bottom_radius=rb, top_radius=rt,
        rayleigh_density=[(0.0, 1.0, r_scale, 0.0, 0.0)],
        rayleigh_scattering=rows["rayleigh_scattering"],
        mie_density=[(0.0, 1.0, m_scale, 0.0, 0.0)],
        mie_scattering=rows["mie_scattering"], mie_extinction=rows["mie_extinction"],
        mie_phase_function_g=g,
        absorption_density=[(25000.0, 0.0, 0.0, oz_a, oz_b), (0.0, 0.0, 0.0, oz_c, oz_d)],
        absorption_extinction=rows["absorption_extinction"],
        ground_albedo=rows["ground_albedo"], max_sun_zenith_angle=max_sz,
        length_unit_in_meters=unit, num_precomputed_wavelengths=3,
        combine_scattering_textures=False, half_precision=False)


def test_cpp_facade_client_equals_python_path(tmp_path):
    exe = facade_build.client()
    print("case: C++ facade client, Init(%d), %d rays" % (ORDERS, N_RAYS), flush=True)
    r = subprocess.run([exe, str(tmp_path), str(ORDERS), str(N_RAYS), str(SEED)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    kw = read_params(tmp_path / "params.txt")
    m = atmo.Model(**kw)
    m.init(ORDERS)
    for name in LUTS:
        got = np.fromfile(tmp_path / (name + ".f32"), dtype=np.float32)
        assert np.array_equal(got, m.lut(name).reshape(-1)), name
    unit = kw["length_unit_in_meters"]
    q = atmo.generate_sky_queries(SEED, N_RAYS, kw["bottom_radius"] / unit,
                                  kw["top_radius"] / unit)
    outs = [atmo.DeviceArray((N_RAYS, 3)) for _ in range(6)]
    m.sky_radiance_device(N_RAYS, q["camera"], q["view_ray"], q["shadow_length"],
                          q["sun_direction"], outs[0], outs[1])
    m.sky_radiance_to_point_device(N_RAYS, q["camera"], q["point"], q["shadow_length"],
                                   q["sun_direction"], outs[2], outs[3])
    m.sun_and_sky_irradiance_device(N_RAYS, q["point"], q["normal"], q["sun_direction"],
                                    outs[4], outs[5])
    atmo.synchronize()
    got = np.fromfile(tmp_path / "lookups.f32", dtype=np.float32).reshape(6, N_RAYS, 3)
    for k in range(6):
        assert np.array_equal(got[k], outs[k].to_host().reshape(N_RAYS, 3)), k
    solar = np.fromfile(tmp_path / "solar.f32", dtype=np.float32)
    assert np.array_equal(solar[:3], m.solar_radiance())
    rgb = atmo.Model.ConvertSpectrumToLinearSrgb(kw["wavelengths"], kw["solar_irradiance"])
    assert np.array_equal(solar[3:], np.float32(rgb))
    m.close()
