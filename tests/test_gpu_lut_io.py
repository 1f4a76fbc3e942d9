"""LUT files: export after Init, import into a fresh model, cached Init
(the reference CPU model's cache directory). Bars: imported LUTs equal the
exported ones bit for bit (FP16 models re-encode to the same FP16 texels and
scales), and lookups on an imported model equal those on the original."""
import os

import numpy as np
import pytest

import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu

NAMES = ("transmittance", "scattering", "single_mie_scattering", "irradiance")


def model(half=False, combined=False):
    cfg = configs.earth(lut_sizes=parity.SMALL_SIZES, orders=3, half_precision=half)
    cfg["combine_scattering_textures"] = combined
    kw, orders = configs.model_kwargs(cfg)
    return atmo.Model(**kw), orders


@pytest.mark.parametrize("half", [False, True])
def test_export_import_round_trip(tmp_path, half):
    m, orders = model(half)
    m.init(orders)
    m.export_luts(tmp_path)
    files = sorted(os.listdir(tmp_path))
    assert files == ["irradiance.dat", "scattering.dat", "single_mie_scattering.dat",
                     "transmittance.dat"]
    # Raw float32 RGBA, x fastest: the file is the decoded LUT.
    raw = np.fromfile(tmp_path / "scattering.dat", dtype=np.float32)
    assert np.array_equal(raw, m.lut("scattering").reshape(-1))
    m2, _ = model(half)
    m2.import_luts(tmp_path)
    for n in NAMES:
        assert np.array_equal(m2.lut(n), m.lut(n)), n
    if half:
        assert list(m2.lut_view("scattering").scale) == list(m.lut_view("scattering").scale)
    n = 4096
    q = {k: atmo.DeviceArray.from_host(v)
         for k, v in configs.sky_queries(3, n, 6360.0, 6420.0).items()}
    outs = []
    for mm in (m, m2):
        o = [atmo.DeviceArray((n, 3)) for _ in range(2)]
        mm.sky_radiance_device(n, q["camera"], q["view_ray"], q["shadow_length"],
                               q["sun_direction"], o[0], o[1])
        outs.append([a.to_host() for a in o])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    m.close()
    m2.close()


def test_combined_textures_export_has_no_single_mie(tmp_path):
    m, orders = model(combined=True)
    m.init(orders)
    m.export_luts(tmp_path)
    assert not (tmp_path / "single_mie_scattering.dat").exists()
    m2, _ = model(combined=True)
    m2.import_luts(tmp_path)
    assert np.array_equal(m2.lut("scattering"), m.lut("scattering"))


def test_init_cached(tmp_path):
    m, orders = model()
    m.init_cached(orders, tmp_path)          # computes and saves
    assert (tmp_path / "transmittance.dat").exists()
    m2, _ = model()
    m2.init_cached(orders, tmp_path)         # loads
    for n in NAMES:
        assert np.array_equal(m2.lut(n), m.lut(n)), n


def test_import_size_mismatch_is_an_error(tmp_path):
    m, orders = model()
    m.init(orders)
    m.export_luts(tmp_path)
    cfg = configs.earth(lut_sizes=parity.MID_SIZES)
    big = atmo.Model(**configs.model_kwargs(cfg)[0])
    with pytest.raises(atmo.AtmoError, match="bytes"):
        big.import_luts(tmp_path)
    with pytest.raises(atmo.AtmoError):
        big.import_luts(tmp_path / "missing")
