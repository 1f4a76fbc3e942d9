"""Error of every texture against the FP64 oracle for every scattering order
(Init(n), n = 1..4), for the Earth atmosphere (3 wavelengths) and the
spectral atmosphere (15 wavelengths, ozone) at the reference LUT sizes - the
spectral FP64 chain takes about a minute of host time on 16 threads - and
the spectral atmosphere at half those sizes in every dimension but nu, each
with FP32 and FP16 storage. The tables are printed (pytest -s, or the log of -v)
and asserted against the bars: 1e-3 for FP32 textures, 5e-3 for FP16-stored
ones (metric of parity.rel_err, floor 1e-6 of the texture's channel maximum).
"""
import numpy as np
import pytest

import oracle_lib as ol
import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu

ORDERS = 4
NAMES = ("transmittance", "scattering", "single_mie_scattering", "irradiance")
HALF = ("scattering", "single_mie_scattering")


def table(cfg, label):
    spec, s, _ = parity.oracle_spec(cfg)
    snaps = ol.model_init_orders(spec, s, ORDERS)
    lines = ["%s: max relative error per texture (rows: scattering orders)" % label,
             "| orders | storage | " + " | ".join(NAMES) + " |",
             "|---|---|" + "---|" * len(NAMES)]
    worst = {}
    for half in (False, True):
        c = dict(cfg)
        c["half_precision"] = half
        kw, _ = configs.model_kwargs(c)
        m = atmo.Model(**kw)
        for n in range(1, ORDERS + 1):
            m.init(n)
            row = []
            for name in NAMES:
                r = snaps[n - 1][name]
                g = m.lut(name).reshape(r.shape)
                e = parity.rel_err(g, r, channels=4 if name == "scattering" else 3)[0]
                tol = parity.TOL_FP16 if (half and name in HALF) else parity.TOL_FP32
                worst[(half, n, name)] = (e, tol)
                row.append("%.2e" % e)
            lines.append("| %d | %s | %s |" % (n, "fp16" if half else "fp32", " | ".join(row)))
        m.close()
    print("\n".join(lines), flush=True)
    for key, (e, tol) in worst.items():
        assert e <= tol, (label, key, e)


def test_earth_reference_sizes():
    table(configs.earth(), "Earth, 3 wavelengths, LUTs 256x64 / 32x128x32x8 / 64x16")


def test_spectral_half_sizes():
    table(configs.spectral(lut_sizes=parity.MID_SIZES),
          "spectral, 15 wavelengths, LUTs 128x32 / 16x64x16x8 / 32x16")


def test_spectral_reference_sizes():
    table(configs.spectral(),
          "spectral, 15 wavelengths, LUTs 256x64 / 32x128x32x8 / 64x16")
