"""Helpers comparing GPU LUTs with the FP64 oracle.

Error metric (per texel and channel):
    rel = |gpu - ref| / max(|ref|, FLOOR * max over the texture's channel |ref|)
FLOOR = 1e-6: texels whose reference value is below a millionth of the
channel's largest value (night side, ground-hitting rays, Mie tails) are
compared in absolute terms at that scale. Bars: 1e-3 (FP32 LUTs), 5e-3 (FP16).
"""
import numpy as np

import oracle_lib as ol
from precomputed_atmospheric_scattering_amd import configs

FLOOR = 1e-6
TOL_FP32 = 1e-3
TOL_FP16 = 5e-3

SMALL_SIZES = dict(transmittance_width=64, transmittance_height=16, scattering_r=8,
                   scattering_mu=32, scattering_mu_s=8, scattering_nu=4,
                   irradiance_width=16, irradiance_height=8)
MID_SIZES = dict(transmittance_width=128, transmittance_height=32, scattering_r=16,
                 scattering_mu=64, scattering_mu_s=16, scattering_nu=8,
                 irradiance_width=32, irradiance_height=16)


def rel_err(gpu, ref, channels=3, floor=FLOOR):
    """(max rel error, flat index of the worst texel-channel)."""
    g = np.asarray(gpu, dtype=np.float64)[..., :channels].reshape(-1, channels)
    r = np.asarray(ref, dtype=np.float64)[..., :channels].reshape(-1, channels)
    scale = np.abs(r).max(axis=0)
    denom = np.maximum(np.abs(r), floor * np.maximum(scale, 1e-300))
    e = np.abs(g - r) / denom
    i = int(np.argmax(e))
    return float(e.reshape(-1)[i]), i


def oracle_spec(cfg):
    kw, orders = configs.model_kwargs(cfg)
    sizes = kw.pop("lut_sizes")
    return ol.ModelSpec(**kw), ol.sizes(**sizes), orders


def oracle_init(cfg, orders=None):
    spec, s, o = oracle_spec(cfg)
    return ol.model_init(spec, s, o if orders is None else orders)


def gpu_model(cfg, **extra):
    import precomputed_atmospheric_scattering_amd as atmo
    kw, _ = configs.model_kwargs(cfg)
    kw.update(extra)
    return atmo.Model(**kw)


def compare_init(cfg, orders, tol=TOL_FP32, textures=None, verbose=True):
    """Full Init on the GPU and in the oracle; returns {texture: max rel err}."""
    ref = oracle_init(cfg, orders)
    m = gpu_model(cfg)
    m.init(orders)
    errs = {}
    for name, r in ref.items():
        if textures and name not in textures:
            continue
        g = m.lut(name)
        ch = 4 if name == "scattering" else 3
        errs[name] = rel_err(g.reshape(r.shape), r, channels=ch)[0]
    if verbose:
        print("orders=%d " % orders + " ".join("%s=%.2e" % kv for kv in errs.items()),
              flush=True)
    m.close()
    return errs
