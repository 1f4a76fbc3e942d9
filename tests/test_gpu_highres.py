"""Error against the FP64 reference at the high-resolution sizes (scattering
LUT 64x256x64x16, transmittance 512x128, irradiance 128x32; the spectral
atmosphere with ozone, 15 precomputed wavelengths).

A full FP64 Init at these sizes takes about 40 minutes on 16 host threads,
so the check is pass-local: after the GPU's Init(n - 1), its retained delta
LUTs (of the last Precompute iteration) are fed in FP64 to the oracle's
passes at 2000 random texels, and the oracle's order-n values are compared
with the GPU's Init(n) deltas there. Each pass is checked on the GPU's own
inputs, so every pass's error is measured at full size: single scattering,
scattering density at orders 2 and 3 (the order-2 and higher-order paths),
and multiple scattering. The transmittance the iteration used is recomputed
by the oracle (the GPU's per-iteration copy is a temporary; the two agree to
about 6e-8). FP16 storage is checked at full size, with all 8 orders,
against the FP32 LUTs of the same atmosphere."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs

pytestmark = pytest.mark.gpu

N_SAMPLE = 2000


def f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def test_high_resolution_passes_match_oracle():
    cfg = configs.high_resolution(half_precision=False)
    spec, s, _ = parity.oracle_spec(cfg)
    kw, _ = configs.model_kwargs(cfg)
    m = atmo.Model(retain_intermediates=True, **kw)
    last = m.params.num_iterations() - 1
    lambdas, _ = m.params.iteration(last)
    a = ol.lower(spec, lambdas)
    T = np.zeros(ol.transmittance_shape(s))
    ol.lib().or_pass_transmittance(C.byref(a), s, ol.ptr(T))
    n_s = s.scattering_r * s.scattering_mu * s.scattering_mu_s * s.scattering_nu
    rng = np.random.Generator(np.random.PCG64(41))
    idx = np.ascontiguousarray(rng.integers(0, n_s, N_SAMPLE), dtype=np.int32)
    ip = idx.ctypes.data_as(C.POINTER(C.c_int))
    errs = {}

    def at(name):
        return m.lut(name).reshape(-1, 4)[idx]

    # Order 1: single scattering.
    m.init(1)
    ray, mie = np.zeros((N_SAMPLE, 4)), np.zeros((N_SAMPLE, 4))
    ol.lib().or_texels_single_scattering(C.byref(a), s, ol.ptr(T), N_SAMPLE, ip,
                                         ol.ptr(ray), ol.ptr(mie))
    errs["single scattering (Rayleigh)"] = parity.rel_err(at("delta_rayleigh"), ray)[0]
    errs["single scattering (Mie)"] = parity.rel_err(at("delta_mie"), mie)[0]
    dR, dM = f64(m.lut("delta_rayleigh")), f64(m.lut("delta_mie"))
    dE = f64(m.lut("delta_irradiance"))
    # Order 2: density from the order-1 deltas, multiple scattering from the
    # GPU's density.
    m.init(2)
    J = np.zeros((N_SAMPLE, 4))
    ol.lib().or_texels_scattering_density(C.byref(a), s, ol.ptr(T), ol.ptr(dR), ol.ptr(dM),
                                          ol.ptr(dR), ol.ptr(dE), 2, N_SAMPLE, ip,
                                          ol.ptr(J))
    errs["scattering density, order 2"] = parity.rel_err(at("delta_scattering_density"),
                                                         J)[0]
    del dR, dM
    dJ = f64(m.lut("delta_scattering_density"))
    ms = np.zeros((N_SAMPLE, 4))
    ol.lib().or_texels_multiple_scattering(C.byref(a), s, ol.ptr(T), ol.ptr(dJ), N_SAMPLE, ip,
                                           ol.ptr(ms))
    errs["multiple scattering, order 2"] = parity.rel_err(at("delta_rayleigh"), ms)[0]
    del dJ
    # Order 3: density from the order-2 multiple scattering and the order-1
    # indirect irradiance (the higher-order path of the kernel).
    dMS, dE1 = f64(m.lut("delta_rayleigh")), f64(m.lut("delta_irradiance"))
    m.init(3)
    J3 = np.zeros((N_SAMPLE, 4))
    ol.lib().or_texels_scattering_density(C.byref(a), s, ol.ptr(T), ol.ptr(dMS), ol.ptr(dMS),
                                          ol.ptr(dMS), ol.ptr(dE1), 3, N_SAMPLE, ip,
                                          ol.ptr(J3))
    errs["scattering density, order 3"] = parity.rel_err(at("delta_scattering_density"),
                                                         J3)[0]
    m.close()
    for name, e in errs.items():
        print("high resolution, %s: max rel err %.2e" % (name, e), flush=True)
    for name, e in errs.items():
        assert e <= parity.TOL_FP32, (name, e)


def test_high_resolution_fp16_storage():
    """FP16 storage (FP32 accumulation, per-channel scales) against the FP32
    LUTs of the same atmosphere, at full size and 8 orders: the error FP16
    adds to the FP32 error bounded above."""
    out = {}
    for half in (False, True):
        cfg = configs.high_resolution(half_precision=half)
        kw, orders = configs.model_kwargs(cfg)
        m = atmo.Model(**kw)
        m.init(orders)
        out[half] = {k: m.lut(k) for k in ("scattering", "single_mie_scattering")}
        m.close()
    for k in out[False]:
        e = parity.rel_err(out[True][k], out[False][k])[0]
        print("high resolution, FP16 %s vs FP32: max rel err %.2e" % (k, e), flush=True)
        assert e <= parity.TOL_FP16 - parity.TOL_FP32, (k, e)
