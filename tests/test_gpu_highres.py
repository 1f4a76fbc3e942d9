"""Error against the FP64 reference at the high-resolution sizes (scattering
LUT 64x256x64x16, transmittance 512x128, irradiance 128x32; the spectral
atmosphere with ozone, 15 precomputed wavelengths).

Two checks:

- End to end: a full FP64 Init (8 orders, 5 Precompute iterations) takes
  about 3.6 hours of host time on 6 threads, so it was computed once by
  tests/golden/make_high_resolution_golden.py; the GPU's final LUTs, with
  FP32 and with FP16 storage, are compared with it at its 10000 sampled
  texels of the scattering LUTs and over the whole irradiance LUT. The final
  transmittance LUT is recomputed in full by the oracle.
- Pass by pass: after the GPU's Init(n - 1), its retained delta LUTs (of the
  last Precompute iteration) are fed in FP64 to the oracle's passes, and the
  oracle's order-n values are compared with the GPU's Init(n) deltas: single
  scattering, scattering density at orders 2 and 3 (the order-2 and
  higher-order paths) and multiple scattering at 2000 random texels, the
  indirect irradiance of orders 1 and 2 over the whole irradiance LUT. The
  transmittance the iteration used is recomputed by the oracle (the GPU's
  per-iteration copy is a temporary; the two agree to about 6e-8).

FP16 storage is also compared at full size with the FP32 LUTs of the same
atmosphere."""
import os
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

    # The final transmittance LUT (680, 550, 440 nm), in full.
    m.init(1)
    T_rgb = np.zeros(ol.transmittance_shape(s))
    ol.lib().or_pass_transmittance(C.byref(ol.lower(spec, (680.0, 550.0, 440.0))), s,
                                   ol.ptr(T_rgb))
    errs["transmittance (final, whole LUT)"] = parity.rel_err(
        m.lut("transmittance").reshape(T_rgb.shape), T_rgb)[0]
    # Order 1: single scattering.
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
    # Indirect irradiance of order 1, from the single scattering (whole LUT).
    E1 = np.zeros(ol.irradiance_shape(s))
    ol.lib().or_pass_indirect_irradiance(C.byref(a), s, ol.ptr(dR), ol.ptr(dM), ol.ptr(dR),
                                         1, ol.ptr(E1))
    errs["indirect irradiance, order 1 (whole LUT)"] = parity.rel_err(
        m.lut("delta_irradiance").reshape(E1.shape), E1)[0]
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
    # Indirect irradiance of order 2, from the order-2 multiple scattering.
    E2 = np.zeros(ol.irradiance_shape(s))
    ol.lib().or_pass_indirect_irradiance(C.byref(a), s, ol.ptr(dMS), ol.ptr(dMS),
                                         ol.ptr(dMS), 2, ol.ptr(E2))
    errs["indirect irradiance, order 2 (whole LUT)"] = parity.rel_err(
        m.lut("delta_irradiance").reshape(E2.shape), E2)[0]
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


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "high_resolution_fp64.npz")


def golden_err(gpu, ref, channel_max, floor=parity.FLOOR):
    """parity.rel_err at sampled texels, with the floor taken from the whole
    reference LUT's channel maxima."""
    g = np.asarray(gpu, dtype=np.float64)[:, :ref.shape[1]]
    denom = np.maximum(np.abs(ref), floor * np.maximum(channel_max[:ref.shape[1]], 1e-300))
    return float((np.abs(g - ref) / denom).max())


@pytest.mark.skipif(not os.path.exists(GOLDEN),
                    reason="tests/golden/high_resolution_fp64.npz not generated")
def test_high_resolution_full_chain_against_fp64_golden():
    gold = np.load(GOLDEN)
    idx = gold["texels"]
    lines = ["high resolution, full Init(8) against the FP64 oracle "
             "(10000 sampled scattering texels, whole irradiance and transmittance LUTs):",
             "| storage | transmittance | scattering | single_mie_scattering | irradiance |",
             "|---|---|---|---|---|"]
    worst = []
    T_ref = None
    for half in (False, True):
        cfg = configs.high_resolution(half_precision=half)
        kw, orders = configs.model_kwargs(cfg)
        assert orders == int(gold["orders"])
        m = atmo.Model(**kw)
        m.init(orders)
        if T_ref is None:
            spec, s, _ = parity.oracle_spec(cfg)
            T_ref = np.zeros(ol.transmittance_shape(s))
            ol.lib().or_pass_transmittance(C.byref(ol.lower(spec, (680.0, 550.0, 440.0))),
                                           s, ol.ptr(T_ref))
        e_t = parity.rel_err(m.lut("transmittance").reshape(T_ref.shape), T_ref)[0]
        e_s = golden_err(m.lut("scattering").reshape(-1, 4)[idx], gold["scattering"],
                         gold["scattering_max"])
        e_m = golden_err(m.lut("single_mie_scattering").reshape(-1, 4)[idx],
                         gold["single_mie_scattering"], gold["single_mie_scattering_max"])
        e_e = parity.rel_err(m.lut("irradiance").reshape(-1, 4)[:, :3], gold["irradiance"])[0]
        m.close()
        lines.append("| %s | %.2e | %.2e | %.2e | %.2e |" % (
            "fp16" if half else "fp32", e_t, e_s, e_m, e_e))
        tol = parity.TOL_FP16 if half else parity.TOL_FP32
        worst += [(half, "transmittance", e_t, parity.TOL_FP32),
                  (half, "scattering", e_s, tol), (half, "single_mie_scattering", e_m, tol),
                  (half, "irradiance", e_e, parity.TOL_FP32)]
    print("\n".join(lines), flush=True)
    for half, name, e, tol in worst:
        assert e <= tol, (half, name, e)
