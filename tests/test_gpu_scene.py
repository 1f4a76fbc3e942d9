"""Image-level test: the reference's test scene (a 1 km sphere on the
planet, soft shadows and light shafts; atmosphere/reference/model_test.glsl,
model_test.cc:806-870), rendered at 640x360 from the GPU lookups on the GPU
LUTs and from the FP64 oracle's lookups on the oracle's full-chain LUTs, for
the reference's radiance test cases. The bars are the reference's own PSNR
thresholds for its GPU-vs-CPU comparisons (47, 46 and 40 dB; model_test.cc:
820-866). The geometry is shared (the package's frames module), so the
renders differ only by the lookups and the LUTs behind them. Constant
albedos (ground 0.1, sphere 0.8), as in the reference's constant-albedo
cases. The demo's nine preset views (demo.cc:471-489, including the two
from space) are rendered the same way at 192x108 against the lowest bar."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as ol
import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import configs
from precomputed_atmospheric_scattering_amd import frames

pytestmark = pytest.mark.gpu

GROUND_ALBEDO = (0.1, 0.1, 0.1)
SPHERE_ALBEDO = (0.8, 0.8, 0.8)
EXPOSURE = 10.0  # radiance images (model_test.cc:471)


@pytest.fixture(scope="module")
def oracle_luts():
    cfg = configs.earth()
    spec, s, orders = parity.oracle_spec(cfg)
    luts = ol.model_init(spec, s, orders)
    return cfg, spec, s, luts


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def oracle_lookups(q, spec, s, luts, combined):
    a = ol.lower(spec, (680.0, 550.0, 440.0))
    T, S, E = luts["transmittance"], luts["scattering"], luts["irradiance"]
    M = None if combined else luts["single_mie_scattering"]
    mp = ol.ptr(M) if M is not None else None
    L = ol.lib()
    out = {}
    sky = q["sky"]
    n = len(sky["camera"])
    r, t = np.zeros((n, 3)), np.zeros((n, 3))
    L.or_sky_radiance(C.byref(a), s, ol.ptr(T), ol.ptr(S), mp, int(combined), n,
                      ol.ptr(f64(sky["camera"])), ol.ptr(f64(sky["view"])),
                      ol.ptr(f64(sky["shadow"])), ol.ptr(f64(sky["sun"])), ol.ptr(r), ol.ptr(t))
    out["sky"] = (r, t)
    for name in ("sphere", "ground"):
        g = q[name]
        k = len(g["camera"])
        sun, skyi = np.zeros((k, 3)), np.zeros((k, 3))
        L.or_sun_and_sky_irradiance(C.byref(a), s, ol.ptr(T), ol.ptr(E), k,
                                    ol.ptr(f64(g["point"])), ol.ptr(f64(g["normal"])),
                                    ol.ptr(f64(g["sun"])), ol.ptr(sun), ol.ptr(skyi))
        r, t = np.zeros((k, 3)), np.zeros((k, 3))
        L.or_sky_radiance_to_point(C.byref(a), s, ol.ptr(T), ol.ptr(S), mp, int(combined), k,
                                   ol.ptr(f64(g["camera"])), ol.ptr(f64(g["point"])),
                                   ol.ptr(f64(g["shadow"])), ol.ptr(f64(g["sun"])),
                                   ol.ptr(r), ol.ptr(t))
        out[name] = (sun, skyi, r, t)
    return out


def gpu_lookups(q, m):
    out = {"sky": m.sky_radiance(q["sky"]["camera"], q["sky"]["view"], q["sky"]["shadow"],
                                 q["sky"]["sun"])}
    for name in ("sphere", "ground"):
        g = q[name]
        sun, sky = m.sun_and_sky_irradiance(g["point"], g["normal"], g["sun"])
        r, t = m.sky_radiance_to_point(g["camera"], g["point"], g["shadow"], g["sun"])
        out[name] = (sun, sky, r, t)
    return out


@pytest.mark.parametrize("name,combined,sun_theta,bar", [
    ("radiance, separate textures", False, 65.0, 47.0),
    ("radiance, combined textures", True, 65.0, 46.0),
    ("radiance, combined textures, sunset", True, 88.0, 40.0),
])
def test_scene_psnr(oracle_luts, name, combined, sun_theta, bar):
    cfg, spec, s, luts = oracle_luts
    c = dict(cfg)
    c["combine_scattering_textures"] = combined
    kw, orders = configs.model_kwargs(c)
    m = atmo.Model(**kw)
    m.init(orders)
    sc = frames.Frame(kw["bottom_radius"] / 1000.0, kw["sun_angular_radius"],
                      frames.sun_direction(math.radians(sun_theta), math.pi / 2.0),
                      frames.Camera.image_test())
    q = sc.queries()
    solar = m.solar_radiance()
    img_gpu = sc.composite(q, gpu_lookups(q, m), solar, GROUND_ALBEDO, SPHERE_ALBEDO)
    img_ref = sc.composite(q, oracle_lookups(q, spec, s, luts, combined), solar,
                           GROUND_ALBEDO, SPHERE_ALBEDO)
    assert np.isfinite(img_gpu).all() and np.isfinite(img_ref).all()
    p = frames.psnr(frames.to_8bit(img_gpu, EXPOSURE), frames.to_8bit(img_ref, EXPOSURE))
    lin = np.abs(img_gpu - img_ref).max() / np.abs(img_ref).max()
    print("%s: PSNR %.1f dB (bar %.0f), max linear difference %.2e of the image maximum"
          % (name, p, bar, lin), flush=True)
    assert p > bar
    m.close()


@pytest.fixture(scope="module")
def gpu_model(oracle_luts):
    cfg = oracle_luts[0]
    kw, orders = configs.model_kwargs(cfg)
    m = atmo.Model(**kw)
    m.init(orders)
    yield m
    m.close()


@pytest.mark.parametrize("view", sorted(frames.DEMO_VIEWS))
def test_demo_view_psnr(oracle_luts, gpu_model, view):
    cfg, spec, s, luts = oracle_luts
    kw, _ = configs.model_kwargs(cfg)
    fr, exposure = frames.demo_frame(view, kw["bottom_radius"] / 1000.0,
                                     kw["sun_angular_radius"], 192, 108)
    q = fr.queries()
    solar = gpu_model.solar_radiance()
    img_gpu = fr.composite(q, gpu_lookups(q, gpu_model), solar, GROUND_ALBEDO,
                           SPHERE_ALBEDO)
    img_ref = fr.composite(q, oracle_lookups(q, spec, s, luts, False), solar,
                           GROUND_ALBEDO, SPHERE_ALBEDO)
    assert np.isfinite(img_gpu).all() and np.isfinite(img_ref).all()
    p = frames.psnr(frames.to_8bit(img_gpu, exposure), frames.to_8bit(img_ref, exposure))
    print("demo view %d: PSNR %.1f dB, %d sphere and %d ground pixels"
          % (view, p, len(q["sphere"]["index"]), len(q["ground"]["index"])), flush=True)
    assert p > 40.0
