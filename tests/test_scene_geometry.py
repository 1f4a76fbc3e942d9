"""CPU checks of the test-scene geometry (tests/scene.py) and of a render
through the oracle's lookups on small LUTs, so that the GPU image test only
has to compare lookups."""
import numpy as np

import parity
import scene
import test_gpu_scene as tgs
import oracle_lib as ol
from precomputed_atmospheric_scattering_amd import configs

TINY = dict(transmittance_width=32, transmittance_height=8, scattering_r=4,
            scattering_mu=16, scattering_mu_s=4, scattering_nu=4,
            irradiance_width=8, irradiance_height=4)


def test_geometry():
    sc = scene.Scene(6360.0, 0.00935, scene.sun_direction(65.0, 90.0))
    assert sc.sphere_hit.sum() > 1000 and sc.ground_hit.sum() > 10000
    a = sc.sphere_alpha[sc.sphere_hit]
    assert (a > 0).all() and (a <= 1).all()
    # The sphere's umbra cone crosses some view rays (light shafts) and
    # shades some ground points.
    assert (sc.sky_shadow > 0).sum() > 1000
    v = sc.sun_visibility[sc.ground_hit]
    assert v.min() < 0.5 and v.max() == 1.0
    # Sky visibility (view factor of the sphere) lies in (0, 1].
    k = sc.sky_visibility[sc.ground_hit]
    assert k.min() > 0.0 and k.max() <= 1.0
    # View rays are unit vectors; the image centre looks along +y, pitched
    # up by pi / 30 (6 degrees).
    assert np.allclose(np.linalg.norm(sc.view, axis=-1), 1.0)
    c = sc.view[scene.HEIGHT // 2, scene.WIDTH // 2]
    assert abs(np.degrees(np.arcsin(c[2])) - 6.0) < 0.2 and c[1] > 0.99


def test_oracle_render_is_finite():
    cfg = configs.earth(lut_sizes=TINY, orders=2)
    spec, s, orders = parity.oracle_spec(cfg)
    luts = ol.model_init(spec, s, orders)
    sc = scene.Scene(6360.0, 0.00935, scene.sun_direction(65.0, 90.0))
    q = sc.queries()
    img = sc.composite(q, tgs.oracle_lookups(q, spec, s, luts, False), (1.0, 1.0, 1.0),
                       tgs.GROUND_ALBEDO, tgs.SPHERE_ALBEDO)
    assert np.isfinite(img).all() and img.min() >= 0.0
    assert scene.psnr(scene.to_8bit(img, 10.0), scene.to_8bit(img, 10.0)) == float("inf")
