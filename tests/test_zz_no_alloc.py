"""The allocation-free Init (atmo_model_init_static_scratch) against Init.

Two calls in a row share the GPU's static scratch: an Earth-like model at the
reference LUT sizes (FP32, 4 orders; 46.2 MB of temporaries), then a smaller
one (quarter sizes, FP16 storage, 3 orders), which runs on the first call's
leftovers. Each must give the LUTs and lookups of atmo_model_init bit for bit.
A model whose temporaries exceed the scratch is rejected.
"""
import ctypes as C

import numpy as np
import pytest

import parity
import precomputed_atmospheric_scattering_amd as atmo
from precomputed_atmospheric_scattering_amd import _capi, configs

pytestmark = pytest.mark.gpu

LUTS = ("transmittance", "scattering", "single_mie_scattering", "irradiance")


def static_init(m, orders, stream):
    st = atmo.lib().atmo_model_init_static_scratch(m.handle, int(orders), stream.handle)
    return st, atmo.lib().atmo_last_error().decode()


def lookups(m):
    q = configs.sky_queries(5, 4096, m.params.bottom_radius / 1000.0,
                            m.params.top_radius / 1000.0)
    return m.sky_radiance(q["camera"], q["view_ray"], q["shadow_length"],
                          q["sun_direction"])


def scratch_bytes(m):
    res, scr = m.footprint()
    return scr


@pytest.mark.parametrize("first,second", [
    ((None, False, 4), (parity.SMALL_SIZES, True, 3)),
])
def test_static_scratch_init_equals_init_twice(first, second):
    stream = atmo.Stream()
    for sizes, half, orders in (first, second):
        print("static-scratch Init: sizes %s half %s orders %d" % (sizes, half, orders),
              flush=True)
        extra = dict(half_precision=half)
        if sizes is not None:
            extra["lut_sizes"] = sizes
        ref = parity.gpu_model(configs.earth(), **extra)
        got = parity.gpu_model(configs.earth(), **extra)
        assert 0 < scratch_bytes(got) <= _capi.STATIC_SCRATCH_BYTES
        ref.init(orders, stream)
        st, msg = static_init(got, orders, stream)
        assert st == _capi.ATMO_OK, msg
        stream.synchronize()
        for name in LUTS:
            a, b = ref.lut(name), got.lut(name)
            assert np.array_equal(a, b), name
        ra, ta = lookups(ref)
        rb, tb = lookups(got)
        assert np.array_equal(ra, rb) and np.array_equal(ta, tb)
        # The model still works with the allocating Init afterwards.
        got.init(orders, stream)
        stream.synchronize()
        assert np.array_equal(ref.lut("scattering"), got.lut("scattering"))
        ref.close()
        got.close()


def test_static_scratch_rejects_large_temporaries():
    # 15 precomputed wavelengths: 5 iterations of temporaries, 568 MB.
    m = parity.gpu_model(configs.spectral(half_precision=False))
    assert scratch_bytes(m) > _capi.STATIC_SCRATCH_BYTES
    stream = atmo.Stream()
    st, msg = static_init(m, 2, stream)
    assert st == 1 and "static scratch" in msg, (st, msg)
    m.close()
