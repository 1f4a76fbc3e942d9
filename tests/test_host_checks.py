"""Host-side argument checks of the Python mirror that must fire before any
device call (no GPU needed): the lookup kernels index their arrays with n
only, so a short or mistyped array must be rejected on the host."""
import ctypes as C

import numpy as np
import pytest

import precomputed_atmospheric_scattering_amd as atmo


class FakeDeviceArray(atmo.DeviceArray):
    """A DeviceArray-shaped record with no device memory behind it: the
    checks under test must reject it before its pointer is used."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p(0x1000)

    def free(self):
        self.ptr = None


def fake_model():
    m = atmo.Model.__new__(atmo.Model)
    m._handle = None
    return m


def sky_args(n, short=None, dtype=None):
    shapes = [(n, 3), (n, 3), (n,), (n, 3), (n, 3), (n, 3)]
    arrs = []
    for i, s in enumerate(shapes):
        if i == short:
            s = (s[0] - 1,) + s[1:]
        arrs.append(FakeDeviceArray(s, dtype if i == short and dtype else np.float32))
    return arrs


@pytest.mark.parametrize("method", ["sky_radiance_device", "sky_radiance_to_point_device",
                                    "sky_luminance_device", "sky_luminance_to_point_device"])
@pytest.mark.parametrize("short", range(6))
def test_sky_lookup_rejects_short_arrays(method, short):
    m = fake_model()
    with pytest.raises(ValueError, match="bytes"):
        getattr(m, method)(16, *sky_args(16, short=short))


@pytest.mark.parametrize("method", ["sun_and_sky_irradiance_device",
                                    "sun_and_sky_illuminance_device"])
@pytest.mark.parametrize("short", range(5))
def test_sun_sky_lookup_rejects_short_arrays(method, short):
    m = fake_model()
    arrs = [FakeDeviceArray((16 - (i == short), 3)) for i in range(5)]
    with pytest.raises(ValueError, match="bytes"):
        getattr(m, method)(16, *arrs)


def test_lookup_rejects_wrong_dtype_negative_n_and_host_arrays():
    m = fake_model()
    args = sky_args(8)
    args[1] = FakeDeviceArray((8, 3), np.float64)   # large enough, wrong type
    with pytest.raises(ValueError, match="float32"):
        m.sky_radiance_device(8, *args)
    with pytest.raises(ValueError, match="non-negative"):
        m.sky_radiance_device(-1, *sky_args(8))
    with pytest.raises(ValueError, match="non-negative"):
        m.sky_radiance_device(2.5, *sky_args(8))
    host = [np.zeros((8, 3), np.float32)] + sky_args(8)[1:]
    with pytest.raises(ValueError, match="DeviceArray"):
        m.sky_radiance_device(8, *host)
    # A larger array than needed is fine as far as the check goes.
    big = sky_args(9)
    assert atmo._check_device_arrays(8, big, atmo._SKY_INPUTS + (3, 3), "t") == 8


@pytest.mark.parametrize("method,arrays", [
    ("sky_radiance", [(4, 3), (4, 3), (3,), (4, 3)]),
    ("sky_radiance_to_point", [(4, 3), (5, 3), (4,), (4, 3)]),
    ("sky_luminance", [(4, 3), (4, 3), (4,), (4, 2)]),
    ("sun_and_sky_irradiance", [(4, 3), (4, 3), (3, 3)]),
    ("sun_and_sky_illuminance", [(4, 3), (4,), (4, 3)]),
])
def test_host_wrappers_require_one_leading_length(method, arrays):
    m = fake_model()
    with pytest.raises(ValueError, match="shape"):
        getattr(m, method)(*[np.zeros(s, np.float32) for s in arrays])


def test_server_submit_lookup_checks_arrays():
    srv = atmo.Server.__new__(atmo.Server)
    srv.handle = None
    srv._keep = {}
    a = sky_args(10)
    with pytest.raises(ValueError, match="bytes"):
        srv.submit_lookup(1, "sky", 10, a[:4], sky_args(10, short=4)[4:])
    with pytest.raises(ValueError, match="inputs"):
        srv.submit_lookup(1, "sun_and_sky", 10, a[:4], a[4:])
    with pytest.raises(ValueError, match="bytes"):
        srv.submit_lookup(1, "sun_and_sky", 11, [FakeDeviceArray((11, 3))] * 2 +
                          [FakeDeviceArray((10, 3))], [FakeDeviceArray((11, 3))] * 2)
    srv.handle = None
