"""Pins the FP64 oracle against the closed forms and invariants of the
reference's own unit tests (atmosphere/reference/functions_test.cc).

Each test names the reference test it restates. The synthetic planet is the
one of functions_test.cc:49-61 and :296-313, in km: bottom radius 1000, top
1500, Rayleigh scale height 60 (beta 1e-3/km), Mie scale height 30 (beta
1.5e-3/km scattering, 2e-3/km extinction), albedo 0.1, solar 123, mu_s_min -1.
All three channels carry the same values (the reference checks channel 0 of a
47-sample spectrum). Tolerances are the reference's (kEpsilon = 1e-3, 1 m).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as ol

EPS = 1e-3
R_B, R_T = 1000.0, 1500.0
H_R, H_M = 60.0, 30.0
BETA_R, BETA_M, BETA_M_EXT = 0.001, 0.0015, 0.002
ALBEDO = 0.1
SOLAR = 123.0
METER = 1e-3  # km


def planet(uniform=False, no_aerosols=False):
    a = ol.Atmosphere()
    for c in range(3):
        a.solar_irradiance[c] = SOLAR
        a.rayleigh_scattering[c] = BETA_R
        a.mie_scattering[c] = 0.0 if no_aerosols else BETA_M
        a.mie_extinction[c] = 0.0 if no_aerosols else BETA_M_EXT
        a.ground_albedo[c] = ALBEDO
    a.sun_angular_radius = 0.00935 / 2.0
    a.bottom_radius, a.top_radius = R_B, R_T
    if uniform:  # SetUniformAtmosphere, functions_test.cc:1452-1461
        a.rayleigh_density[1] = ol.layer(constant_term=1.0)
        a.mie_density[1] = ol.layer(constant_term=1.0)
    else:
        a.rayleigh_density[1] = ol.layer(exp_term=1.0, exp_scale=-1.0 / H_R)
        a.mie_density[1] = ol.layer(exp_term=1.0, exp_scale=-1.0 / H_M)
    a.mie_phase_function_g = 0.8
    a.mu_s_min = -1.0
    return a


def horizon_mu(r):  # CosineOfHorizonZenithAngle, functions_test.cc:69-72
    return -math.sqrt(1.0 - (R_B / r) ** 2)


S = ol.sizes()  # reference compile-time sizes (constants.h:47-61)
L = ol.lib


def out3():
    return (C.c_double * 3)()


def t_lut(a, s=S):
    T = np.zeros(ol.transmittance_shape(s))
    L().or_pass_transmittance(C.byref(a), s, ol.ptr(T))
    return T


def lazy_scattering(a, s, fill, r, mu, mu_s, nu, ground):
    """LUT holding only the texels a GetScattering(r, mu, mu_s, nu) lookup reads
    (the lazy fixtures of functions_test.cc:82-287, evaluated on demand).
    fill(indices) -> (n, 4) texel values. Other texels are NaN so that any
    read outside the tap set shows up."""
    uvwz = (C.c_double * 4)()
    L().or_scattering_uvwz_from_rmumusnu(C.byref(a), s, r, mu, mu_s, nu, int(ground), uvwz)
    u_nu, u_mu_s, u_mu, u_r = list(uvwz)
    nu_n, mus_n, mu_n, r_n = s.scattering_nu, s.scattering_mu_s, s.scattering_mu, s.scattering_r
    w = nu_n * mus_n

    def taps(u, n):
        x = u * n - 0.5
        f = math.floor(x)
        return {min(max(f, 0), n - 1), min(max(f + 1, 0), n - 1)}

    tx = math.floor(u_nu * (nu_n - 1))
    xs = set()
    for t in (tx, tx + 1):
        xs |= taps((t + u_mu_s) / nu_n, w)
    idx = sorted(x + w * (y + mu_n * z) for x in xs for y in taps(u_mu, mu_n)
                 for z in taps(u_r, r_n))
    lut = np.full(ol.scattering_shape(s), np.nan)
    vals = fill(np.array(idx, dtype=np.int32))
    flat = lut.reshape(-1, 4)
    flat[idx] = vals
    return lut


def get_scattering(a, s, lut, r, mu, mu_s, nu, ground):
    o = (C.c_double * 4)()
    L().or_get_scattering(C.byref(a), s, ol.ptr(lut), r, mu, mu_s, nu, int(ground), o)
    return list(o)


# functions_test.cc:321-333
def test_distance_to_top_atmosphere_boundary():
    a = planet()
    r = R_B * 0.2 + R_T * 0.8
    assert abs(L().or_distance_to_top(C.byref(a), r, 1.0) - (R_T - r)) < METER
    assert abs(L().or_distance_to_top(C.byref(a), r, 0.0) -
               math.sqrt(R_T * R_T - r * r)) < METER


# functions_test.cc:343-352
def test_ray_intersects_ground():
    a = planet()
    r = R_B * 0.9 + R_T * 0.1
    mh = horizon_mu(r)
    assert not L().or_ray_intersects_ground(C.byref(a), r, 1.0)
    assert not L().or_ray_intersects_ground(C.byref(a), r, mh + EPS)
    assert L().or_ray_intersects_ground(C.byref(a), r, mh - EPS)
    assert L().or_ray_intersects_ground(C.byref(a), r, -1.0)


# functions_test.cc:370-388
def test_compute_optical_length_to_top_atmosphere_boundary():
    a = planet()
    r = R_B * 0.2 + R_T * 0.8
    h_r, h_top = r - R_B, R_T - R_B
    expected = H_R * (math.exp(-h_r / H_R) - math.exp(-h_top / H_R))
    got = L().or_optical_length(C.byref(a), a.rayleigh_density, r, 1.0)
    assert abs(got - expected) < METER
    a = planet(uniform=True)
    got = L().or_optical_length(C.byref(a), a.rayleigh_density, r, 0.0)
    assert abs(got - math.sqrt(R_T * R_T - r * r)) < METER


# functions_test.cc:396-419
def test_get_profile_density():
    dens = lambda p, h: L().or_profile_density(p, h)
    p = ol.profile(ol.layer(exp_term=1.0, exp_scale=-1.0))
    assert dens(p, 2.0) == math.exp(-2.0)
    p = ol.profile(ol.layer(linear_term=-0.5, constant_term=1.0))
    assert dens(p, 0.0) == 1.0
    assert dens(p, 1.0) == 0.5
    assert dens(p, 3.0) == 0.0
    p = ol.profile(ol.layer(25.0, 0.0, 0.0, 1.0 / 15.0, -2.0 / 3.0),
                   ol.layer(0.0, 0.0, 0.0, -1.0 / 15.0, 8.0 / 3.0))
    assert dens(p, 0.0) == 0.0
    assert abs(dens(p, 10.0)) < EPS
    assert abs(dens(p, 25.0) - 1.0) < EPS
    assert abs(dens(p, 40.0)) < EPS
    assert dens(p, 50.0) == 0.0


# functions_test.cc:433-469
def test_compute_transmittance_to_top_atmosphere_boundary():
    a = planet()
    r = R_B * 0.2 + R_T * 0.8
    h_r, h_top = r - R_B, R_T - R_B
    tau_r = BETA_R * H_R * (math.exp(-h_r / H_R) - math.exp(-h_top / H_R))
    tau_m = BETA_M_EXT * H_M * (math.exp(-h_r / H_M) - math.exp(-h_top / H_M))
    o = out3()
    L().or_transmittance_to_top(C.byref(a), r, 1.0, o)
    assert all(abs(v - math.exp(-(tau_r + tau_m))) < EPS for v in o)
    # Absorption only, tent profile whose integral is 15 km.
    a.rayleigh_density[1] = ol.layer()
    a.mie_density[1] = ol.layer()
    a.absorption_density[0] = ol.layer(25.0, 0.0, 0.0, 1.0 / 15.0, -2.0 / 3.0)
    a.absorption_density[1] = ol.layer(0.0, 0.0, 0.0, -1.0 / 15.0, 8.0 / 3.0)
    for c in range(3):
        a.absorption_extinction[c] = 0.02
    L().or_transmittance_to_top(C.byref(a), R_B, 1.0, o)
    assert all(abs(v - math.exp(-0.02 * 15.0)) < EPS for v in o)
    a = planet(uniform=True, no_aerosols=True)
    L().or_transmittance_to_top(C.byref(a), r, 0.0, o)
    assert all(abs(v - math.exp(-BETA_R * math.sqrt(R_T * R_T - r * r))) < EPS for v in o)


# functions_test.cc:479-488
def test_get_texture_coord_from_unit_range():
    f, g = L().or_tex_coord_from_unit_range, L().or_unit_range_from_tex_coord
    assert abs(f(0.0, 10) - 0.5 / 10) < EPS
    assert abs(f(1.0, 10) - 9.5 / 10) < EPS
    assert abs(g(0.5 / 10, 10)) < EPS
    assert abs(g(9.5 / 10, 10) - 1.0) < EPS
    assert abs(g(f(1.0 / 3.0, 10), 10) - 1.0 / 3.0) < EPS


# functions_test.cc:497-518
def test_get_transmittance_texture_uv_from_rmu():
    a = planet()
    W, H = S.transmittance_width, S.transmittance_height
    uv = (C.c_double * 2)()
    cases = [
        (R_T, 1.0, 0.5 / W, 1.0 - 0.5 / H),
        (R_T, horizon_mu(R_T), 1.0 - 0.5 / W, 1.0 - 0.5 / H),
        (R_B, 1.0, 0.5 / W, 0.5 / H),
        (R_B, 0.0, 1.0 - 0.5 / W, 0.5 / H),
    ]
    for r, mu, eu, ev in cases:
        L().or_transmittance_uv_from_rmu(C.byref(a), S, r, mu, uv)
        assert abs(uv[0] - eu) < EPS and abs(uv[1] - ev) < EPS


# functions_test.cc:529-574
def test_get_rmu_from_transmittance_texture_uv():
    a = planet()
    W, H = S.transmittance_width, S.transmittance_height
    r, mu = C.c_double(), C.c_double()
    f = lambda u, v: L().or_rmu_from_transmittance_uv(C.byref(a), S, u, v,
                                                      C.byref(r), C.byref(mu))
    f(0.5 / W, 1.0 - 0.5 / H)
    assert abs(r.value - R_T) < METER and abs(mu.value - 1.0) < EPS
    f(1.0 - 0.5 / W, 1.0 - 0.5 / H)
    assert abs(r.value - R_T) < METER and abs(mu.value - horizon_mu(R_T)) < EPS
    f(0.5 / W, 0.5 / H)
    assert abs(r.value - R_B) < METER and abs(mu.value - 1.0) < EPS
    f(1.0 - 0.5 / W, 0.5 / H)
    assert abs(r.value - R_B) < METER and abs(mu.value) < EPS
    uv = (C.c_double * 2)()
    r0 = R_B * 0.2 + R_T * 0.8
    L().or_transmittance_uv_from_rmu(C.byref(a), S, r0, 0.25, uv)
    f(uv[0], uv[1])
    assert abs(r.value - r0) < METER and abs(mu.value - 0.25) < EPS


# functions_test.cc:583-594
def test_get_transmittance_to_top_atmosphere_boundary():
    a = planet()
    T = t_lut(a)
    r, mu = R_B * 0.2 + R_T * 0.8, 0.4
    o1, o2 = out3(), out3()
    L().or_transmittance_to_top(C.byref(a), r, mu, o1)
    L().or_get_transmittance_to_top(C.byref(a), S, ol.ptr(T), r, mu, o2)
    assert all(abs(x - y) < EPS for x, y in zip(o1, o2))


# functions_test.cc:605-632
def test_compute_and_get_transmittance():
    a = planet(uniform=True, no_aerosols=True)
    T = t_lut(a)
    r = R_B * 0.2 + R_T * 0.8
    d = (R_T - R_B) * 0.1
    expected = math.exp(-BETA_R * d)
    o = out3()
    for rr, mu, ground in ((R_B, 0.0, 0), (r, 0.7, 0),
                           (r, -0.7, L().or_ray_intersects_ground(C.byref(a), r, -0.7))):
        L().or_get_transmittance(C.byref(a), S, ol.ptr(T), rr, mu, d, ground, o)
        assert abs(o[0] - expected) < EPS


# functions_test.cc:676-736
def test_compute_single_scattering_integrand():
    a = planet()
    T = t_lut(a)
    h_top = R_T - R_B
    h = h_top / 2.0
    ray, mie = out3(), out3()
    L().or_single_scattering_integrand(C.byref(a), S, ol.ptr(T), R_B, 1.0, 1.0, 1.0, h, 0,
                                       ray, mie)
    tau_r = BETA_R * H_R * (1.0 - math.exp(-h_top / H_R))
    tau_m = BETA_M_EXT * H_M * (1.0 - math.exp(-h_top / H_M))
    assert abs(ray[0] - math.exp(-tau_r - tau_m) * math.exp(-h / H_R)) < EPS
    assert abs(mie[0] - math.exp(-tau_r - tau_m) * math.exp(-h / H_M)) < EPS
    L().or_single_scattering_integrand(C.byref(a), S, ol.ptr(T), R_T, -1.0, 1.0, -1.0, h, 1,
                                       ray, mie)
    tau_r = 2.0 * BETA_R * H_R * (math.exp(-h / H_R) - math.exp(-h_top / H_R))
    tau_m = 2.0 * BETA_M_EXT * H_M * (math.exp(-h / H_M) - math.exp(-h_top / H_M))
    assert abs(ray[0] - math.exp(-tau_r - tau_m) * math.exp(-h / H_R)) < EPS
    assert abs(mie[0] - math.exp(-tau_r - tau_m) * math.exp(-h / H_M)) < EPS
    a = planet(uniform=True, no_aerosols=True)
    T = t_lut(a)
    L().or_single_scattering_integrand(C.byref(a), S, ol.ptr(T), R_B, 0.0, 0.0, 1.0, 50.0, 0,
                                       ray, mie)
    tau_r = BETA_R * math.sqrt(R_T * R_T - R_B * R_B)
    assert abs(ray[0] - math.exp(-tau_r)) < EPS


# functions_test.cc:745-768
def test_distance_to_nearest_atmosphere_boundary():
    a = planet()
    r = R_B * 0.2 + R_T * 0.8
    g = lambda mu: L().or_ray_intersects_ground(C.byref(a), r, mu)
    assert abs(L().or_distance_to_nearest(C.byref(a), r, 1.0, g(1.0)) - (R_T - r)) < METER
    assert abs(L().or_distance_to_nearest(C.byref(a), r, 0.0, g(0.0)) -
               math.sqrt(R_T * R_T - r * r)) < METER
    assert abs(L().or_distance_to_nearest(C.byref(a), r, -1.0, g(-1.0)) - (r - R_B)) < METER


# functions_test.cc:818-858
def test_compute_single_scattering():
    a = planet()
    T = t_lut(a)
    h_top = R_T - R_B
    ray, mie = out3(), out3()
    L().or_single_scattering(C.byref(a), S, ol.ptr(T), R_B, 1.0, 1.0, 1.0, 0, ray, mie)
    tau_r = BETA_R * H_R * (1.0 - math.exp(-h_top / H_R))
    tau_m = BETA_M_EXT * H_M * (1.0 - math.exp(-h_top / H_M))
    assert abs(ray[0] / (SOLAR * tau_r * math.exp(-tau_r - tau_m)) - 1.0) < 10 * EPS
    assert abs(mie[0] / (SOLAR * tau_m * BETA_M / BETA_M_EXT * math.exp(-tau_r - tau_m))
               - 1.0) < 10 * EPS
    a = planet(no_aerosols=True)
    T = t_lut(a)
    L().or_single_scattering(C.byref(a), S, ol.ptr(T), R_T, -1.0, 1.0, -1.0, 1, ray, mie)
    expected = SOLAR * 0.5 * (1.0 - math.exp(-2.0 * H_R * BETA_R *
                                             (1.0 - math.exp(-h_top / H_R))))
    assert abs(ray[0] / expected - 1.0) < 2 * EPS
    assert abs(mie[0]) < EPS


# functions_test.cc:865-879
def test_phase_functions():
    n = 100
    ri = mi = 0.0
    for i in range(n):
        theta = (i + 0.5) * math.pi / n
        dw = math.sin(theta) * (math.pi / n) * (2.0 * math.pi)
        ri += L().or_rayleigh_phase(math.cos(theta)) * dw
        mi += L().or_mie_phase(0.8, math.cos(theta)) * dw
    assert abs(ri - 1.0) < 2 * EPS
    assert abs(mi - 1.0) < 2 * EPS


def uvwz(a, r, mu, mu_s, nu, ground):
    o = (C.c_double * 4)()
    L().or_scattering_uvwz_from_rmumusnu(C.byref(a), S, r, mu, mu_s, nu, int(ground), o)
    return list(o)


# functions_test.cc:888-949
def test_get_scattering_texture_uvwz_from_rmumusnu():
    a = planet()
    Rn, MU, MUS = S.scattering_r, S.scattering_mu, S.scattering_mu_s
    assert abs(uvwz(a, R_B, 0.0, 0.0, 0.0, False)[3] - 0.5 / Rn) < EPS
    assert abs(uvwz(a, R_T, 0.0, 0.0, 0.0, False)[3] - (1.0 - 0.5 / Rn)) < EPS
    r = (R_T + R_B) / 2.0
    mu = horizon_mu(r)
    assert abs(uvwz(a, r, mu, 0.0, 0.0, True)[2] - 0.5 / MU) < EPS
    assert abs(uvwz(a, r, mu, 0.0, 0.0, False)[2] - (1.0 - 0.5 / MU)) < EPS
    assert uvwz(a, r, -1.0, 0.0, 0.0, True)[2] < 0.5
    assert uvwz(a, r, 1.0, 0.0, 0.0, False)[2] > 0.5
    for rr in (R_B, R_T):
        assert abs(uvwz(a, rr, 0.0, -1.0, 0.0, False)[1] - 0.5 / MUS) < EPS
        assert abs(uvwz(a, rr, 0.0, 1.0, 0.0, False)[1] - (1.0 - 0.5 / MUS)) < EPS
    assert abs(uvwz(a, R_B, 0.0, 0.0, -1.0, False)[0]) < EPS
    assert abs(uvwz(a, R_B, 0.0, 0.0, 1.0, False)[0] - 1.0) < EPS


def rmumusnu(a, u):
    o = (C.c_double * 4)()
    g = C.c_int()
    L().or_rmumusnu_from_scattering_uvwz(C.byref(a), S, (C.c_double * 4)(*u), o, C.byref(g))
    return list(o), bool(g.value)


# functions_test.cc:960-1060
def test_get_rmumusnu_from_scattering_texture_uvwz():
    a = planet()
    Rn, MU, MUS = S.scattering_r, S.scattering_mu, S.scattering_mu_s
    (r, mu, mu_s, nu), g = rmumusnu(a, (0.0, 0.5 / MUS, 0.5 / MU, 0.5 / Rn))
    assert abs(r - R_B) < METER
    (r, mu, mu_s, nu), g = rmumusnu(a, (0.0, 0.5 / MUS, 0.5 / MU, 1.0 - 0.5 / Rn))
    assert abs(r - R_T) < METER
    (r, mu, mu_s, nu), g = rmumusnu(a, (0.0, 0.5 / MUS, 0.5 / MU + EPS, 0.5))
    mh = horizon_mu(r)
    assert abs(mu - mh) < EPS and mu <= mh and g
    (r, mu, mu_s, nu), g = rmumusnu(a, (0.0, 0.5 / MUS, 1.0 - 0.5 / MU - EPS, 0.5))
    assert abs(mu - mh) < 5 * EPS and mu >= mh and not g
    (r, mu, mu_s, nu), g = rmumusnu(a, (0.0, 0.5 / MUS, 0.5 / MU, 0.5 / Rn))
    assert abs(mu_s + 1.0) < EPS
    (r, mu, mu_s, nu), g = rmumusnu(a, (0.0, 1.0 - 0.5 / MUS, 0.5 / MU, 0.5 / Rn))
    assert abs(mu_s - 1.0) < EPS
    (r, mu, mu_s, nu), g = rmumusnu(a, (0.0, 0.5 / MUS, 0.5 / MU, 0.5 / Rn))
    assert abs(nu + 1.0) < EPS
    (r, mu, mu_s, nu), g = rmumusnu(a, (1.0, 0.5 / MUS, 0.5 / MU, 0.5 / Rn))
    assert abs(nu - 1.0) < EPS
    for rr in (R_B, R_T):
        (r, mu, mu_s, nu), g = rmumusnu(a, uvwz(a, rr, -1.0, 1.0, -1.0, True))
        assert abs(r - rr) < METER and abs(mu + 1) < EPS and abs(mu_s - 1) < EPS
        assert abs(nu + 1) < EPS and g
    (r, mu, mu_s, nu), g = rmumusnu(a, uvwz(a, (R_B + R_T) / 2, 0.2, 0.3, 0.4, False))
    assert abs(r - (R_B + R_T) / 2) < METER
    assert abs(mu - 0.2) < EPS and abs(mu_s - 0.3) < EPS and abs(nu - 0.4) < EPS
    assert not g


# functions_test.cc:1069-1131
def test_compute_and_get_single_scattering():
    a = planet()
    T = t_lut(a)

    def lazy(r, mu, mu_s, nu, ground, rayleigh):
        def fill(idx):
            ro = np.zeros((len(idx), 4))
            mo = np.zeros((len(idx), 4))
            L().or_texels_single_scattering(C.byref(a), S, ol.ptr(T), len(idx),
                                            idx.ctypes.data_as(C.POINTER(C.c_int)),
                                            ol.ptr(ro), ol.ptr(mo))
            return ro if rayleigh else mo
        return lazy_scattering(a, S, fill, r, mu, mu_s, nu, ground)

    mh = horizon_mu(R_T)
    cases = [((R_B, 1.0, 1.0, 1.0, False), EPS), ((R_T, -1.0, 1.0, -1.0, True), EPS),
             ((R_B, 0.0, 0.0, 0.0, False), 1e-1), ((R_T, mh, 1.0, mh, False), EPS)]
    ray, mie = out3(), out3()
    for args, tol in cases:
        L().or_single_scattering(C.byref(a), S, ol.ptr(T), *args[:4], int(args[4]), ray, mie)
        got_r = get_scattering(a, S, lazy(*args, True), *args)
        got_m = get_scattering(a, S, lazy(*args, False), *args)
        assert abs(got_r[0] / ray[0] - 1.0) < tol
        assert abs(got_m[0] / mie[0] - 1.0) < tol


def uniform_lut(shape, value):
    lut = np.zeros(shape)
    lut[..., :3] = value
    return lut


# functions_test.cc:1155-1189
def test_compute_scattering_density():
    a = planet()
    full_t = uniform_lut(ol.transmittance_shape(S), 1.0)
    no_single = uniform_lut(ol.scattering_shape(S), 0.0)
    uniform_ms = uniform_lut(ol.scattering_shape(S), 13.0)
    no_e = uniform_lut(ol.irradiance_shape(S), 0.0)
    o = out3()
    L().or_scattering_density(C.byref(a), S, ol.ptr(full_t), ol.ptr(no_single),
                              ol.ptr(no_single), ol.ptr(uniform_ms), ol.ptr(no_e),
                              R_B, 0.0, 0.0, 1.0, 3, o)
    assert abs(o[0] / ((BETA_R + BETA_M) * 13.0) - 1.0) < 2 * EPS
    uniform_e = uniform_lut(ol.irradiance_shape(S), 13.0)
    L().or_scattering_density(C.byref(a), S, ol.ptr(full_t), ol.ptr(no_single),
                              ol.ptr(no_single), ol.ptr(no_single), ol.ptr(uniform_e),
                              R_B, 0.0, 0.0, 1.0, 3, o)
    expected = (BETA_R + BETA_M) * ALBEDO / (2.0 * math.pi) * 13.0
    assert abs(o[0] / expected - 1.0) < 2 * EPS


# functions_test.cc:1202-1226
def test_compute_multiple_scattering():
    a = planet()
    full_t = uniform_lut(ol.transmittance_shape(S), 1.0)
    j = uniform_lut(ol.scattering_shape(S), 0.17)
    o = out3()
    r = R_B * 0.2 + R_T * 0.8
    L().or_multiple_scattering(C.byref(a), S, ol.ptr(full_t), ol.ptr(j), r, -1.0, 1.0, -1.0,
                               1, o)
    assert abs(o[0] - 0.17 * (r - R_B)) < 0.17 * (r - R_B) * EPS
    mu = horizon_mu(R_T)
    dh = math.sqrt(R_T * R_T - R_B * R_B)
    L().or_multiple_scattering(C.byref(a), S, ol.ptr(full_t), ol.ptr(j), R_T, mu, 1.0, mu,
                               1, o)
    assert abs(o[0] - 0.17 * dh) < 0.17 * dh * EPS


# functions_test.cc:1237-1276
def test_compute_and_get_scattering_density():
    a = planet()
    full_t = uniform_lut(ol.transmittance_shape(S), 1.0)
    no_single = uniform_lut(ol.scattering_shape(S), 0.0)
    for ms_value, e_value, expected in (
            (13.0, 0.0, (BETA_R + BETA_M) * 13.0),
            (0.0, 13.0, (BETA_R + BETA_M) * ALBEDO / (2.0 * math.pi) * 13.0)):
        ms = uniform_lut(ol.scattering_shape(S), ms_value)
        e = uniform_lut(ol.irradiance_shape(S), e_value)

        def fill(idx):
            o = np.zeros((len(idx), 4))
            L().or_texels_scattering_density(
                C.byref(a), S, ol.ptr(full_t), ol.ptr(no_single), ol.ptr(no_single),
                ol.ptr(ms), ol.ptr(e), 3, len(idx),
                idx.ctypes.data_as(C.POINTER(C.c_int)), ol.ptr(o))
            return o
        lut = lazy_scattering(a, S, fill, R_B, 0.0, 0.0, 1.0, False)
        got = get_scattering(a, S, lut, R_B, 0.0, 0.0, 1.0, False)
        assert abs(got[0] / expected - 1.0) < 2 * EPS


# functions_test.cc:1287-1313
def test_compute_and_get_multiple_scattering():
    a = planet()
    full_t = uniform_lut(ol.transmittance_shape(S), 1.0)
    j = uniform_lut(ol.scattering_shape(S), 0.17)

    def fill(idx):
        o = np.zeros((len(idx), 4))
        L().or_texels_multiple_scattering(C.byref(a), S, ol.ptr(full_t), ol.ptr(j), len(idx),
                                          idx.ctypes.data_as(C.POINTER(C.c_int)), ol.ptr(o))
        return o
    r = R_B * 0.2 + R_T * 0.8
    mu = horizon_mu(R_T)
    dh = math.sqrt(R_T * R_T - R_B * R_B)
    for args, expected in (((r, -1.0, 1.0, -1.0, True), 0.17 * (r - R_B)),
                           ((R_T, mu, 1.0, mu, True), 0.17 * dh)):
        got = get_scattering(a, S, lazy_scattering(a, S, fill, *args), *args)
        assert abs(got[0] - expected) < expected * EPS


# functions_test.cc:1325-1337
def test_compute_indirect_irradiance():
    a = planet()
    no_single = uniform_lut(ol.scattering_shape(S), 0.0)
    ms = uniform_lut(ol.scattering_shape(S), 1.0)
    o = out3()
    L().or_indirect_irradiance(C.byref(a), S, ol.ptr(no_single), ol.ptr(no_single),
                               ol.ptr(ms), R_B, 1.0, 2, o)
    assert abs(o[0] - math.pi) < 10 * EPS


# functions_test.cc:1346-1367
def test_get_irradiance_texture_uv_from_rmus():
    a = planet()
    W, H = S.irradiance_width, S.irradiance_height
    uv = (C.c_double * 2)()
    f = lambda r, m: (L().or_irradiance_uv_from_rmus(C.byref(a), S, r, m, uv), list(uv))[1]
    assert abs(f(R_B, 0.0)[1] - 0.5 / H) < EPS
    assert abs(f(R_T, 0.0)[1] - (1.0 - 0.5 / H)) < EPS
    assert abs(f(R_B, -1.0)[0] - 0.5 / W) < EPS
    assert abs(f(R_B, 1.0)[0] - (1.0 - 0.5 / W)) < EPS


# functions_test.cc:1376-1399
def test_get_rmus_from_irradiance_texture_uv():
    a = planet()
    W, H = S.irradiance_width, S.irradiance_height
    r, m = C.c_double(), C.c_double()
    f = lambda u, v: (L().or_rmus_from_irradiance_uv(C.byref(a), S, u, v, C.byref(r),
                                                     C.byref(m)), (r.value, m.value))[1]
    assert abs(f(0.5 / W, 0.5 / H)[0] - R_B) < METER
    assert abs(f(0.5 / W, 1.0 - 0.5 / H)[0] - R_T) < METER
    assert abs(f(0.5 / W, 0.5 / H)[1] + 1.0) < EPS
    assert abs(f(1.0 - 0.5 / W, 0.5 / H)[1] - 1.0) < EPS


# functions_test.cc:1408-1443
def test_compute_and_get_irradiance():
    a = planet()
    no_single = uniform_lut(ol.scattering_shape(S), 0.0)
    w = S.scattering_nu * S.scattering_mu_s
    h, d = S.scattering_mu, S.scattering_r
    x, y, z = np.meshgrid(np.arange(w), np.arange(h), np.arange(d), indexing="ij")
    v = z + d * (y + h * x)                      # indexed [x, y, z]
    fake = np.zeros(ol.scattering_shape(S))
    fake[..., :3] = np.transpose(v, (2, 1, 0))[..., None]
    E = np.zeros(ol.irradiance_shape(S))
    L().or_pass_indirect_irradiance(C.byref(a), S, ol.ptr(no_single), ol.ptr(no_single),
                                    ol.ptr(fake), 2, ol.ptr(E))
    r, mu_s = R_B * 0.8 + R_T * 0.2, 0.25
    looked, direct, other = out3(), out3(), out3()
    L().or_get_irradiance(C.byref(a), S, ol.ptr(E), r, mu_s, looked)
    L().or_indirect_irradiance(C.byref(a), S, ol.ptr(no_single), ol.ptr(no_single),
                               ol.ptr(fake), r, mu_s, 2, direct)
    L().or_indirect_irradiance(C.byref(a), S, ol.ptr(no_single), ol.ptr(no_single),
                               ol.ptr(fake), r, 0.5, 2, other)
    assert abs(looked[0] / direct[0] - 1.0) < EPS
    assert abs(looked[0] / other[0] - 1.0) >= EPS
