"""The test scene of the reference's image tests (atmosphere/reference/
model_test.glsl:59-348, camera and sun set-up of model_test.cc:306-319 and
:436-476): a 1 km sphere resting on a spherical planet, seen from 500 m
above the ground, with soft shadows and the light shafts of the sphere's
umbra cone, rendered by ray casting.

The geometry (view rays, sphere and ground hits, shadow-volume entry and
exit) is computed here on the host in float64; the atmosphere enters only
through the three lookups, which the caller supplies (GPU kernels or the
FP64 oracle), so two renders differ only by their lookups. Lengths are in km
(the model's length unit)."""
import math

import numpy as np

WIDTH, HEIGHT = 640, 360                      # model_test.cc:180-181
SPHERE_RADIUS = 1.0                           # model_test.cc:88
SPHERE_CENTER = np.array([0.0, 0.0, 1.0])     # model_test.cc:89
CAMERA = np.array([2.0, -8.0, 0.5])           # model_test.cc:440
PITCH = math.pi / 30.0                        # model_test.cc:441
FOV_Y = 50.0 / 180.0 * math.pi                # model_test.cc:451


def model_from_clip():
    """Clip space to world space, 3x3 (model_test.cc:438-467)."""
    model_from_view = np.array([
        [1.0, 0.0, 0.0, CAMERA[0]],
        [0.0, -math.sin(PITCH), -math.cos(PITCH), CAMERA[1]],
        [0.0, math.cos(PITCH), -math.sin(PITCH), CAMERA[2]],
        [0.0, 0.0, 0.0, 1.0]])
    t = math.tan(FOV_Y / 2.0)
    view_from_clip = np.array([
        [t * WIDTH / HEIGHT, 0.0, 0.0, 0.0],
        [0.0, t, 0.0, 0.0],
        [0.0, 0.0, 0.0, -1.0],
        [0.0, 0.0, 1.0, 1.0]])
    m = np.zeros((3, 3))
    for row in range(3):
        for col in range(3):
            col2 = col if col < 2 else 3
            m[row, col] = sum(model_from_view[row, k] * view_from_clip[k, col2]
                              for k in range(3))
    return m


def sun_direction(theta_deg, phi_deg):
    th, ph = math.radians(theta_deg), math.radians(phi_deg)
    return np.array([math.cos(ph) * math.sin(th), math.sin(ph) * math.sin(th), math.cos(th)])


def _dot(a, b):
    return np.einsum("...k,...k->...", a, b)


class Scene:
    """Per-pixel geometry for one sun direction (model_test.glsl
    GetViewRayRadiance, :214-348, up to the lookups)."""

    def __init__(self, bottom_radius, sun_angular_radius, sun_dir):
        self.earth_center = np.array([0.0, 0.0, -bottom_radius])
        self.sun = np.asarray(sun_dir, dtype=np.float64)
        self.sun_size = (math.tan(sun_angular_radius), math.cos(sun_angular_radius))
        m = model_from_clip()
        j, i = np.mgrid[0:HEIGHT, 0:WIDTH]
        y = 1.0 - 2.0 * (j + 0.5) / HEIGHT
        x = 2.0 * (i + 0.5) / WIDTH - 1.0
        dx, dy = 2.0 / WIDTH, -2.0 / HEIGHT
        ray = np.stack([m[r, 0] * x + m[r, 1] * y + m[r, 2] for r in range(3)], -1)
        diff = np.array([m[r, 0] * dx + m[r, 1] * dy for r in range(3)])
        self.view = ray / np.linalg.norm(ray, axis=-1, keepdims=True)
        self.fragment_angular_size = np.linalg.norm(diff) / np.linalg.norm(ray, axis=-1)
        self.shadow_in, self.shadow_out = self._shadow_in_out()
        self._sphere()
        self._ground()
        self.sky_shadow = np.maximum(0.0, self.shadow_out - self.shadow_in)

    # model_test.glsl:166-198
    def _shadow_in_out(self):
        pos = CAMERA - SPHERE_CENTER
        v, s = self.view, self.sun
        pos_dot_sun = pos @ s
        view_dot_sun = v @ s
        k = self.sun_size[0]
        l = 1.0 + k * k
        a = 1.0 - l * view_dot_sun * view_dot_sun
        b = v @ pos - l * pos_dot_sun * view_dot_sun - k * SPHERE_RADIUS * view_dot_sun
        c = (pos @ pos - l * pos_dot_sun * pos_dot_sun -
             2.0 * k * SPHERE_RADIUS * pos_dot_sun - SPHERE_RADIUS * SPHERE_RADIUS)
        disc = b * b - a * c
        with np.errstate(invalid="ignore", divide="ignore"):
            sq = np.sqrt(np.maximum(disc, 0.0))
            d_in = np.maximum(0.0, (-b - sq) / a)
            d_out = (-b + sq) / a
            d_base = -pos_dot_sun / view_dot_sun
            d_apex = -(pos_dot_sun + SPHERE_RADIUS / k) / view_dot_sun
        up = view_dot_sun > 0.0
        d_in2 = np.where(up, np.maximum(d_in, d_apex), np.where(a > 0.0, np.maximum(d_in, d_base),
                                                                 d_base))
        d_out2 = np.where(up, np.where(a > 0.0, np.minimum(d_out, d_base), d_base),
                          np.minimum(d_out, d_apex))
        hit = disc > 0.0
        return np.where(hit, d_in2, 0.0), np.where(hit, d_out2, 0.0)

    # model_test.glsl:226-271
    def _sphere(self):
        p = CAMERA - SPHERE_CENTER
        p_dot_v = self.view @ p
        p_dot_p = p @ p
        center_sq = p_dot_p - p_dot_v * p_dot_v
        with np.errstate(invalid="ignore"):
            dist = -p_dot_v - np.sqrt(SPHERE_RADIUS * SPHERE_RADIUS - center_sq)
        self.sphere_hit = dist > 0.0
        ray_sphere_distance = SPHERE_RADIUS - np.sqrt(np.maximum(center_sq, 0.0))
        angular = -ray_sphere_distance / p_dot_v
        self.sphere_alpha = np.where(self.sphere_hit,
                                     np.minimum(angular / self.fragment_angular_size, 1.0),
                                     0.0)
        d = np.where(self.sphere_hit, dist, 0.0)
        self.sphere_point = CAMERA + self.view * d[..., None]
        nrm = self.sphere_point - SPHERE_CENTER
        self.sphere_normal = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
        self.sphere_shadow = np.maximum(0.0, np.minimum(self.shadow_out, d) - self.shadow_in)

    # model_test.glsl:280-310 and the visibility factors of :59-98
    def _ground(self):
        p = CAMERA - self.earth_center
        p_dot_v = self.view @ p
        p_dot_p = p @ p
        center_sq = p_dot_p - p_dot_v * p_dot_v
        with np.errstate(invalid="ignore"):
            dist = -p_dot_v - np.sqrt(self.earth_center[2] ** 2 - center_sq)
        self.ground_hit = dist > 0.0
        d = np.where(self.ground_hit, dist, 0.0)
        self.ground_point = CAMERA + self.view * d[..., None]
        nrm = self.ground_point - self.earth_center
        self.ground_normal = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
        self.ground_shadow = np.maximum(0.0, np.minimum(self.shadow_out, d) - self.shadow_in)
        # GetSunVisibility (model_test.glsl:59-75)
        q = self.ground_point - SPHERE_CENTER
        q_dot_v = q @ self.sun
        q_dot_q = _dot(q, q)
        csq = q_dot_q - q_dot_v * q_dot_v
        with np.errstate(invalid="ignore"):
            dint = -q_dot_v - np.sqrt(SPHERE_RADIUS * SPHERE_RADIUS - csq)
        rsd = SPHERE_RADIUS - np.sqrt(np.maximum(csq, 0.0))
        with np.errstate(invalid="ignore", divide="ignore"):
            ang = -rsd / q_dot_v
        t = np.clip(ang / self.sun_size[0], 0.0, 1.0)
        smooth = 1.0 - t * t * (3.0 - 2.0 * t)  # smoothstep(1, 0, x)
        self.sun_visibility = np.where(dint > 0.0, smooth, 1.0)
        # GetSkyVisibility (model_test.glsl:84-89)
        self.sky_visibility = 1.0 + q[..., 2] / np.sqrt(q_dot_q) * SPHERE_RADIUS ** 2 / q_dot_q

    def queries(self):
        """Lookup inputs, float32, relative to the planet centre: 'sky' for
        every pixel; 'sphere' and 'ground' for the pixels whose view ray hits
        the sphere or the ground (the shader calls those lookups only
        there)."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        hw = HEIGHT * WIDTH
        cam = CAMERA - self.earth_center
        out = dict(sky=dict(camera=f(np.broadcast_to(cam, (hw, 3))),
                            view=f(self.view.reshape(hw, 3)),
                            shadow=f(self.sky_shadow.reshape(hw)),
                            sun=f(np.broadcast_to(self.sun, (hw, 3)))))
        for name, hit, point, normal, shadow in (
                ("sphere", self.sphere_hit, self.sphere_point, self.sphere_normal,
                 self.sphere_shadow),
                ("ground", self.ground_hit, self.ground_point, self.ground_normal,
                 self.ground_shadow)):
            idx = np.flatnonzero(hit.reshape(-1))
            k = len(idx)
            out[name] = dict(index=idx,
                             camera=f(np.broadcast_to(cam, (k, 3))),
                             point=f(point.reshape(hw, 3)[idx] - self.earth_center),
                             normal=f(normal.reshape(hw, 3)[idx]),
                             shadow=f(shadow.reshape(hw)[idx]),
                             sun=f(np.broadcast_to(self.sun, (k, 3))))
        return out

    def composite(self, q, lookups, solar_radiance, ground_albedo, sphere_albedo):
        """Final radiance per pixel (model_test.glsl:251-347) from the lookup
        results: lookups = dict(sky=(radiance, transmittance) per pixel,
        sphere=(sun irradiance, sky irradiance, in-scatter, transmittance)
        per sphere pixel, ground=(same) per ground pixel)."""
        hw = HEIGHT * WIDTH
        f = lambda a: np.asarray(a, np.float64)
        sky, sky_t = map(f, lookups["sky"])
        radiance = sky.copy()
        sun_disc = (self.view @ self.sun > self.sun_size[1]).reshape(hw)
        radiance[sun_disc] += sky_t[sun_disc] * np.asarray(solar_radiance)
        g_sun, g_sky, g_in, g_t = map(f, lookups["ground"])
        gi = q["ground"]["index"]
        vis_sun = self.sun_visibility.reshape(hw)[gi][:, None]
        vis_sky = self.sky_visibility.reshape(hw)[gi][:, None]
        ground = np.asarray(ground_albedo) / math.pi * (g_sun * vis_sun + g_sky * vis_sky)
        radiance[gi] = ground * g_t + g_in
        s_sun, s_sky, s_in, s_t = map(f, lookups["sphere"])
        si = q["sphere"]["index"]
        sphere = np.asarray(sphere_albedo) / math.pi * (s_sun + s_sky)
        sphere = sphere * s_t + s_in
        sa = self.sphere_alpha.reshape(hw)[si][:, None]
        radiance[si] = radiance[si] * (1.0 - sa) + sphere * sa
        return radiance.reshape(HEIGHT, WIDTH, 3)


def to_8bit(radiance, exposure):
    """model_test.cc:724-733 (and the GL shader, :155)."""
    c = np.power(1.0 - np.exp(-radiance * exposure), 1.0 / 2.2)
    return np.floor(c * 255.0).astype(np.int64)


def psnr(img1, img2):
    """ComputePSNR of model_test.cc:750-765 (its definition, with the square
    root of the mean squared error)."""
    sq = np.sum((img1 - img2) ** 2)
    mse = math.sqrt(sq / (WIDTH * HEIGHT))
    if mse == 0.0:
        return math.inf
    return 10.0 * math.log10(255.0 * 255.0 / mse)
