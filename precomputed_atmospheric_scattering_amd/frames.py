"""Frames of the reference's scene, as coherent lookup batches: a 1 km sphere
resting on a spherical planet, ray cast with soft shadows and the light
shafts of the sphere's umbra cone (atmosphere/demo/demo.glsl:59-387,
atmosphere/reference/model_test.glsl:59-348). Two cameras:

- the image tests' camera (model_test.cc:436-476): 500 m above the ground,
  pitched up by pi/30;
- the demo's orbit camera (demo.cc:357-380) at a distance, view zenith and
  azimuth, with the nine preset views of its number keys (demo.cc:471-489).

The geometry (view rays, sphere and ground hits, shadow-volume entry and
exit, visibility factors) is computed on the host in float64. The
atmosphere enters only through the three lookups, which the caller runs on
the batches of Frame.queries() (the GPU kernels, or any other
implementation) and hands back to Frame.composite(). Lengths are in km,
the model's length unit of these scenes.
"""
import math

import numpy as np

SPHERE_RADIUS = 1.0                        # demo.glsl:62, model_test.cc:88
SPHERE_CENTER = np.array([0.0, 0.0, 1.0])  # demo.glsl:61, model_test.cc:89
FOV_Y = 50.0 / 180.0 * math.pi             # demo.cc:425, model_test.cc:451

# demo.cc:471-489: view distance (m), view zenith, view azimuth, sun zenith,
# sun azimuth (rad), exposure.
DEMO_VIEWS = {
    1: (9000.0, 1.47, 0.0, 1.3, 3.0, 10.0),
    2: (9000.0, 1.47, 0.0, 1.564, -3.0, 10.0),
    3: (7000.0, 1.57, 0.0, 1.54, -2.96, 10.0),
    4: (7000.0, 1.57, 0.0, 1.328, -3.044, 10.0),
    5: (9000.0, 1.39, 0.0, 1.2, 0.7, 10.0),
    6: (9000.0, 1.5, 0.0, 1.628, 1.05, 200.0),
    7: (7000.0, 1.43, 0.0, 1.57, 1.34, 40.0),
    8: (2.7e6, 0.81, 0.0, 1.57, 2.0, 10.0),
    9: (1.2e7, 0.0, 0.0, 0.93, -2.0, 10.0),
}


def _view_from_clip(width, height):
    t = math.tan(FOV_Y / 2.0)
    return np.array([[t * width / height, 0.0, 0.0, 0.0],
                     [0.0, t, 0.0, 0.0],
                     [0.0, 0.0, 0.0, -1.0],
                     [0.0, 0.0, 1.0, 1.0]])


def _model_from_clip(model_from_view, width, height):
    """The 3x3 map from (clip x, clip y, 1) to the world-space view ray: the
    vertex shaders' model_from_view * (view_from_clip * vertex).xyz."""
    vfc = _view_from_clip(width, height)
    m = np.zeros((3, 3))
    for row in range(3):
        for col in range(3):
            col2 = col if col < 2 else 3
            m[row, col] = sum(model_from_view[row, k] * vfc[k, col2] for k in range(3))
    return m


class Camera:
    """A pinhole camera: position (km, world space) and the clip-to-world
    ray map of a width x height frame."""

    def __init__(self, position, model_from_view, width, height):
        self.position = np.asarray(position, dtype=np.float64)
        self.width, self.height = int(width), int(height)
        self.model_from_clip = _model_from_clip(model_from_view, width, height)

    @classmethod
    def image_test(cls, width=640, height=360):
        """model_test.cc:436-476: camera at (2, -8, 0.5) km, pitched up by
        pi / 30, looking along +y."""
        cam = np.array([2.0, -8.0, 0.5])
        pitch = math.pi / 30.0
        mfv = np.array([[1.0, 0.0, 0.0, cam[0]],
                        [0.0, -math.sin(pitch), -math.cos(pitch), cam[1]],
                        [0.0, math.cos(pitch), -math.sin(pitch), cam[2]],
                        [0.0, 0.0, 0.0, 1.0]])
        return cls(cam, mfv, width, height)

    @classmethod
    def demo(cls, view_distance_m, view_zenith, view_azimuth, width=1920, height=1080,
             length_unit_m=1000.0):
        """demo.cc:357-380: camera frame (ux, uy, uz) from the view zenith and
        azimuth angles, camera at distance * uz from the origin."""
        cz, sz = math.cos(view_zenith), math.sin(view_zenith)
        ca, sa = math.cos(view_azimuth), math.sin(view_azimuth)
        ux = (-sa, ca, 0.0)
        uy = (-cz * ca, -cz * sa, sz)
        uz = (sz * ca, sz * sa, cz)
        l = view_distance_m / length_unit_m
        mfv = np.array([[ux[0], uy[0], uz[0], uz[0] * l],
                        [ux[1], uy[1], uz[1], uz[1] * l],
                        [ux[2], uy[2], uz[2], uz[2] * l],
                        [0.0, 0.0, 0.0, 1.0]])
        return cls(mfv[:3, 3], mfv, width, height)


def sun_direction(zenith, azimuth):
    """Unit vector from zenith and azimuth angles in radians (demo.cc:386-388)."""
    return np.array([math.cos(azimuth) * math.sin(zenith),
                     math.sin(azimuth) * math.sin(zenith), math.cos(zenith)])


def _dot(a, b):
    return np.einsum("...k,...k->...", a, b)


def _smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


class Frame:
    """Per-pixel geometry of one frame (GetViewRayRadiance of
    model_test.glsl:214-348, main() of demo.glsl:250-387, up to the
    lookups). light_shaft_fadein: the demo's fade-out of the light shafts
    when the sun nears the horizon (demo.glsl:261-263)."""

    def __init__(self, bottom_radius, sun_angular_radius, sun_dir, camera,
                 light_shaft_fadein=False):
        self.camera = camera
        self.width, self.height = camera.width, camera.height
        self.earth_center = np.array([0.0, 0.0, -bottom_radius])
        self.sun = np.asarray(sun_dir, dtype=np.float64)
        self.sun_size = (math.tan(sun_angular_radius), math.cos(sun_angular_radius))
        m = camera.model_from_clip
        j, i = np.mgrid[0:self.height, 0:self.width]
        y = 1.0 - 2.0 * (j + 0.5) / self.height
        x = 2.0 * (i + 0.5) / self.width - 1.0
        dx, dy = 2.0 / self.width, -2.0 / self.height
        ray = np.stack([m[r, 0] * x + m[r, 1] * y + m[r, 2] for r in range(3)], -1)
        diff = np.array([m[r, 0] * dx + m[r, 1] * dy for r in range(3)])
        self.view = ray / np.linalg.norm(ray, axis=-1, keepdims=True)
        self.fragment_angular_size = np.linalg.norm(diff) / np.linalg.norm(ray, axis=-1)
        self.fadein = 1.0
        if light_shaft_fadein:
            up = camera.position - self.earth_center
            self.fadein = float(_smoothstep(0.02, 0.04, up @ self.sun / np.linalg.norm(up)))
        self.shadow_in, self.shadow_out = self._shadow_in_out()
        self._sphere()
        self._ground()
        self.sky_shadow = np.maximum(0.0, self.shadow_out - self.shadow_in) * self.fadein

    # model_test.glsl:166-198, demo.glsl:201-248
    def _shadow_in_out(self):
        pos = self.camera.position - SPHERE_CENTER
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

    # model_test.glsl:226-271, demo.glsl:270-318
    def _sphere(self):
        cam = self.camera.position
        p = cam - SPHERE_CENTER
        p_dot_v = self.view @ p
        p_dot_p = p @ p
        center_sq = p_dot_p - p_dot_v * p_dot_v
        with np.errstate(invalid="ignore"):
            dist = -p_dot_v - np.sqrt(SPHERE_RADIUS * SPHERE_RADIUS - center_sq)
        self.sphere_hit = dist > 0.0
        ray_sphere_distance = SPHERE_RADIUS - np.sqrt(np.maximum(center_sq, 0.0))
        with np.errstate(invalid="ignore", divide="ignore"):
            angular = -ray_sphere_distance / p_dot_v
        self.sphere_alpha = np.where(self.sphere_hit,
                                     np.minimum(angular / self.fragment_angular_size, 1.0),
                                     0.0)
        d = np.where(self.sphere_hit, dist, 0.0)
        self.sphere_point = cam + self.view * d[..., None]
        nrm = self.sphere_point - SPHERE_CENTER
        self.sphere_normal = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
        self.sphere_shadow = np.maximum(0.0, np.minimum(self.shadow_out, d) -
                                        self.shadow_in) * self.fadein

    # model_test.glsl:280-310, demo.glsl:320-362, and the visibility
    # factors of model_test.glsl:59-98 (demo.glsl:92-125)
    def _ground(self):
        cam = self.camera.position
        p = cam - self.earth_center
        p_dot_v = self.view @ p
        p_dot_p = p @ p
        center_sq = p_dot_p - p_dot_v * p_dot_v
        with np.errstate(invalid="ignore"):
            dist = -p_dot_v - np.sqrt(self.earth_center[2] ** 2 - center_sq)
        self.ground_hit = dist > 0.0
        d = np.where(self.ground_hit, dist, 0.0)
        self.ground_point = cam + self.view * d[..., None]
        nrm = self.ground_point - self.earth_center
        self.ground_normal = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
        self.ground_shadow = np.maximum(0.0, np.minimum(self.shadow_out, d) -
                                        self.shadow_in) * self.fadein
        # GetSunVisibility
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
        # GetSkyVisibility
        self.sky_visibility = 1.0 + q[..., 2] / np.sqrt(q_dot_q) * SPHERE_RADIUS ** 2 / q_dot_q

    def queries(self):
        """Lookup inputs, float32, relative to the planet centre: 'sky' for
        every pixel (camera, view, shadow, sun); 'sphere' and 'ground' for
        the pixels whose view ray hits the sphere or the ground (camera,
        point, normal, shadow, sun, and the pixel 'index'): the shaders call
        those lookups only there."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        hw = self.height * self.width
        cam = self.camera.position - self.earth_center
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
        """Final radiance per pixel (model_test.glsl:251-347, demo.glsl:
        364-385) from the lookup results: lookups = dict(sky=(radiance,
        transmittance) per pixel, sphere=(sun irradiance, sky irradiance,
        in-scatter, transmittance) per sphere pixel, ground=(same) per ground
        pixel)."""
        hw = self.height * self.width
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
        return radiance.reshape(self.height, self.width, 3)


def demo_frame(view, bottom_radius, sun_angular_radius, width=1920, height=1080):
    """(Frame, exposure) of demo preset view 1..9 (demo.cc:471-489), with the
    demo's light-shaft fade-in. bottom_radius in km."""
    dist, vz, va, sz, sa, exposure = DEMO_VIEWS[view]
    cam = Camera.demo(dist, vz, va, width, height)
    return Frame(bottom_radius, sun_angular_radius, sun_direction(sz, sa), cam,
                 light_shaft_fadein=True), exposure


def to_8bit(radiance, exposure, white_point=(1.0, 1.0, 1.0)):
    """Tone mapping of the shaders (model_test.cc:724-733, demo.glsl:384-385)."""
    c = np.power(1.0 - np.exp(-radiance / np.asarray(white_point) * exposure), 1.0 / 2.2)
    return np.floor(c * 255.0).astype(np.int64)


def psnr(img1, img2):
    """ComputePSNR of model_test.cc:750-765 (its definition, with the square
    root of the mean squared error)."""
    sq = np.sum((img1 - img2) ** 2)
    mse = math.sqrt(sq / (img1.shape[0] * img1.shape[1]))
    if mse == 0.0:
        return math.inf
    return 10.0 * math.log10(255.0 * 255.0 / mse)
