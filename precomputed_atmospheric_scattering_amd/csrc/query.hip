// Sky-shading lookups for gfx950: GetSkyRadiance, GetSkyRadianceToPoint and
// GetSunAndSkyIrradiance (atmosphere/functions.glsl:1633-1896) over batches
// of rays, one ray per thread.
//
// Precision plan. FP64 where FP32 cancellation moves a result: the camera
// radius and r * mu from the FP32 inputs, the two ray/sphere discriminants
// r^2 (mu^2 - 1) + rb^2 and + rt^2 and their square roots (the ground/sky
// choice and the mu coordinate near the horizon), the distances derived from
// them, sqrt(r^2 - rb^2), and the blending of texels in
// GetSkyRadianceToPoint, whose difference of two nearby lookups keeps the
// texels' precision only if each lookup is summed in FP64. The discriminants
// depend only on the ray's closest approach, so one square root of each
// serves every point of a ray. FP32 for the rest: LUT coordinates once their
// numerators are formed in FP64 (coordinates are below 4096 texels, so FP32
// keeps them to ~1e-4 texel), the mu_s mapping in cancellation-free form,
// phase functions and the GetSkyRadiance / irradiance blends.
#include <hip/hip_runtime.h>

#include "atmo_device.h"
#include "kernels.h"

namespace atmo {
namespace {

__device__ __forceinline__ float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_f(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt_f(float x) { return __builtin_amdgcn_sqrtf(x); }

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 load3(const float* p, int64_t i) {
  return D3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
template <class V>
__device__ __forceinline__ void store3(float* p, int64_t i, V v) {
  p[3 * i] = static_cast<float>(v.x);
  p[3 * i + 1] = static_cast<float>(v.y);
  p[3 * i + 2] = static_cast<float>(v.z);
}
__device__ __forceinline__ double dot3(D3 a, D3 b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ D4 mul(D4 a, D4 b) {
  return D4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w};
}
__device__ __forceinline__ D4 scale(D4 a, double s) {
  return D4{a.x * s, a.y * s, a.z * s, a.w * s};
}
__device__ __forceinline__ D4 sub(D4 a, D4 b) {
  return D4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
}
__device__ __forceinline__ D4 rgb(D4 a) { return D4{a.x, a.y, a.z, 0.0}; }

// A point of a lookup ray: radius r, r * mu, and the ray's discriminant
// roots st = sqrt(rmu^2 - r^2 + rt^2), sg = sqrt(rmu^2 - r^2 + rb^2) (0 when
// negative). rho = sqrt(r^2 - rb^2).
struct RayPoint {
  double r, rmu, st, sg, rho;
};

__device__ __forceinline__ RayPoint ray_point(const Job& J, double r, double rmu) {
  RayPoint p;
  p.r = r;
  p.rmu = rmu;
  const double q = rmu * rmu - r * r;
  p.st = sqrt(fmax(q + J.rt2, 0.0));
  p.sg = sqrt(fmax(q + J.rb2, 0.0));
  p.rho = sqrt(fmax((r - J.rb) * (r + J.rb), 0.0));
  return p;
}

// functions.glsl:240-246
__device__ __forceinline__ bool ray_hits_ground(const Job& J, const RayPoint& p) {
  return p.rmu < 0.0 && p.rmu * p.rmu - p.r * p.r + J.rb2 >= 0.0;
}

// Distance to the top boundary along the ray (functions.glsl:207-214),
// cancellation-free for rays going up.
__device__ __forceinline__ double top_distance(const Job& J, const RayPoint& p) {
  const double d = p.rmu > 0.0 ? (J.rt - p.r) * (J.rt + p.r) / (p.st + p.rmu)
                               : p.st - p.rmu;
  return d > 0.0 ? d : 0.0;
}

// x_mu of the transmittance mapping and of the sky half of the scattering
// mu mapping (functions.glsl:408-419, :800-811): (d - d_min) / (d_max - d_min).
__device__ __forceinline__ float top_unit(const Job& J, const RayPoint& p, double d_top) {
  const double d_min = J.rt - p.r;
  return static_cast<float>(d_top - d_min) / static_cast<float>(p.rho + J.H - d_min);
}

// Bilinear fetch at GL texel-space (x, y) of a 2D RGBA LUT.
__device__ __forceinline__ float3 fetch2(const float4* __restrict__ lut, int w, int h,
                                         float x, float y) {
  x = clamp_index_coord(x, static_cast<float>(w - 1));
  y = clamp_index_coord(y, static_cast<float>(h - 1));
  const float fx = floorf(x), fy = floorf(y);
  const float wx = x - fx, wy = y - fy;
  const int x0 = static_cast<int>(fx), y0 = static_cast<int>(fy);
  const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
  const float4 a = gload(lut, y0 * w + x0), b = gload(lut, y0 * w + x1);
  const float4 c = gload(lut, y1 * w + x0), d = gload(lut, y1 * w + x1);
  float3 o;
  o.x = fmaf(wy, fmaf(wx, d.x - c.x, c.x) - fmaf(wx, b.x - a.x, a.x), fmaf(wx, b.x - a.x, a.x));
  o.y = fmaf(wy, fmaf(wx, d.y - c.y, c.y) - fmaf(wx, b.y - a.y, a.y), fmaf(wx, b.y - a.y, a.y));
  o.z = fmaf(wy, fmaf(wx, d.z - c.z, c.z) - fmaf(wx, b.z - a.z, a.z), fmaf(wx, b.z - a.z, a.z));
  return o;
}

// GetTransmittanceToTopAtmosphereBoundary (functions.glsl:402-421, :473-480)
// at a ray point going up (rmu > 0) or down.
__device__ __forceinline__ float3 t_top(const Job& J, const RayPoint& p) {
  const float xm = top_unit(J, p, top_distance(J, p));
  const float xr = static_cast<float>(p.rho) * J.q_inv_H;
  return fetch2(J.T, J.tw, J.th, xm * static_cast<float>(J.tw - 1),
                xr * static_cast<float>(J.th - 1));
}

// The same point looking along the reversed ray (mu -> -mu): same
// discriminants.
__device__ __forceinline__ RayPoint reversed(RayPoint p) {
  p.rmu = -p.rmu;
  return p;
}

// The point at distance d along the ray from p (functions.glsl:499-502):
// r_d clamped to the atmosphere, mu_d = (r mu + d) / r_d clamped.
__device__ __forceinline__ RayPoint advance(const Job& J, const RayPoint& p, double d) {
  const double r_d = clamp_radius(J, sqrt(d * d + 2.0 * p.rmu * d + p.r * p.r));
  const double mu_d = clamp_cos((p.rmu + d) / r_d);
  return ray_point(J, r_d, r_d * mu_d);
}

// GetTransmittance (functions.glsl:493-519) between p and q = advance(p, d).
__device__ __forceinline__ D4 transmittance_pq(const Job& J, const RayPoint& p,
                                               const RayPoint& q, bool ground) {
  float3 num, den;
  if (ground) {
    num = t_top(J, reversed(q));
    den = t_top(J, reversed(p));
  } else {
    num = t_top(J, p);
    den = t_top(J, q);
  }
  const double x = static_cast<double>(num.x) / den.x;
  const double y = static_cast<double>(num.y) / den.y;
  const double z = static_cast<double>(num.z) / den.z;
  return D4{x < 1.0 ? x : 1.0, y < 1.0 ? y : 1.0, z < 1.0 ? z : 1.0, 0.0};
}

// Texel-space coordinates of a scattering lookup (functions.glsl:773-831):
// z (r), y (mu; ground and sky halves of MU/2 rows each), x (mu_s within a
// nu slice), s (nu slice coordinate u_nu * (NU - 1), functions.glsl:967).
struct SCoordsF {
  float z, y, x, s;
};

__device__ __forceinline__ SCoordsF s_coords_f(const Job& J, const RayPoint& p,
                                               float mu_s, float nu, bool ground) {
  SCoordsF c;
  c.z = static_cast<float>(p.rho) * J.q_inv_H * static_cast<float>(J.nr - 1);
  const float h1 = static_cast<float>(J.nmu / 2 - 1);
  if (ground) {
    const double d = -p.rmu - p.sg;
    const double d_min = p.r - J.rb;
    const double d_max = p.rho;
    const float x = d_max == d_min ? 0.0f
                                   : static_cast<float>(d - d_min) /
                                         static_cast<float>(d_max - d_min);
    c.y = h1 * (1.0f - x);
  } else {
    // The unclamped distance to the top, as in functions.glsl:807.
    const double d = p.rmu > 0.0 ? (J.rt - p.r) * (J.rt + p.r) / (p.st + p.rmu)
                                 : p.st - p.rmu;
    c.y = static_cast<float>(J.nmu / 2) + h1 * top_unit(J, p, d);
  }
  // mu_s (functions.glsl:814-827): distance to the top from the ground at
  // mu_s in cancellation-free form.
  const float tt = J.q_rb * mu_s;
  const float sq = sqrt_f(fmaf(tt, tt, J.q_H2));
  const float dd = tt > 0.0f ? J.q_H2 * rcp_f(sq + tt) : sq - tt;
  const float a = (dd - J.q_dmin) * J.q_inv_range;
  const float xm = fmaxf(fmaf(-a, J.q_inv_A, 1.0f), 0.0f) * rcp_f(1.0f + a);
  c.x = xm * static_cast<float>(J.nmus - 1);
  c.s = (nu + 1.0f) * 0.5f * static_cast<float>(J.nnu - 1);
  return c;
}

// One axis of a GL LINEAR + CLAMP_TO_EDGE fetch: two in-bounds taps and the
// weight of the second.
struct TapsF {
  int i0, i1;
  float w;
};
__device__ __forceinline__ TapsF taps_f(float x, int n) {
  x = clamp_index_coord(x, static_cast<float>(n - 1));
  const float f = floorf(x);
  TapsF t;
  t.i0 = static_cast<int>(f);
  t.i1 = min(t.i0 + 1, n - 1);
  t.w = x - f;
  return t;
}

// GetScattering (functions.glsl:958-976): two nu slices, each a trilinear
// (mu_s, mu, r) GL fetch. A nu slice coordinate past the last slice only
// gets weight 0 (the reference's x > 1 is clamped by the sampler). Sums in
// Acc (float or double), weights in FP32.
template <class Acc, class Texels>
struct Acc4 {
  Acc x, y, z, w;
};

template <class Acc, class Texels>
__device__ __forceinline__ Acc4<Acc, Texels> s_fetch(const Job& J, const Texels& lut,
                                                     const SCoordsF& c) {
  const TapsF tz = taps_f(c.z, J.nr), ty = taps_f(c.y, J.nmu), tx = taps_f(c.x, J.nmus);
  float cs = c.s;
  cs = cs >= 0.0f ? cs : 0.0f;
  cs = cs <= static_cast<float>(J.nnu - 1) ? cs : static_cast<float>(J.nnu - 1);
  const float fs = floorf(cs);
  const int s0 = static_cast<int>(fs);
  const int s1 = min(s0 + 1, J.nnu - 1);
  const float ws = cs - fs;
  const int W = J.nnu * J.nmus;
  const int rows[4] = {(tz.i0 * J.nmu + ty.i0) * W, (tz.i0 * J.nmu + ty.i1) * W,
                       (tz.i1 * J.nmu + ty.i0) * W, (tz.i1 * J.nmu + ty.i1) * W};
  const float wr[4] = {(1.0f - tz.w) * (1.0f - ty.w), (1.0f - tz.w) * ty.w,
                       tz.w * (1.0f - ty.w), tz.w * ty.w};
  const int cols[4] = {s0 * J.nmus + tx.i0, s0 * J.nmus + tx.i1, s1 * J.nmus + tx.i0,
                       s1 * J.nmus + tx.i1};
  const float wc[4] = {(1.0f - ws) * (1.0f - tx.w), (1.0f - ws) * tx.w,
                       ws * (1.0f - tx.w), ws * tx.w};
  Acc4<Acc, Texels> acc{0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const Acc wt = static_cast<Acc>(wr[a] * wc[b]);
      const float4 v = lut[static_cast<int64_t>(rows[a]) + cols[b]];
      acc.x += wt * v.x;
      acc.y += wt * v.y;
      acc.z += wt * v.z;
      acc.w += wt * v.w;
    }
  return acc;
}

// functions.glsl:1633-1645 (combined textures: Mie from the alpha channel).
__device__ __forceinline__ D4 extrapolated_mie(const Job& J, D4 s) {
  if (s.x <= 0.0) return D4{0.0, 0.0, 0.0, 0.0};
  const double k = s.w / s.x * (J.rayleigh[0] / J.mie_sca[0]);
  return D4{s.x * k * (J.mie_sca[0] / J.rayleigh[0]),
            s.y * k * (J.mie_sca[1] / J.rayleigh[1]),
            s.z * k * (J.mie_sca[2] / J.rayleigh[2]), 0.0};
}

// GetCombinedScattering (functions.glsl:1658-1690) at a ray point.
template <class Acc>
__device__ D4 combined_scattering(const Job& J, const RayPoint& p, float mu_s, float nu,
                                  bool ground, D4* single_mie) {
  const SCoordsF c = s_coords_f(J, p, mu_s, nu, ground);
  D4 s;
  if (J.half_luts) {
    const float* sc = J.scales;
    const F16Texels st{J.S16, make_float4(1.0f / sc[0], 1.0f / sc[1], 1.0f / sc[2],
                                          1.0f / sc[3])};
    const auto a = s_fetch<Acc>(J, st, c);
    s = D4{a.x, a.y, a.z, a.w};
    if (!J.combined) {
      const F16Texels mt{J.SM16, make_float4(1.0f / sc[4], 1.0f / sc[5], 1.0f / sc[6],
                                             0.0f)};
      const auto m = s_fetch<Acc>(J, mt, c);
      *single_mie = D4{m.x, m.y, m.z, 0.0};
    }
  } else {
    const auto a = s_fetch<Acc>(J, F32Texels{J.S}, c);
    s = D4{a.x, a.y, a.z, a.w};
    if (!J.combined) {
      const auto m = s_fetch<Acc>(J, F32Texels{J.SM}, c);
      *single_mie = D4{m.x, m.y, m.z, 0.0};
    }
  }
  if (J.combined) *single_mie = extrapolated_mie(J, s);
  return s;
}

// scattering * P_R(nu) + single_mie * P_M(nu) (functions.glsl:739-747,
// :1768, :1862), optionally times the sky radiance-to-luminance factors.
__device__ __forceinline__ D4 shade(const Job& J, D4 scat, D4 mie, float nu,
                                    int luminance) {
  const float x1 = fmaf(nu, nu, 1.0f);
  const float pr = J.q_kR * x1;
  const float y = fmaf(J.q_mg2, nu, J.q_g1);
  const float ry = rsq_f(y);
  const float pm = J.q_kM * x1 * ry * ry * ry;
  D4 l{scat.x * pr + mie.x * pm, scat.y * pr + mie.y * pm, scat.z * pr + mie.z * pm,
       0.0};
  if (luminance) l = D4{l.x * J.sky_k[0], l.y * J.sky_k[1], l.z * J.sky_k[2], 0.0};
  return l;
}

// The camera moved to the top boundary if it is in space and the view ray
// enters the atmosphere (functions.glsl:1715-1729, :1797-1808): returns
// false if the ray misses the atmosphere.
__device__ __forceinline__ bool enter_atmosphere(const Job& J, D3* cam, D3 v, double* r,
                                                 double* rmu) {
  const double r2 = dot3(*cam, *cam);
  *r = sqrt(r2);
  *rmu = dot3(*cam, v);
  const double to_top = -*rmu - sqrt(fmax(*rmu * *rmu - r2 + J.rt2, 0.0));
  if (to_top > 0.0) {
    *cam = D3{cam->x + v.x * to_top, cam->y + v.y * to_top, cam->z + v.z * to_top};
    *r = J.rt;
    *rmu += to_top;
    return true;
  }
  return *r <= J.rt;
}

__global__ __launch_bounds__(kBlock) void k_sky_radiance(
    const Job* __restrict__ job, int64_t n, const float* __restrict__ camera,
    const float* __restrict__ view_ray, const float* __restrict__ shadow_length,
    const float* __restrict__ sun_direction, float* __restrict__ radiance,
    float* __restrict__ transmittance_out, int luminance) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const Job& J = *job;
  D3 cam = load3(camera, i);
  const D3 v = load3(view_ray, i);
  const D3 sd = load3(sun_direction, i);
  const float shadow = shadow_length[i];
  double r, rmu;
  if (!enter_atmosphere(J, &cam, v, &r, &rmu)) {
    store3(transmittance_out, i, make_float3(1.0f, 1.0f, 1.0f));
    store3(radiance, i, make_float3(0.0f, 0.0f, 0.0f));
    return;
  }
  const RayPoint p = ray_point(J, r, rmu);
  const float mu_s = static_cast<float>(dot3(cam, sd) / r);
  const float nu = static_cast<float>(dot3(v, sd));
  const bool ground = ray_hits_ground(J, p);
  const float3 t = ground ? make_float3(0.0f, 0.0f, 0.0f) : t_top(J, p);
  D4 mie, scat;
  if (shadow == 0.0f) {
    scat = rgb(combined_scattering<float>(J, p, mu_s, nu, ground, &mie));
  } else {
    // Light shafts (functions.glsl:1746-1766): the lookups move to the end
    // of the shadowed segment, the ground flag stays the ray's.
    const double d = shadow;
    const RayPoint q = advance(J, p, d);
    const float mu_s_p = static_cast<float>((r * mu_s + d * nu) / q.r);
    scat = rgb(combined_scattering<float>(J, q, mu_s_p, nu, ground, &mie));
    const D4 st = transmittance_pq(J, p, q, ground);
    scat = mul(scat, st);
    mie = mul(mie, st);
  }
  store3(radiance, i, shade(J, scat, mie, nu, luminance));
  store3(transmittance_out, i, t);
}

__global__ __launch_bounds__(kBlock) void k_sky_radiance_to_point(
    const Job* __restrict__ job, int64_t n, const float* __restrict__ camera,
    const float* __restrict__ point, const float* __restrict__ shadow_length,
    const float* __restrict__ sun_direction, float* __restrict__ radiance,
    float* __restrict__ transmittance_out, int luminance) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const Job& J = *job;
  D3 cam = load3(camera, i);
  const D3 pt = load3(point, i);
  const D3 sd = load3(sun_direction, i);
  const double shadow = shadow_length[i];
  D3 v = D3{pt.x - cam.x, pt.y - cam.y, pt.z - cam.z};
  const double inv_len = 1.0 / sqrt(dot3(v, v));
  v = D3{v.x * inv_len, v.y * inv_len, v.z * inv_len};
  double r, rmu;
  (void)enter_atmosphere(J, &cam, v, &r, &rmu);
  const RayPoint p = ray_point(J, r, rmu);
  const double mu_s = dot3(cam, sd) / r;
  const double nu = dot3(v, sd);
  const float fnu = static_cast<float>(nu);
  const D3 pc = D3{pt.x - cam.x, pt.y - cam.y, pt.z - cam.z};
  double d = sqrt(dot3(pc, pc));
  const bool ground = ray_hits_ground(J, p);
  const RayPoint q = advance(J, p, d);
  const D4 t = transmittance_pq(J, p, q, ground);
  D4 mie;
  D4 scat = rgb(combined_scattering<double>(J, p, static_cast<float>(mu_s), fnu, ground,
                                            &mie));
  // functions.glsl:1831-1840: the far lookup at the end of the lit segment.
  d = d - shadow > 0.0 ? d - shadow : 0.0;
  const double r_p = clamp_radius(J, sqrt(d * d + 2.0 * rmu * d + r * r));
  const RayPoint pp = ray_point(J, r_p, rmu + d);
  const float mu_s_p = static_cast<float>((r * mu_s + d * nu) / r_p);
  D4 mie_p;
  const D4 scat_p = rgb(combined_scattering<double>(J, pp, mu_s_p, fnu, ground, &mie_p));
  D4 st = t;
  if (shadow > 0.0) st = transmittance_pq(J, p, advance(J, p, d), ground);
  scat = sub(scat, mul(st, scat_p));
  mie = sub(mie, mul(st, mie_p));
  if (J.combined) mie = extrapolated_mie(J, D4{scat.x, scat.y, scat.z, mie.x});
  // functions.glsl:1857-1859
  mie = scale(mie, smoothstep_d(0.0, 0.01, mu_s));
  store3(radiance, i, shade(J, scat, mie, fnu, luminance));
  store3(transmittance_out, i, t);
}

__global__ __launch_bounds__(kBlock) void k_sun_and_sky_irradiance(
    const Job* __restrict__ job, int64_t n, const float* __restrict__ point,
    const float* __restrict__ normal, const float* __restrict__ sun_direction,
    float* __restrict__ sun_out, float* __restrict__ sky_out, int luminance) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const Job& J = *job;
  const D3 p = load3(point, i);
  const D3 nrm = load3(normal, i);
  const D3 sd = load3(sun_direction, i);
  const double r = sqrt(dot3(p, p));
  const double inv_r = 1.0 / r;
  const double mu_s = dot3(p, sd) * inv_r;
  // GetIrradiance (functions.glsl:1524-1533, :1595-1601): affine mapping.
  const float xr = static_cast<float>((r - J.rb) / (J.rt - J.rb));
  const float xm = static_cast<float>(mu_s) * 0.5f + 0.5f;
  float3 sky = fetch2(J.E, J.ew, J.eh, xm * static_cast<float>(J.ew - 1),
                      xr * static_cast<float>(J.eh - 1));
  const float fs = static_cast<float>((1.0 + dot3(nrm, p) * inv_r) * 0.5);
  sky = make_float3(sky.x * fs, sky.y * fs, sky.z * fs);
  // GetTransmittanceToSun (functions.glsl:552-563).
  const double sin_h = J.rb * inv_r;
  const double cos_h = -sqrt(fmax(1.0 - sin_h * sin_h, 0.0));
  const float vis = static_cast<float>(smoothstep_d(
      -sin_h * J.sun_angular_radius, sin_h * J.sun_angular_radius, mu_s - cos_h));
  const float3 ts = t_top(J, ray_point(J, r, r * mu_s));
  const double c = dot3(nrm, sd);
  const float cf = static_cast<float>(c > 0.0 ? c : 0.0) * vis;
  float3 sun = make_float3(static_cast<float>(J.solar[0]) * ts.x * cf,
                           static_cast<float>(J.solar[1]) * ts.y * cf,
                           static_cast<float>(J.solar[2]) * ts.z * cf);
  if (luminance) {
    // GetSunAndSkyIlluminance (atmosphere/model.cc:273-281).
    sky = make_float3(sky.x * J.sky_k[0], sky.y * J.sky_k[1], sky.z * J.sky_k[2]);
    sun = make_float3(sun.x * J.sun_k[0], sun.y * J.sun_k[1], sun.z * J.sun_k[2]);
  }
  store3(sky_out, i, sky);
  store3(sun_out, i, sun);
}

// Seeded synthetic query batches. Variate k of ray i is output number
// 6 i + k + 1 of the SplitMix64 sequence started at 'seed' (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014): the
// sequence's n-th output is the finalizer applied to seed + n * golden gamma.
__device__ __forceinline__ double splitmix_uniform(uint64_t seed, uint64_t n) {
  uint64_t z = seed + n * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<double>(z >> 11) * 0x1.0p-53;
}

__global__ __launch_bounds__(kBlock) void k_generate_queries(
    uint64_t seed, int64_t n, double rb, double rt, double max_altitude,
    double max_distance, double max_sun_zenith, float* __restrict__ camera,
    float* __restrict__ view_ray, float* __restrict__ sun_direction,
    float* __restrict__ point, float* __restrict__ normal,
    float* __restrict__ shadow_length) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t base = 6ull * static_cast<uint64_t>(i) + 1ull;
  const double alt = max_altitude * splitmix_uniform(seed, base);
  const double vz = 2.0 * splitmix_uniform(seed, base + 1) - 1.0;
  const double vphi = 2.0 * kPi * splitmix_uniform(seed, base + 2);
  const double zen = max_sun_zenith * splitmix_uniform(seed, base + 3);
  const double az = 2.0 * kPi * splitmix_uniform(seed, base + 4);
  const double dist = max_distance * splitmix_uniform(seed, base + 5);
  const double vs = sqrt(fmax(0.0, 1.0 - vz * vz));
  const double v[3] = {vs * cos(vphi), vs * sin(vphi), vz};
  const double cz = rb + alt;
  double p[3] = {v[0] * dist, v[1] * dist, cz + v[2] * dist};
  const double pr = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  // Points are kept 1e-3 length units inside the atmosphere's shell.
  const double lo = rb + 1e-3, hi = rt - 1e-3;
  const double k = (pr < lo ? lo : (pr > hi ? hi : pr)) / pr;
  p[0] *= k;
  p[1] *= k;
  p[2] *= k;
  const double pn = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const double sun[3] = {sin(zen) * cos(az), sin(zen) * sin(az), cos(zen)};
  for (int c = 0; c < 3; ++c) {
    camera[3 * i + c] = c == 2 ? static_cast<float>(cz) : 0.0f;
    view_ray[3 * i + c] = static_cast<float>(v[c]);
    sun_direction[3 * i + c] = static_cast<float>(sun[c]);
    point[3 * i + c] = static_cast<float>(p[c]);
    normal[3 * i + c] = static_cast<float>(p[c] / pn);
  }
  shadow_length[i] = 0.0f;
}

inline unsigned grid_for(int64_t n) {
  return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

}  // namespace

hipError_t launch_sky_radiance(const Job* job, int64_t n, const float* camera,
                               const float* view_ray,
                               const float* shadow_length,
                               const float* sun_direction, float* radiance,
                               float* transmittance, int luminance, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sky_radiance, dim3(grid_for(n)), dim3(kBlock), 0, s, job, n,
                     camera, view_ray, shadow_length, sun_direction, radiance,
                     transmittance, luminance);
  return hipGetLastError();
}

hipError_t launch_sky_radiance_to_point(const Job* job, int64_t n,
                                        const float* camera,
                                        const float* point,
                                        const float* shadow_length,
                                        const float* sun_direction,
                                        float* radiance, float* transmittance,
                                        int luminance, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sky_radiance_to_point, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     job, n, camera, point, shadow_length, sun_direction, radiance,
                     transmittance, luminance);
  return hipGetLastError();
}

hipError_t launch_sun_and_sky_irradiance(const Job* job, int64_t n,
                                         const float* point,
                                         const float* normal,
                                         const float* sun_direction,
                                         float* sun_irradiance,
                                         float* sky_irradiance, int luminance,
                                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sun_and_sky_irradiance, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     job, n, point, normal, sun_direction, sun_irradiance,
                     sky_irradiance, luminance);
  return hipGetLastError();
}

hipError_t launch_generate_queries(uint64_t seed, int64_t n, double bottom_radius,
                                  double top_radius, double max_altitude,
                                  double max_distance, double max_sun_zenith,
                                  float* camera, float* view_ray, float* sun_direction,
                                  float* point, float* normal, float* shadow_length,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_generate_queries, dim3(grid_for(n)), dim3(kBlock), 0, s, seed, n,
                     bottom_radius, top_radius, max_altitude, max_distance,
                     max_sun_zenith, camera, view_ray, sun_direction, point, normal,
                     shadow_length);
  return hipGetLastError();
}

}  // namespace atmo
