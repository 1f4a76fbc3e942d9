// Sky-shading lookups for gfx950: GetSkyRadiance, GetSkyRadianceToPoint and
// GetSunAndSkyIrradiance (atmosphere/functions.glsl:1633-1896) over batches
// of rays, one ray per thread. The scalar geometry (radius, ray/sphere
// discriminants, ground flag, LUT coordinates) is FP64, because those are
// where FP32 cancellation near the horizon flips the ground/sky choice or
// shifts a lookup; texel fetches and blending are FP32.
#include <hip/hip_runtime.h>

#include "atmo_device.h"
#include "kernels.h"

namespace atmo {
namespace {

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 load3(const float* p, int64_t i) {
  return D3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
__device__ __forceinline__ void store3(float* p, int64_t i, float3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ __forceinline__ double dot3(D3 a, D3 b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ float3 mul3(float3 a, float3 b) {
  return make_float3(a.x * b.x, a.y * b.y, a.z * b.z);
}
__device__ __forceinline__ float3 scale3(float3 a, float s) {
  return make_float3(a.x * s, a.y * s, a.z * s);
}
__device__ __forceinline__ float3 sub3(float3 a, float3 b) {
  return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}

// functions.glsl:1633-1645
__device__ __forceinline__ float3 extrapolated_mie(const Job& J, float4 s) {
  if (s.x <= 0.0f) return make_float3(0.f, 0.f, 0.f);
  const float k = s.w / s.x * static_cast<float>(J.rayleigh[0] / J.mie_sca[0]);
  return make_float3(s.x * k * static_cast<float>(J.mie_sca[0] / J.rayleigh[0]),
                     s.y * k * static_cast<float>(J.mie_sca[1] / J.rayleigh[1]),
                     s.z * k * static_cast<float>(J.mie_sca[2] / J.rayleigh[2]));
}

// functions.glsl:1658-1690
__device__ float4 combined_scattering(const Job& J, double r, double mu,
                                      double mu_s, double nu, bool ground,
                                      float3* single_mie) {
  const SCoords c = s_coords(J, r, mu, mu_s, nu, ground);
  const float4 s = s_lookup(J, J.S, c);
  if (J.combined) {
    *single_mie = extrapolated_mie(J, s);
  } else {
    *single_mie = make3(s_lookup(J, J.SM, c));
  }
  return s;
}

__global__ __launch_bounds__(kBlock) void k_sky_radiance(
    const Job* __restrict__ job, int64_t n, const float* __restrict__ camera,
    const float* __restrict__ view_ray, const float* __restrict__ shadow_length,
    const float* __restrict__ sun_direction, float* __restrict__ radiance,
    float* __restrict__ transmittance_out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const Job& J = *job;
  D3 cam = load3(camera, i);
  const D3 v = load3(view_ray, i);
  const D3 sd = load3(sun_direction, i);
  const double shadow = shadow_length[i];
  // functions.glsl:1715-1729
  double r = sqrt(dot3(cam, cam));
  double rmu = dot3(cam, v);
  const double to_top = -rmu - sqrt(rmu * rmu - r * r + J.rt * J.rt);
  if (to_top > 0.0) {
    cam = D3{cam.x + v.x * to_top, cam.y + v.y * to_top, cam.z + v.z * to_top};
    r = J.rt;
    rmu += to_top;
  } else if (r > J.rt) {
    store3(transmittance_out, i, make_float3(1.f, 1.f, 1.f));
    store3(radiance, i, make_float3(0.f, 0.f, 0.f));
    return;
  }
  const double mu = rmu / r;
  const double mu_s = dot3(cam, sd) / r;
  const double nu = dot3(v, sd);
  const bool ground = hits_ground(J, r, mu);
  const float3 t = ground ? make_float3(0.f, 0.f, 0.f) : t_lookup(J, J.T, r, mu);
  float3 mie;
  float3 scat;
  if (shadow == 0.0) {
    scat = make3(combined_scattering(J, r, mu, mu_s, nu, ground, &mie));
  } else {
    // Light shafts (functions.glsl:1746-1766).
    const double d = shadow;
    const double r_p = clamp_radius(J, sqrt(d * d + 2.0 * r * mu * d + r * r));
    const double mu_p = (r * mu + d) / r_p;
    const double mu_s_p = (r * mu_s + d * nu) / r_p;
    scat = make3(combined_scattering(J, r_p, mu_p, mu_s_p, nu, ground, &mie));
    const float3 st = transmittance(J, J.T, r, mu, shadow, ground);
    scat = mul3(scat, st);
    mie = mul3(mie, st);
  }
  const float fnu = static_cast<float>(nu);
  const float pr = rayleigh_phase(fnu);
  const float pm = mie_phase(static_cast<float>(J.g), fnu);
  store3(radiance, i,
         make_float3(scat.x * pr + mie.x * pm, scat.y * pr + mie.y * pm,
                     scat.z * pr + mie.z * pm));
  store3(transmittance_out, i, t);
}

__global__ __launch_bounds__(kBlock) void k_sky_radiance_to_point(
    const Job* __restrict__ job, int64_t n, const float* __restrict__ camera,
    const float* __restrict__ point, const float* __restrict__ shadow_length,
    const float* __restrict__ sun_direction, float* __restrict__ radiance,
    float* __restrict__ transmittance_out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const Job& J = *job;
  D3 cam = load3(camera, i);
  const D3 p = load3(point, i);
  const D3 sd = load3(sun_direction, i);
  const double shadow = shadow_length[i];
  // functions.glsl:1797-1808
  D3 v = D3{p.x - cam.x, p.y - cam.y, p.z - cam.z};
  const double inv_len = 1.0 / sqrt(dot3(v, v));
  v = D3{v.x * inv_len, v.y * inv_len, v.z * inv_len};
  double r = sqrt(dot3(cam, cam));
  double rmu = dot3(cam, v);
  const double to_top = -rmu - sqrt(rmu * rmu - r * r + J.rt * J.rt);
  if (to_top > 0.0) {
    cam = D3{cam.x + v.x * to_top, cam.y + v.y * to_top, cam.z + v.z * to_top};
    r = J.rt;
    rmu += to_top;
  }
  const double mu = rmu / r;
  const double mu_s = dot3(cam, sd) / r;
  const double nu = dot3(v, sd);
  const D3 pc = D3{p.x - cam.x, p.y - cam.y, p.z - cam.z};
  double d = sqrt(dot3(pc, pc));
  const bool ground = hits_ground(J, r, mu);
  const float3 t = transmittance(J, J.T, r, mu, d, ground);
  float3 mie;
  float3 scat = make3(combined_scattering(J, r, mu, mu_s, nu, ground, &mie));
  // functions.glsl:1831-1840
  d = d - shadow > 0.0 ? d - shadow : 0.0;
  const double r_p = clamp_radius(J, sqrt(d * d + 2.0 * r * mu * d + r * r));
  const double mu_p = (r * mu + d) / r_p;
  const double mu_s_p = (r * mu_s + d * nu) / r_p;
  float3 mie_p;
  const float3 scat_p = make3(combined_scattering(J, r_p, mu_p, mu_s_p, nu, ground, &mie_p));
  float3 st = t;
  if (shadow > 0.0) st = transmittance(J, J.T, r, mu, d, ground);
  scat = sub3(scat, mul3(st, scat_p));
  mie = sub3(mie, mul3(st, mie_p));
  if (J.combined) {
    mie = extrapolated_mie(J, make_float4(scat.x, scat.y, scat.z, mie.x));
  }
  // functions.glsl:1857-1859
  mie = scale3(mie, static_cast<float>(smoothstep_d(0.0, 0.01, mu_s)));
  const float fnu = static_cast<float>(nu);
  const float pr = rayleigh_phase(fnu);
  const float pm = mie_phase(static_cast<float>(J.g), fnu);
  store3(radiance, i,
         make_float3(scat.x * pr + mie.x * pm, scat.y * pr + mie.y * pm,
                     scat.z * pr + mie.z * pm));
  store3(transmittance_out, i, t);
}

__global__ __launch_bounds__(kBlock) void k_sun_and_sky_irradiance(
    const Job* __restrict__ job, int64_t n, const float* __restrict__ point,
    const float* __restrict__ normal, const float* __restrict__ sun_direction,
    float* __restrict__ sun_out, float* __restrict__ sky_out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const Job& J = *job;
  const D3 p = load3(point, i);
  const D3 nrm = load3(normal, i);
  const D3 sd = load3(sun_direction, i);
  const double r = sqrt(dot3(p, p));
  const double mu_s = dot3(p, sd) / r;
  const float3 e = e_lookup(J, J.E, r, mu_s);
  store3(sky_out, i, scale3(e, static_cast<float>((1.0 + dot3(nrm, p) / r) * 0.5)));
  const float3 ts = transmittance_to_sun(J, J.T, r, mu_s);
  const double c = dot3(nrm, sd);
  const float cf = static_cast<float>(c > 0.0 ? c : 0.0);
  store3(sun_out, i,
         make_float3(static_cast<float>(J.solar[0]) * ts.x * cf,
                     static_cast<float>(J.solar[1]) * ts.y * cf,
                     static_cast<float>(J.solar[2]) * ts.z * cf));
}

inline unsigned grid_for(int64_t n) {
  return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

}  // namespace

hipError_t launch_sky_radiance(const Job* job, int64_t n, const float* camera,
                               const float* view_ray,
                               const float* shadow_length,
                               const float* sun_direction, float* radiance,
                               float* transmittance, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sky_radiance, dim3(grid_for(n)), dim3(kBlock), 0, s, job, n,
                     camera, view_ray, shadow_length, sun_direction, radiance,
                     transmittance);
  return hipGetLastError();
}

hipError_t launch_sky_radiance_to_point(const Job* job, int64_t n,
                                        const float* camera,
                                        const float* point,
                                        const float* shadow_length,
                                        const float* sun_direction,
                                        float* radiance, float* transmittance,
                                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sky_radiance_to_point, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     job, n, camera, point, shadow_length, sun_direction, radiance,
                     transmittance);
  return hipGetLastError();
}

hipError_t launch_sun_and_sky_irradiance(const Job* job, int64_t n,
                                         const float* point,
                                         const float* normal,
                                         const float* sun_direction,
                                         float* sun_irradiance,
                                         float* sky_irradiance, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sun_and_sky_irradiance, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     job, n, point, normal, sun_direction, sun_irradiance,
                     sky_irradiance);
  return hipGetLastError();
}

}  // namespace atmo
