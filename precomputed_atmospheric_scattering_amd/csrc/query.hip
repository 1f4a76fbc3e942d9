// Sky-shading lookups for gfx950: GetSkyRadiance, GetSkyRadianceToPoint and
// GetSunAndSkyIrradiance (atmosphere/functions.glsl:1633-1896) over batches
// of rays, one ray per thread. Texels are FP32; everything else is FP64: the
// scalar geometry (radius, ray/sphere discriminants, ground flag, LUT
// coordinates), where FP32 cancellation near the horizon flips the ground/sky
// choice or shifts a lookup, and the texel blending, so that the difference
// of two nearby lookups in GetSkyRadianceToPoint keeps the texels' precision.
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
__device__ __forceinline__ void store3(float* p, int64_t i, D4 v) {
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

// functions.glsl:1633-1645
__device__ __forceinline__ D4 extrapolated_mie(const Job& J, D4 s) {
  if (s.x <= 0.0) return D4{0.0, 0.0, 0.0, 0.0};
  const double k = s.w / s.x * (J.rayleigh[0] / J.mie_sca[0]);
  return D4{s.x * k * (J.mie_sca[0] / J.rayleigh[0]),
            s.y * k * (J.mie_sca[1] / J.rayleigh[1]),
            s.z * k * (J.mie_sca[2] / J.rayleigh[2]), 0.0};
}

// functions.glsl:1658-1690. kF32: FP32 tap weights and sums (lookups whose
// result is not a difference of two lookups); otherwise FP64.
template <bool kF32>
__device__ D4 combined_scattering(const Job& J, double r, double mu, double mu_s,
                                  double nu, bool ground, D4* single_mie) {
  const SCoords c = s_coords(J, r, mu, mu_s, nu, ground);
  auto lookup = [&](const auto& lut) {
    return kF32 ? s_lookup_f(J, lut, c) : s_lookup_d(J, lut, c);
  };
  auto lookup_quads = [&](const auto& lut, D4* s, D4* m) {
    if (kF32)
      s_lookup_quads_f(J, lut, c, s, m);
    else
      s_lookup_quads_d(J, lut, c, s, m);
  };
  D4 s;
  if (J.Q != nullptr && !J.combined) {
    // Lookup copy: the S and SM texels of the 4 (nu, mu_s) taps of each
    // (r, mu) tap in one contiguous quad.
    if (J.half_luts) {
      const float* sc = J.scales;
      const F16Quads qt{static_cast<const uint4*>(J.Q),
                        make_float4(1.0f / sc[0], 1.0f / sc[1], 1.0f / sc[2], 1.0f / sc[3]),
                        make_float4(1.0f / sc[4], 1.0f / sc[5], 1.0f / sc[6], 0.0f)};
      lookup_quads(qt, &s, single_mie);
    } else {
      lookup_quads(F32Quads{static_cast<const float4*>(J.Q)}, &s, single_mie);
    }
    *single_mie = rgb(*single_mie);
    return s;
  }
  if (J.half_luts) {
    const float* sc = J.scales;
    const F16Texels st{J.S16, make_float4(1.0f / sc[0], 1.0f / sc[1], 1.0f / sc[2],
                                          1.0f / sc[3])};
    s = lookup(st);
    if (!J.combined) {
      const F16Texels mt{J.SM16, make_float4(1.0f / sc[4], 1.0f / sc[5], 1.0f / sc[6],
                                             0.0f)};
      *single_mie = rgb(lookup(mt));
    }
  } else {
    s = lookup(F32Texels{J.S});
    if (!J.combined) *single_mie = rgb(lookup(F32Texels{J.SM}));
  }
  if (J.combined) *single_mie = extrapolated_mie(J, s);
  return s;
}

// scattering * P_R(nu) + single_mie * P_M(nu), optionally times the sky
// radiance-to-luminance factors.
__device__ __forceinline__ D4 shade(const Job& J, D4 scat, D4 mie, double nu,
                                    int luminance) {
  const double pr = rayleigh_phase_d(nu);
  const double pm = mie_phase_d(J.g, nu);
  D4 l{scat.x * pr + mie.x * pm, scat.y * pr + mie.y * pm, scat.z * pr + mie.z * pm,
       0.0};
  if (luminance) l = D4{l.x * J.sky_k[0], l.y * J.sky_k[1], l.z * J.sky_k[2], 0.0};
  return l;
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
    store3(transmittance_out, i, D4{1.0, 1.0, 1.0, 0.0});
    store3(radiance, i, D4{0.0, 0.0, 0.0, 0.0});
    return;
  }
  const double mu = rmu / r;
  const double mu_s = dot3(cam, sd) / r;
  const double nu = dot3(v, sd);
  const bool ground = hits_ground(J, r, mu);
  const D4 t = ground ? D4{0.0, 0.0, 0.0, 0.0} : t_lookup_d(J, J.T, r, mu);
  D4 mie, scat;
  if (shadow == 0.0) {
    scat = rgb(combined_scattering<true>(J, r, mu, mu_s, nu, ground, &mie));
  } else {
    // Light shafts (functions.glsl:1746-1766).
    const double d = shadow;
    const double r_p = clamp_radius(J, sqrt(d * d + 2.0 * r * mu * d + r * r));
    const double mu_p = (r * mu + d) / r_p;
    const double mu_s_p = (r * mu_s + d * nu) / r_p;
    scat = rgb(combined_scattering<true>(J, r_p, mu_p, mu_s_p, nu, ground, &mie));
    const D4 st = transmittance_d(J, J.T, r, mu, shadow, ground);
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
  const D4 t = transmittance_d(J, J.T, r, mu, d, ground);
  D4 mie;
  D4 scat = rgb(combined_scattering<false>(J, r, mu, mu_s, nu, ground, &mie));
  // functions.glsl:1831-1840
  d = d - shadow > 0.0 ? d - shadow : 0.0;
  const double r_p = clamp_radius(J, sqrt(d * d + 2.0 * r * mu * d + r * r));
  const double mu_p = (r * mu + d) / r_p;
  const double mu_s_p = (r * mu_s + d * nu) / r_p;
  D4 mie_p;
  const D4 scat_p = rgb(combined_scattering<false>(J, r_p, mu_p, mu_s_p, nu, ground, &mie_p));
  D4 st = t;
  if (shadow > 0.0) st = transmittance_d(J, J.T, r, mu, d, ground);
  scat = sub(scat, mul(st, scat_p));
  mie = sub(mie, mul(st, mie_p));
  if (J.combined) mie = extrapolated_mie(J, D4{scat.x, scat.y, scat.z, mie.x});
  // functions.glsl:1857-1859
  mie = scale(mie, smoothstep_d(0.0, 0.01, mu_s));
  store3(radiance, i, shade(J, scat, mie, nu, luminance));
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
  const double mu_s = dot3(p, sd) / r;
  // Two lookups, neither a difference: FP32 weights and sums (as in
  // GetSkyRadiance).
  D4 sky = scale(e_lookup_d<float>(J, J.E, r, mu_s), (1.0 + dot3(nrm, p) / r) * 0.5);
  const D4 ts = transmittance_to_sun_d<float>(J, J.T, r, mu_s);
  const double c = dot3(nrm, sd);
  const double cf = c > 0.0 ? c : 0.0;
  D4 sun{J.solar[0] * ts.x * cf, J.solar[1] * ts.y * cf, J.solar[2] * ts.z * cf, 0.0};
  if (luminance) {
    // GetSunAndSkyIlluminance (atmosphere/model.cc:273-281).
    sky = D4{sky.x * J.sky_k[0], sky.y * J.sky_k[1], sky.z * J.sky_k[2], 0.0};
    sun = D4{sun.x * J.sun_k[0], sun.y * J.sun_k[1], sun.z * J.sun_k[2], 0.0};
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
