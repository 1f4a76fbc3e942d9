// FP64 CPU oracle for the atmosphere LUT precomputation and sky lookups.
//
// TEST INFRASTRUCTURE ONLY: the reference results the HIP product is
// compared against.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it. See oracle.h for the scope, layout and conventions.
//
// Every function cites the reference lines it restates (paths relative to the
// reference repository root). The code is a from-scratch double-precision
// restatement of that mathematics; it shares no source with the reference.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "../precomputed_atmospheric_scattering_amd/csrc/cie_1931.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Small value types.
// ---------------------------------------------------------------------------

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double Length(V3 a) { return std::sqrt(Dot(a, a)); }
inline V3 Normalize(V3 a) { return a * (1.0 / Length(a)); }

// A spectral value at the three precomputed wavelengths, plus a fourth
// channel used by the packed RGBA scattering texel.
struct C4 {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  double& operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
};
inline C4 Splat3(double s) {
  C4 c;
  c[0] = c[1] = c[2] = s;
  return c;
}
inline C4 From3(const double* p) {
  C4 c;
  c[0] = p[0];
  c[1] = p[1];
  c[2] = p[2];
  return c;
}
inline C4 Add(const C4& a, const C4& b) {
  C4 c;
  for (int i = 0; i < 4; ++i) c[i] = a[i] + b[i];
  return c;
}
inline C4 Sub(const C4& a, const C4& b) {
  C4 c;
  for (int i = 0; i < 4; ++i) c[i] = a[i] - b[i];
  return c;
}
inline C4 Mul(const C4& a, const C4& b) {
  C4 c;
  for (int i = 0; i < 4; ++i) c[i] = a[i] * b[i];
  return c;
}
inline C4 Scale(const C4& a, double s) {
  C4 c;
  for (int i = 0; i < 4; ++i) c[i] = a[i] * s;
  return c;
}
inline void Put3(const C4& c, double* out) {
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
}

using Atmo = or_atmosphere;
using Sizes = or_sizes;

// ---------------------------------------------------------------------------
// Thread pool: a parallel loop over rows or layers, the role RunJobs plays in
// atmosphere/reference/model.cc:140-237.
// ---------------------------------------------------------------------------

int g_threads = 0;

int ThreadCount() {
  if (g_threads > 0) return g_threads;
  const char* env = std::getenv("ORACLE_THREADS");
  if (env != nullptr && std::atoi(env) > 0) return std::atoi(env);
  // Default: the host's threads, at most 16, so that a run shares its host
  // with the GPU work beside it.
  unsigned hc = std::thread::hardware_concurrency();
  return hc == 0 ? 1 : static_cast<int>(hc < 16 ? hc : 16);
}

template <class F>
void ParallelFor(int n, F&& body) {
  int threads = std::min(ThreadCount(), n);
  if (threads <= 1) {
    for (int i = 0; i < n; ++i) body(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&]() {
      for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) body(i);
    });
  }
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// OpenGL 3.3 texture() with LINEAR filtering and CLAMP_TO_EDGE wrapping
// (GL 3.3 core specification, section 3.8.11): for each axis of size n the
// texel-space coordinate is x = u * n - 1/2, the taps are floor(x) and
// floor(x) + 1 clamped to [0, n - 1], and the weight of the upper tap is
// x - floor(x). Weights are exact FP64 (no fixed-point filter weights).
// ---------------------------------------------------------------------------

struct Taps {
  int i0, i1;
  double w;
};

inline Taps AxisTaps(double u, int n) {
  double x = u * n - 0.5;
  double f = std::floor(x);
  Taps t;
  t.w = x - f;
  int i = static_cast<int>(f);
  t.i0 = std::min(std::max(i, 0), n - 1);
  t.i1 = std::min(std::max(i + 1, 0), n - 1);
  return t;
}

inline C4 Fetch(const double* lut, size_t texel) {
  C4 c;
  const double* p = lut + texel * 4;
  c[0] = p[0];
  c[1] = p[1];
  c[2] = p[2];
  c[3] = p[3];
  return c;
}

C4 Texture2D(const double* lut, int w, int h, double u, double v) {
  Taps tx = AxisTaps(u, w), ty = AxisTaps(v, h);
  C4 r;
  const int xs[2] = {tx.i0, tx.i1};
  const int ys[2] = {ty.i0, ty.i1};
  const double wx[2] = {1.0 - tx.w, tx.w};
  const double wy[2] = {1.0 - ty.w, ty.w};
  for (int b = 0; b < 2; ++b)
    for (int a = 0; a < 2; ++a) {
      C4 t = Fetch(lut, static_cast<size_t>(ys[b]) * w + xs[a]);
      r = Add(r, Scale(t, wx[a] * wy[b]));
    }
  return r;
}

C4 Texture3D(const double* lut, int w, int h, int d, double u, double v,
             double s) {
  Taps tx = AxisTaps(u, w), ty = AxisTaps(v, h), tz = AxisTaps(s, d);
  C4 r;
  const int xs[2] = {tx.i0, tx.i1};
  const int ys[2] = {ty.i0, ty.i1};
  const int zs[2] = {tz.i0, tz.i1};
  const double wx[2] = {1.0 - tx.w, tx.w};
  const double wy[2] = {1.0 - ty.w, ty.w};
  const double wz[2] = {1.0 - tz.w, tz.w};
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < 2; ++b)
      for (int a = 0; a < 2; ++a) {
        size_t texel =
            (static_cast<size_t>(zs[c]) * h + ys[b]) * w + xs[a];
        r = Add(r, Scale(Fetch(lut, texel), wx[a] * wy[b] * wz[c]));
      }
  return r;
}

inline int ScatteringWidth(const Sizes& s) {
  return s.scattering_nu * s.scattering_mu_s;
}

// ---------------------------------------------------------------------------
// functions.glsl: utilities and geometry.
// ---------------------------------------------------------------------------

// functions.glsl:113-127
inline double ClampCosine(double mu) { return std::min(std::max(mu, -1.0), 1.0); }
inline double ClampDistance(double d) { return std::max(d, 0.0); }
inline double ClampRadius(const Atmo& a, double r) {
  return std::min(std::max(r, a.bottom_radius), a.top_radius);
}
inline double SafeSqrt(double area) { return std::sqrt(std::max(area, 0.0)); }

// functions.glsl:207-214
double DistanceToTopAtmosphereBoundary(const Atmo& a, double r, double mu) {
  double disc = r * r * (mu * mu - 1.0) + a.top_radius * a.top_radius;
  return ClampDistance(-r * mu + SafeSqrt(disc));
}

// functions.glsl:222-229
double DistanceToBottomAtmosphereBoundary(const Atmo& a, double r, double mu) {
  double disc = r * r * (mu * mu - 1.0) + a.bottom_radius * a.bottom_radius;
  return ClampDistance(-r * mu - SafeSqrt(disc));
}

// functions.glsl:240-246
bool RayIntersectsGround(const Atmo& a, double r, double mu) {
  return mu < 0.0 &&
         r * r * (mu * mu - 1.0) + a.bottom_radius * a.bottom_radius >= 0.0;
}

// functions.glsl:263-267
double GetLayerDensity(const or_layer& layer, double altitude) {
  double density = layer.exp_term * std::exp(layer.exp_scale * altitude) +
                   layer.linear_term * altitude + layer.constant_term;
  return std::min(std::max(density, 0.0), 1.0);
}

// functions.glsl:269-273 (a profile is exactly two layers)
double GetProfileDensity(const or_layer* profile, double altitude) {
  return altitude < profile[0].width ? GetLayerDensity(profile[0], altitude)
                                     : GetLayerDensity(profile[1], altitude);
}

// functions.glsl:275-299: trapezoidal rule, 500 intervals.
double ComputeOpticalLengthToTopAtmosphereBoundary(const Atmo& a,
                                                   const or_layer* profile,
                                                   double r, double mu) {
  const int kIntervals = 500;
  double dx = DistanceToTopAtmosphereBoundary(a, r, mu) / kIntervals;
  double result = 0.0;
  for (int i = 0; i <= kIntervals; ++i) {
    double d_i = i * dx;
    double r_i = std::sqrt(d_i * d_i + 2.0 * r * mu * d_i + r * r);
    double y_i = GetProfileDensity(profile, r_i - a.bottom_radius);
    double weight_i = (i == 0 || i == kIntervals) ? 0.5 : 1.0;
    result += y_i * weight_i * dx;
  }
  return result;
}

// functions.glsl:306-320
C4 ComputeTransmittanceToTopAtmosphereBoundary(const Atmo& a, double r,
                                               double mu) {
  double lr = ComputeOpticalLengthToTopAtmosphereBoundary(
      a, a.rayleigh_density, r, mu);
  double lm =
      ComputeOpticalLengthToTopAtmosphereBoundary(a, a.mie_density, r, mu);
  double la = ComputeOpticalLengthToTopAtmosphereBoundary(
      a, a.absorption_density, r, mu);
  C4 t;
  for (int c = 0; c < 3; ++c) {
    t[c] = std::exp(-(a.rayleigh_scattering[c] * lr + a.mie_extinction[c] * lm +
                      a.absorption_extinction[c] * la));
  }
  return t;
}

// functions.glsl:342-348
inline double GetTextureCoordFromUnitRange(double x, int n) {
  return 0.5 / n + x * (1.0 - 1.0 / n);
}
inline double GetUnitRangeFromTextureCoord(double u, int n) {
  return (u - 0.5 / n) / (1.0 - 1.0 / n);
}

// functions.glsl:402-421
void GetTransmittanceTextureUvFromRMu(const Atmo& a, const Sizes& s, double r,
                                      double mu, double* u, double* v) {
  double H = std::sqrt(a.top_radius * a.top_radius -
                       a.bottom_radius * a.bottom_radius);
  double rho = SafeSqrt(r * r - a.bottom_radius * a.bottom_radius);
  double d = DistanceToTopAtmosphereBoundary(a, r, mu);
  double d_min = a.top_radius - r;
  double d_max = rho + H;
  double x_mu = (d - d_min) / (d_max - d_min);
  double x_r = rho / H;
  *u = GetTextureCoordFromUnitRange(x_mu, s.transmittance_width);
  *v = GetTextureCoordFromUnitRange(x_r, s.transmittance_height);
}

// functions.glsl:427-447
void GetRMuFromTransmittanceTextureUv(const Atmo& a, const Sizes& s, double u,
                                      double v, double* r, double* mu) {
  double x_mu = GetUnitRangeFromTextureCoord(u, s.transmittance_width);
  double x_r = GetUnitRangeFromTextureCoord(v, s.transmittance_height);
  double H = std::sqrt(a.top_radius * a.top_radius -
                       a.bottom_radius * a.bottom_radius);
  double rho = H * x_r;
  *r = std::sqrt(rho * rho + a.bottom_radius * a.bottom_radius);
  double d_min = a.top_radius - *r;
  double d_max = rho + H;
  double d = d_min + x_mu * (d_max - d_min);
  double m = d == 0.0 ? 1.0 : (H * H - rho * rho - d * d) / (2.0 * *r * d);
  *mu = ClampCosine(m);
}

// functions.glsl:454-463 (texel (i, j) has frag_coord (i + 1/2, j + 1/2))
C4 ComputeTransmittanceToTopAtmosphereBoundaryTexture(const Atmo& a,
                                                      const Sizes& s, int i,
                                                      int j) {
  double r, mu;
  GetRMuFromTransmittanceTextureUv(a, s, (i + 0.5) / s.transmittance_width,
                                   (j + 0.5) / s.transmittance_height, &r,
                                   &mu);
  return ComputeTransmittanceToTopAtmosphereBoundary(a, r, mu);
}

// functions.glsl:473-480
C4 GetTransmittanceToTopAtmosphereBoundary(const Atmo& a, const Sizes& s,
                                           const double* t_lut, double r,
                                           double mu) {
  double u, v;
  GetTransmittanceTextureUvFromRMu(a, s, r, mu, &u, &v);
  C4 t = Texture2D(t_lut, s.transmittance_width, s.transmittance_height, u, v);
  t[3] = 0.0;
  return t;
}

inline C4 MinOne(const C4& x) {
  C4 r;
  for (int c = 0; c < 3; ++c) r[c] = std::min(x[c], 1.0);  // GLSL min(x, 1)
  return r;
}
inline C4 Div3(const C4& a, const C4& b) {
  C4 r;
  for (int c = 0; c < 3; ++c) r[c] = a[c] / b[c];
  return r;
}

// functions.glsl:493-519
C4 GetTransmittance(const Atmo& a, const Sizes& s, const double* t_lut,
                    double r, double mu, double d, bool ground) {
  double r_d = ClampRadius(a, std::sqrt(d * d + 2.0 * r * mu * d + r * r));
  double mu_d = ClampCosine((r * mu + d) / r_d);
  if (ground) {
    return MinOne(
        Div3(GetTransmittanceToTopAtmosphereBoundary(a, s, t_lut, r_d, -mu_d),
             GetTransmittanceToTopAtmosphereBoundary(a, s, t_lut, r, -mu)));
  }
  return MinOne(
      Div3(GetTransmittanceToTopAtmosphereBoundary(a, s, t_lut, r, mu),
           GetTransmittanceToTopAtmosphereBoundary(a, s, t_lut, r_d, mu_d)));
}

inline double SmoothStep(double e0, double e1, double x) {
  double t = std::min(std::max((x - e0) / (e1 - e0), 0.0), 1.0);
  return t * t * (3.0 - 2.0 * t);
}

// functions.glsl:552-563
C4 GetTransmittanceToSun(const Atmo& a, const Sizes& s, const double* t_lut,
                         double r, double mu_s) {
  double sin_theta_h = a.bottom_radius / r;
  double cos_theta_h = -std::sqrt(std::max(1.0 - sin_theta_h * sin_theta_h, 0.0));
  double visible = SmoothStep(-sin_theta_h * a.sun_angular_radius,
                              sin_theta_h * a.sun_angular_radius,
                              mu_s - cos_theta_h);
  return Scale(GetTransmittanceToTopAtmosphereBoundary(a, s, t_lut, r, mu_s),
               visible);
}

// ---------------------------------------------------------------------------
// Single scattering.
// ---------------------------------------------------------------------------

// functions.glsl:650-668
void ComputeSingleScatteringIntegrand(const Atmo& a, const Sizes& s,
                                      const double* t_lut, double r, double mu,
                                      double mu_s, double nu, double d,
                                      bool ground, C4* rayleigh, C4* mie) {
  double r_d = ClampRadius(a, std::sqrt(d * d + 2.0 * r * mu * d + r * r));
  double mu_s_d = ClampCosine((r * mu_s + d * nu) / r_d);
  C4 transmittance =
      Mul(GetTransmittance(a, s, t_lut, r, mu, d, ground),
          GetTransmittanceToSun(a, s, t_lut, r_d, mu_s_d));
  *rayleigh = Scale(transmittance,
                    GetProfileDensity(a.rayleigh_density, r_d - a.bottom_radius));
  *mie = Scale(transmittance,
               GetProfileDensity(a.mie_density, r_d - a.bottom_radius));
}

// functions.glsl:680-687
double DistanceToNearestAtmosphereBoundary(const Atmo& a, double r, double mu,
                                           bool ground) {
  return ground ? DistanceToBottomAtmosphereBoundary(a, r, mu)
                : DistanceToTopAtmosphereBoundary(a, r, mu);
}

// functions.glsl:695-730: trapezoidal rule, 50 intervals.
void ComputeSingleScattering(const Atmo& a, const Sizes& s,
                             const double* t_lut, double r, double mu,
                             double mu_s, double nu, bool ground, C4* rayleigh,
                             C4* mie) {
  const int kIntervals = 50;
  double dx = DistanceToNearestAtmosphereBoundary(a, r, mu, ground) / kIntervals;
  C4 rayleigh_sum, mie_sum;
  for (int i = 0; i <= kIntervals; ++i) {
    double d_i = i * dx;
    C4 ri, mi;
    ComputeSingleScatteringIntegrand(a, s, t_lut, r, mu, mu_s, nu, d_i, ground,
                                     &ri, &mi);
    double w = (i == 0 || i == kIntervals) ? 0.5 : 1.0;
    rayleigh_sum = Add(rayleigh_sum, Scale(ri, w));
    mie_sum = Add(mie_sum, Scale(mi, w));
  }
  C4 r_out, m_out;
  for (int c = 0; c < 3; ++c) {
    r_out[c] = rayleigh_sum[c] * dx * a.solar_irradiance[c] *
               a.rayleigh_scattering[c];
    m_out[c] = mie_sum[c] * dx * a.solar_irradiance[c] * a.mie_scattering[c];
  }
  *rayleigh = r_out;
  *mie = m_out;
}

// functions.glsl:739-747
double RayleighPhaseFunction(double nu) {
  double k = 3.0 / (16.0 * kPi);
  return k * (1.0 + nu * nu);
}
double MiePhaseFunction(double g, double nu) {
  double k = 3.0 / (8.0 * kPi) * (1.0 - g * g) / (2.0 + g * g);
  return k * (1.0 + nu * nu) / std::pow(1.0 + g * g - 2.0 * g * nu, 1.5);
}

// ---------------------------------------------------------------------------
// 4D scattering LUT parameterisation.
// ---------------------------------------------------------------------------

struct Uvwz {
  double u_nu, u_mu_s, u_mu, u_r;
};

// functions.glsl:773-831
Uvwz GetScatteringTextureUvwzFromRMuMuSNu(const Atmo& a, const Sizes& s,
                                          double r, double mu, double mu_s,
                                          double nu, bool ground) {
  double H = std::sqrt(a.top_radius * a.top_radius -
                       a.bottom_radius * a.bottom_radius);
  double rho = SafeSqrt(r * r - a.bottom_radius * a.bottom_radius);
  Uvwz out;
  out.u_r = GetTextureCoordFromUnitRange(rho / H, s.scattering_r);

  double r_mu = r * mu;
  double disc = r_mu * r_mu - r * r + a.bottom_radius * a.bottom_radius;
  if (ground) {
    double d = -r_mu - SafeSqrt(disc);
    double d_min = r - a.bottom_radius;
    double d_max = rho;
    double x = d_max == d_min ? 0.0 : (d - d_min) / (d_max - d_min);
    out.u_mu = 0.5 - 0.5 * GetTextureCoordFromUnitRange(x, s.scattering_mu / 2);
  } else {
    double d = -r_mu + SafeSqrt(disc + H * H);
    double d_min = a.top_radius - r;
    double d_max = rho + H;
    out.u_mu = 0.5 + 0.5 * GetTextureCoordFromUnitRange(
                             (d - d_min) / (d_max - d_min), s.scattering_mu / 2);
  }

  double d = DistanceToTopAtmosphereBoundary(a, a.bottom_radius, mu_s);
  double d_min = a.top_radius - a.bottom_radius;
  double d_max = H;
  double a_ = (d - d_min) / (d_max - d_min);
  double D = DistanceToTopAtmosphereBoundary(a, a.bottom_radius, a.mu_s_min);
  double A = (D - d_min) / (d_max - d_min);
  out.u_mu_s = GetTextureCoordFromUnitRange(
      std::max(1.0 - a_ / A, 0.0) / (1.0 + a_), s.scattering_mu_s);
  out.u_nu = (nu + 1.0) / 2.0;
  return out;
}

// functions.glsl:837-890
void GetRMuMuSNuFromScatteringTextureUvwz(const Atmo& a, const Sizes& s,
                                          const Uvwz& t, double* r, double* mu,
                                          double* mu_s, double* nu,
                                          bool* ground) {
  double H = std::sqrt(a.top_radius * a.top_radius -
                       a.bottom_radius * a.bottom_radius);
  double rho = H * GetUnitRangeFromTextureCoord(t.u_r, s.scattering_r);
  *r = std::sqrt(rho * rho + a.bottom_radius * a.bottom_radius);
  if (t.u_mu < 0.5) {
    double d_min = *r - a.bottom_radius;
    double d_max = rho;
    double d = d_min + (d_max - d_min) * GetUnitRangeFromTextureCoord(
                           1.0 - 2.0 * t.u_mu, s.scattering_mu / 2);
    *mu = d == 0.0 ? -1.0 : ClampCosine(-(rho * rho + d * d) / (2.0 * *r * d));
    *ground = true;
  } else {
    double d_min = a.top_radius - *r;
    double d_max = rho + H;
    double d = d_min + (d_max - d_min) * GetUnitRangeFromTextureCoord(
                           2.0 * t.u_mu - 1.0, s.scattering_mu / 2);
    *mu = d == 0.0 ? 1.0
                   : ClampCosine((H * H - rho * rho - d * d) / (2.0 * *r * d));
    *ground = false;
  }
  double x_mu_s = GetUnitRangeFromTextureCoord(t.u_mu_s, s.scattering_mu_s);
  double d_min = a.top_radius - a.bottom_radius;
  double d_max = H;
  double D = DistanceToTopAtmosphereBoundary(a, a.bottom_radius, a.mu_s_min);
  double A = (D - d_min) / (d_max - d_min);
  double a_ = (A - x_mu_s * A) / (1.0 + x_mu_s * A);
  double d = d_min + std::min(a_, A) * (d_max - d_min);
  *mu_s = d == 0.0 ? 1.0
                   : ClampCosine((H * H - d * d) / (2.0 * a.bottom_radius * d));
  *nu = ClampCosine(t.u_nu * 2.0 - 1.0);
}

// functions.glsl:905-926. Texel (x, y, z) has frag_coord (x+1/2, y+1/2, z+1/2);
// the nu index is floor(frag.x / MU_S) and is sampled endpoint-exact over NU-1.
void GetRMuMuSNuFromScatteringTextureFragCoord(const Atmo& a, const Sizes& s,
                                               int x, int y, int z, double* r,
                                               double* mu, double* mu_s,
                                               double* nu, bool* ground) {
  double fx = x + 0.5, fy = y + 0.5, fz = z + 0.5;
  double frag_nu = std::floor(fx / s.scattering_mu_s);
  double frag_mu_s = fx - s.scattering_mu_s * std::floor(fx / s.scattering_mu_s);
  Uvwz t;
  t.u_nu = frag_nu / (s.scattering_nu - 1);
  t.u_mu_s = frag_mu_s / s.scattering_mu_s;
  t.u_mu = fy / s.scattering_mu;
  t.u_r = fz / s.scattering_r;
  GetRMuMuSNuFromScatteringTextureUvwz(a, s, t, r, mu, mu_s, nu, ground);
  double m = *mu, ms = *mu_s;
  double half = std::sqrt((1.0 - m * m) * (1.0 - ms * ms));
  *nu = std::min(std::max(*nu, m * ms - half), m * ms + half);
}

// functions.glsl:958-976: manual lerp between two nu slices, each slice a
// trilinear texture() over (mu_s within the slice, mu, r).
C4 GetScattering(const Atmo& a, const Sizes& s, const double* lut, double r,
                 double mu, double mu_s, double nu, bool ground) {
  Uvwz t = GetScatteringTextureUvwzFromRMuMuSNu(a, s, r, mu, mu_s, nu, ground);
  double tex_coord_x = t.u_nu * (s.scattering_nu - 1);
  double tex_x = std::floor(tex_coord_x);
  double lerp = tex_coord_x - tex_x;
  double u0 = (tex_x + t.u_mu_s) / s.scattering_nu;
  double u1 = (tex_x + 1.0 + t.u_mu_s) / s.scattering_nu;
  int w = ScatteringWidth(s);
  C4 s0 = Texture3D(lut, w, s.scattering_mu, s.scattering_r, u0, t.u_mu, t.u_r);
  C4 s1 = Texture3D(lut, w, s.scattering_mu, s.scattering_r, u1, t.u_mu, t.u_r);
  return Add(Scale(s0, 1.0 - lerp), Scale(s1, lerp));
}

// functions.glsl:987-1009
C4 GetScatteringOfOrder(const Atmo& a, const Sizes& s,
                        const double* single_rayleigh, const double* single_mie,
                        const double* multiple, double r, double mu,
                        double mu_s, double nu, bool ground, int order) {
  if (order == 1) {
    C4 ray = GetScattering(a, s, single_rayleigh, r, mu, mu_s, nu, ground);
    C4 mie = GetScattering(a, s, single_mie, r, mu, mu_s, nu, ground);
    C4 out = Add(Scale(ray, RayleighPhaseFunction(nu)),
                 Scale(mie, MiePhaseFunction(a.mie_phase_function_g, nu)));
    out[3] = 0.0;
    return out;
  }
  C4 out = GetScattering(a, s, multiple, r, mu, mu_s, nu, ground);
  out[3] = 0.0;
  return out;
}

// ---------------------------------------------------------------------------
// Irradiance LUT parameterisation.
// ---------------------------------------------------------------------------

// functions.glsl:1524-1533
void GetIrradianceTextureUvFromRMuS(const Atmo& a, const Sizes& s, double r,
                                    double mu_s, double* u, double* v) {
  double x_r = (r - a.bottom_radius) / (a.top_radius - a.bottom_radius);
  double x_mu_s = mu_s * 0.5 + 0.5;
  *u = GetTextureCoordFromUnitRange(x_mu_s, s.irradiance_width);
  *v = GetTextureCoordFromUnitRange(x_r, s.irradiance_height);
}

// functions.glsl:1539-1548
void GetRMuSFromIrradianceTextureUv(const Atmo& a, const Sizes& s, double u,
                                    double v, double* r, double* mu_s) {
  double x_mu_s = GetUnitRangeFromTextureCoord(u, s.irradiance_width);
  double x_r = GetUnitRangeFromTextureCoord(v, s.irradiance_height);
  *r = a.bottom_radius + x_r * (a.top_radius - a.bottom_radius);
  *mu_s = ClampCosine(2.0 * x_mu_s - 1.0);
}

// functions.glsl:1595-1601
C4 GetIrradiance(const Atmo& a, const Sizes& s, const double* e_lut, double r,
                 double mu_s) {
  double u, v;
  GetIrradianceTextureUvFromRMuS(a, s, r, mu_s, &u, &v);
  C4 e = Texture2D(e_lut, s.irradiance_width, s.irradiance_height, u, v);
  e[3] = 0.0;
  return e;
}

// ---------------------------------------------------------------------------
// Multiple scattering.
// ---------------------------------------------------------------------------

// functions.glsl:1163-1260: midpoint rule over the sphere, 16 x 32 samples.
C4 ComputeScatteringDensity(const Atmo& a, const Sizes& s, const double* t_lut,
                            const double* single_rayleigh,
                            const double* single_mie, const double* multiple,
                            const double* irradiance, double r, double mu,
                            double mu_s, double nu, int order) {
  const V3 zenith{0.0, 0.0, 1.0};
  V3 omega{std::sqrt(1.0 - mu * mu), 0.0, mu};
  double sun_x = omega.x == 0.0 ? 0.0 : (nu - mu * mu_s) / omega.x;
  double sun_y = std::sqrt(std::max(1.0 - sun_x * sun_x - mu_s * mu_s, 0.0));
  V3 omega_s{sun_x, sun_y, mu_s};

  const int kSamples = 16;
  const double dphi = kPi / kSamples;
  const double dtheta = kPi / kSamples;
  // The density of the medium at the scattering point does not depend on the
  // incident direction; the reference recomputes it per sample with the same
  // result.
  double rayleigh_density =
      GetProfileDensity(a.rayleigh_density, r - a.bottom_radius);
  double mie_density = GetProfileDensity(a.mie_density, r - a.bottom_radius);
  C4 rayleigh_mie;
  for (int l = 0; l < kSamples; ++l) {
    double theta = (l + 0.5) * dtheta;
    double cos_theta = std::cos(theta);
    double sin_theta = std::sin(theta);
    bool ground_ray = RayIntersectsGround(a, r, cos_theta);
    double distance_to_ground = 0.0;
    C4 transmittance_to_ground;
    C4 ground_albedo;
    if (ground_ray) {
      distance_to_ground = DistanceToBottomAtmosphereBoundary(a, r, cos_theta);
      transmittance_to_ground = GetTransmittance(a, s, t_lut, r, cos_theta,
                                                 distance_to_ground, true);
      ground_albedo = From3(a.ground_albedo);
    }
    for (int m = 0; m < 2 * kSamples; ++m) {
      double phi = (m + 0.5) * dphi;
      V3 omega_i{std::cos(phi) * sin_theta, std::sin(phi) * sin_theta,
                 cos_theta};
      double domega_i = dtheta * dphi * std::sin(theta);

      double nu1 = Dot(omega_s, omega_i);
      C4 incident = GetScatteringOfOrder(a, s, single_rayleigh, single_mie,
                                         multiple, r, omega_i.z, mu_s, nu1,
                                         ground_ray, order - 1);
      V3 ground_normal = Normalize(zenith * r + omega_i * distance_to_ground);
      C4 ground_irradiance = GetIrradiance(a, s, irradiance, a.bottom_radius,
                                           Dot(ground_normal, omega_s));
      incident = Add(incident, Scale(Mul(Mul(transmittance_to_ground,
                                             ground_albedo),
                                         ground_irradiance),
                                     1.0 / kPi));

      double nu2 = Dot(omega, omega_i);
      double pr = RayleighPhaseFunction(nu2);
      double pm = MiePhaseFunction(a.mie_phase_function_g, nu2);
      for (int c = 0; c < 3; ++c) {
        rayleigh_mie[c] += incident[c] *
                           (a.rayleigh_scattering[c] * rayleigh_density * pr +
                            a.mie_scattering[c] * mie_density * pm) *
                           domega_i;
      }
    }
  }
  return rayleigh_mie;
}

// functions.glsl:1285-1330: trapezoidal rule, 50 intervals.
C4 ComputeMultipleScattering(const Atmo& a, const Sizes& s,
                             const double* t_lut, const double* density,
                             double r, double mu, double mu_s, double nu,
                             bool ground) {
  const int kIntervals = 50;
  double dx = DistanceToNearestAtmosphereBoundary(a, r, mu, ground) / kIntervals;
  C4 sum;
  for (int i = 0; i <= kIntervals; ++i) {
    double d_i = i * dx;
    double r_i =
        ClampRadius(a, std::sqrt(d_i * d_i + 2.0 * r * mu * d_i + r * r));
    double mu_i = ClampCosine((r * mu + d_i) / r_i);
    double mu_s_i = ClampCosine((r * mu_s + d_i * nu) / r_i);
    C4 v = Scale(Mul(GetScattering(a, s, density, r_i, mu_i, mu_s_i, nu, ground),
                     GetTransmittance(a, s, t_lut, r, mu, d_i, ground)),
                 dx);
    double w = (i == 0 || i == kIntervals) ? 0.5 : 1.0;
    sum = Add(sum, Scale(v, w));
  }
  sum[3] = 0.0;
  return sum;
}

// functions.glsl:1443-1461
C4 ComputeDirectIrradiance(const Atmo& a, const Sizes& s, const double* t_lut,
                           double r, double mu_s) {
  double alpha_s = a.sun_angular_radius;
  double average_cosine =
      mu_s < -alpha_s ? 0.0
                      : (mu_s > alpha_s ? mu_s
                                        : (mu_s + alpha_s) * (mu_s + alpha_s) /
                                              (4.0 * alpha_s));
  C4 t = GetTransmittanceToTopAtmosphereBoundary(a, s, t_lut, r, mu_s);
  C4 e;
  for (int c = 0; c < 3; ++c) e[c] = a.solar_irradiance[c] * t[c] * average_cosine;
  return e;
}

// functions.glsl:1477-1511: upper hemisphere, 16 x 64 midpoint samples.
C4 ComputeIndirectIrradiance(const Atmo& a, const Sizes& s,
                             const double* single_rayleigh,
                             const double* single_mie, const double* multiple,
                             double r, double mu_s, int order) {
  const int kSamples = 32;
  const double dphi = kPi / kSamples;
  const double dtheta = kPi / kSamples;
  C4 result;
  V3 omega_s{std::sqrt(1.0 - mu_s * mu_s), 0.0, mu_s};
  for (int j = 0; j < kSamples / 2; ++j) {
    double theta = (j + 0.5) * dtheta;
    for (int i = 0; i < 2 * kSamples; ++i) {
      double phi = (i + 0.5) * dphi;
      V3 omega{std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta),
               std::cos(theta)};
      double domega = dtheta * dphi * std::sin(theta);
      double nu = Dot(omega, omega_s);
      C4 v = GetScatteringOfOrder(a, s, single_rayleigh, single_mie, multiple,
                                  r, omega.z, mu_s, nu, false, order);
      result = Add(result, Scale(v, omega.z * domega));
    }
  }
  result[3] = 0.0;
  return result;
}

// ---------------------------------------------------------------------------
// Rendering lookups.
// ---------------------------------------------------------------------------

// functions.glsl:1633-1645
C4 GetExtrapolatedSingleMieScattering(const Atmo& a, const C4& scattering) {
  C4 out;
  if (scattering[0] <= 0.0) return out;
  for (int c = 0; c < 3; ++c) {
    out[c] = scattering[c] * scattering[3] / scattering[0] *
             (a.rayleigh_scattering[0] / a.mie_scattering[0]) *
             (a.mie_scattering[c] / a.rayleigh_scattering[c]);
  }
  return out;
}

// functions.glsl:1658-1690
C4 GetCombinedScattering(const Atmo& a, const Sizes& s, const double* scat,
                         const double* single_mie, bool combined, double r,
                         double mu, double mu_s, double nu, bool ground,
                         C4* single_mie_out) {
  Uvwz t = GetScatteringTextureUvwzFromRMuMuSNu(a, s, r, mu, mu_s, nu, ground);
  double tex_coord_x = t.u_nu * (s.scattering_nu - 1);
  double tex_x = std::floor(tex_coord_x);
  double lerp = tex_coord_x - tex_x;
  double u0 = (tex_x + t.u_mu_s) / s.scattering_nu;
  double u1 = (tex_x + 1.0 + t.u_mu_s) / s.scattering_nu;
  int w = ScatteringWidth(s);
  auto sample = [&](const double* lut) {
    C4 s0 = Texture3D(lut, w, s.scattering_mu, s.scattering_r, u0, t.u_mu, t.u_r);
    C4 s1 = Texture3D(lut, w, s.scattering_mu, s.scattering_r, u1, t.u_mu, t.u_r);
    return Add(Scale(s0, 1.0 - lerp), Scale(s1, lerp));
  };
  C4 scattering = sample(scat);
  if (combined) {
    *single_mie_out = GetExtrapolatedSingleMieScattering(a, scattering);
  } else {
    *single_mie_out = sample(single_mie);
    (*single_mie_out)[3] = 0.0;
  }
  return scattering;
}

inline C4 Keep3(C4 c) {
  c[3] = 0.0;
  return c;
}

// functions.glsl:1705-1769
C4 GetSkyRadiance(const Atmo& a, const Sizes& s, const double* t_lut,
                  const double* scat, const double* single_mie, bool combined,
                  V3 camera, V3 view_ray, double shadow_length,
                  V3 sun_direction, C4* transmittance) {
  double r = Length(camera);
  double rmu = Dot(camera, view_ray);
  double distance_to_top =
      -rmu - std::sqrt(rmu * rmu - r * r + a.top_radius * a.top_radius);
  if (distance_to_top > 0.0) {
    camera = camera + view_ray * distance_to_top;
    r = a.top_radius;
    rmu += distance_to_top;
  } else if (r > a.top_radius) {
    *transmittance = Splat3(1.0);
    return C4();
  }
  double mu = rmu / r;
  double mu_s = Dot(camera, sun_direction) / r;
  double nu = Dot(view_ray, sun_direction);
  bool ground = RayIntersectsGround(a, r, mu);
  *transmittance =
      ground ? C4() : GetTransmittanceToTopAtmosphereBoundary(a, s, t_lut, r, mu);
  C4 single_mie_scattering, scattering;
  if (shadow_length == 0.0) {
    scattering = GetCombinedScattering(a, s, scat, single_mie, combined, r, mu,
                                       mu_s, nu, ground, &single_mie_scattering);
  } else {
    double d = shadow_length;
    double r_p = ClampRadius(a, std::sqrt(d * d + 2.0 * r * mu * d + r * r));
    double mu_p = (r * mu + d) / r_p;
    double mu_s_p = (r * mu_s + d * nu) / r_p;
    scattering = GetCombinedScattering(a, s, scat, single_mie, combined, r_p,
                                       mu_p, mu_s_p, nu, ground,
                                       &single_mie_scattering);
    C4 shadow_t = GetTransmittance(a, s, t_lut, r, mu, shadow_length, ground);
    scattering = Mul(scattering, shadow_t);
    single_mie_scattering = Mul(single_mie_scattering, shadow_t);
  }
  C4 out = Add(Scale(scattering, RayleighPhaseFunction(nu)),
               Scale(single_mie_scattering,
                     MiePhaseFunction(a.mie_phase_function_g, nu)));
  return Keep3(out);
}

// functions.glsl:1787-1863
C4 GetSkyRadianceToPoint(const Atmo& a, const Sizes& s, const double* t_lut,
                         const double* scat, const double* single_mie,
                         bool combined, V3 camera, V3 point,
                         double shadow_length, V3 sun_direction,
                         C4* transmittance) {
  V3 view_ray = Normalize(point - camera);
  double r = Length(camera);
  double rmu = Dot(camera, view_ray);
  double distance_to_top =
      -rmu - std::sqrt(rmu * rmu - r * r + a.top_radius * a.top_radius);
  if (distance_to_top > 0.0) {
    camera = camera + view_ray * distance_to_top;
    r = a.top_radius;
    rmu += distance_to_top;
  }
  double mu = rmu / r;
  double mu_s = Dot(camera, sun_direction) / r;
  double nu = Dot(view_ray, sun_direction);
  double d = Length(point - camera);
  bool ground = RayIntersectsGround(a, r, mu);
  *transmittance = GetTransmittance(a, s, t_lut, r, mu, d, ground);

  C4 single_mie_scattering;
  C4 scattering = GetCombinedScattering(a, s, scat, single_mie, combined, r, mu,
                                        mu_s, nu, ground,
                                        &single_mie_scattering);
  d = std::max(d - shadow_length, 0.0);
  double r_p = ClampRadius(a, std::sqrt(d * d + 2.0 * r * mu * d + r * r));
  double mu_p = (r * mu + d) / r_p;
  double mu_s_p = (r * mu_s + d * nu) / r_p;
  C4 single_mie_scattering_p;
  C4 scattering_p = GetCombinedScattering(a, s, scat, single_mie, combined, r_p,
                                          mu_p, mu_s_p, nu, ground,
                                          &single_mie_scattering_p);
  C4 shadow_t = *transmittance;
  if (shadow_length > 0.0) {
    shadow_t = GetTransmittance(a, s, t_lut, r, mu, d, ground);
  }
  scattering = Keep3(Sub(scattering, Mul(Keep3(shadow_t), scattering_p)));
  single_mie_scattering =
      Sub(single_mie_scattering, Mul(Keep3(shadow_t), single_mie_scattering_p));
  if (combined) {
    // vec4(scattering, single_mie_scattering.r), functions.glsl:1852-1855.
    C4 packed = Keep3(scattering);
    packed[3] = single_mie_scattering[0];
    single_mie_scattering = GetExtrapolatedSingleMieScattering(a, packed);
  }
  double fade = SmoothStep(0.0, 0.01, mu_s);
  single_mie_scattering = Scale(single_mie_scattering, fade);
  C4 out = Add(Scale(Keep3(scattering), RayleighPhaseFunction(nu)),
               Scale(Keep3(single_mie_scattering),
                     MiePhaseFunction(a.mie_phase_function_g, nu)));
  return Keep3(out);
}

// functions.glsl:1878-1896
C4 GetSunAndSkyIrradiance(const Atmo& a, const Sizes& s, const double* t_lut,
                          const double* e_lut, V3 point, V3 normal,
                          V3 sun_direction, C4* sky_irradiance) {
  double r = Length(point);
  double mu_s = Dot(point, sun_direction) / r;
  *sky_irradiance = Scale(GetIrradiance(a, s, e_lut, r, mu_s),
                          (1.0 + Dot(normal, point) / r) * 0.5);
  C4 sun = Scale(Mul(From3(a.solar_irradiance),
                     GetTransmittanceToSun(a, s, t_lut, r, mu_s)),
                 std::max(Dot(normal, sun_direction), 0.0));
  return Keep3(sun);
}

// ---------------------------------------------------------------------------
// Whole-LUT passes.
// ---------------------------------------------------------------------------

inline size_t ScatteringTexels(const Sizes& s) {
  return static_cast<size_t>(ScatteringWidth(s)) * s.scattering_mu *
         s.scattering_r;
}

void Store(double* lut, size_t texel, const C4& v) {
  double* p = lut + texel * 4;
  p[0] = v[0];
  p[1] = v[1];
  p[2] = v[2];
  p[3] = v[3];
}

void TexelXYZ(const Sizes& s, size_t idx, int* x, int* y, int* z) {
  int w = ScatteringWidth(s);
  *x = static_cast<int>(idx % w);
  size_t rest = idx / w;
  *y = static_cast<int>(rest % s.scattering_mu);
  *z = static_cast<int>(rest / s.scattering_mu);
}

// atmosphere/model.cc:1048-1215 with the GL programs of model.cc:125-208.
// luminance_from_radiance is applied as in the shaders; 'blend' adds into the
// outputs (glBlendFunc(ONE, ONE)) instead of overwriting them.
// 'after_order', if set, is called once the outputs hold orders 1..n, for
// n = 1..orders.
void Precompute(const Atmo& a, const Sizes& s, const double* L, bool blend,
                unsigned orders, bool combined, double* transmittance,
                double* scattering, double* single_mie, double* irradiance,
                const std::function<void(unsigned)>& after_order = nullptr) {
  const int tw = s.transmittance_width, th = s.transmittance_height;
  const int ew = s.irradiance_width, eh = s.irradiance_height;
  const size_t n_s = ScatteringTexels(s);
  const size_t n_e = static_cast<size_t>(ew) * eh;
  std::vector<double> delta_irradiance(n_e * 4, 0.0);
  std::vector<double> delta_rayleigh(n_s * 4, 0.0);  // also delta_multiple
  std::vector<double> delta_mie(n_s * 4, 0.0);
  std::vector<double> delta_density(n_s * 4, 0.0);
  auto lum = [L](const C4& v) {
    C4 o;
    for (int r = 0; r < 3; ++r)
      o[r] = L[3 * r] * v[0] + L[3 * r + 1] * v[1] + L[3 * r + 2] * v[2];
    return o;
  };

  ParallelFor(th, [&](int j) {
    for (int i = 0; i < tw; ++i)
      Store(transmittance, static_cast<size_t>(j) * tw + i,
            ComputeTransmittanceToTopAtmosphereBoundaryTexture(a, s, i, j));
  });

  ParallelFor(eh, [&](int j) {
    for (int i = 0; i < ew; ++i) {
      double r, mu_s;
      GetRMuSFromIrradianceTextureUv(a, s, (i + 0.5) / ew, (j + 0.5) / eh, &r,
                                     &mu_s);
      size_t t = static_cast<size_t>(j) * ew + i;
      Store(delta_irradiance.data(), t,
            ComputeDirectIrradiance(a, s, transmittance, r, mu_s));
      if (!blend) Store(irradiance, t, C4());
    }
  });

  const int sw = ScatteringWidth(s);
  ParallelFor(s.scattering_r, [&](int z) {
    for (int y = 0; y < s.scattering_mu; ++y)
      for (int x = 0; x < sw; ++x) {
        double r, mu, mu_s, nu;
        bool ground;
        GetRMuMuSNuFromScatteringTextureFragCoord(a, s, x, y, z, &r, &mu,
                                                  &mu_s, &nu, &ground);
        C4 ray, mie;
        ComputeSingleScattering(a, s, transmittance, r, mu, mu_s, nu, ground,
                                &ray, &mie);
        size_t t = (static_cast<size_t>(z) * s.scattering_mu + y) * sw + x;
        Store(delta_rayleigh.data(), t, ray);
        Store(delta_mie.data(), t, mie);
        C4 lr = lum(ray), lm = lum(mie);
        C4 packed = lr;
        packed[3] = lm[0];
        if (blend) packed = Add(packed, Fetch(scattering, t));
        Store(scattering, t, packed);
        if (!combined) {
          if (blend) lm = Add(lm, Fetch(single_mie, t));
          lm[3] = 0.0;
          Store(single_mie, t, lm);
        }
      }
  });
  if (after_order) after_order(1);

  for (unsigned order = 2; order <= orders; ++order) {
    ParallelFor(s.scattering_r, [&](int z) {
      for (int y = 0; y < s.scattering_mu; ++y)
        for (int x = 0; x < sw; ++x) {
          double r, mu, mu_s, nu;
          bool ground;
          GetRMuMuSNuFromScatteringTextureFragCoord(a, s, x, y, z, &r, &mu,
                                                    &mu_s, &nu, &ground);
          size_t t = (static_cast<size_t>(z) * s.scattering_mu + y) * sw + x;
          Store(delta_density.data(), t,
                ComputeScatteringDensity(
                    a, s, transmittance, delta_rayleigh.data(),
                    delta_mie.data(), delta_rayleigh.data(),
                    delta_irradiance.data(), r, mu, mu_s, nu, order));
        }
    });
    ParallelFor(eh, [&](int j) {
      for (int i = 0; i < ew; ++i) {
        double r, mu_s;
        GetRMuSFromIrradianceTextureUv(a, s, (i + 0.5) / ew, (j + 0.5) / eh,
                                       &r, &mu_s);
        size_t t = static_cast<size_t>(j) * ew + i;
        C4 de = ComputeIndirectIrradiance(a, s, delta_rayleigh.data(),
                                          delta_mie.data(),
                                          delta_rayleigh.data(), r, mu_s,
                                          order - 1);
        // Written after all texels are computed (below): the pass reads the
        // scattering LUTs only, so the order does not matter here.
        Store(delta_irradiance.data(), t, de);
        Store(irradiance, t, Add(Fetch(irradiance, t), lum(de)));
      }
    });
    ParallelFor(s.scattering_r, [&](int z) {
      for (int y = 0; y < s.scattering_mu; ++y)
        for (int x = 0; x < sw; ++x) {
          double r, mu, mu_s, nu;
          bool ground;
          GetRMuMuSNuFromScatteringTextureFragCoord(a, s, x, y, z, &r, &mu,
                                                    &mu_s, &nu, &ground);
          size_t t = (static_cast<size_t>(z) * s.scattering_mu + y) * sw + x;
          C4 dms = ComputeMultipleScattering(a, s, transmittance,
                                             delta_density.data(), r, mu, mu_s,
                                             nu, ground);
          Store(delta_rayleigh.data(), t, dms);
          C4 add = Scale(lum(dms), 1.0 / RayleighPhaseFunction(nu));
          add[3] = 0.0;
          Store(scattering, t, Add(Fetch(scattering, t), add));
        }
    });
    if (after_order) after_order(order);
  }
}

// ---------------------------------------------------------------------------
// Model-level host code (atmosphere/model.cc).
// ---------------------------------------------------------------------------

constexpr double kLambdaR = 680.0;
constexpr double kLambdaG = 550.0;
constexpr double kLambdaB = 440.0;
constexpr int kLambdaMin = 360;
constexpr int kLambdaMax = 830;

// model.cc:535-552
double Interpolate(int n, const double* wl, const double* f, double lambda) {
  if (lambda < wl[0]) return f[0];
  for (int i = 0; i < n - 1; ++i) {
    if (lambda < wl[i + 1]) {
      double u = (lambda - wl[i]) / (wl[i + 1] - wl[i]);
      return f[i] * (1.0 - u) + f[i + 1] * u;
    }
  }
  return f[n - 1];
}

// model.cc:521-533 (column 1, 2, 3 = x_bar, y_bar, z_bar)
double CieValue(double lambda, int column) {
  if (lambda <= kLambdaMin || lambda >= kLambdaMax) return 0.0;
  double u = (lambda - kLambdaMin) / atmo_cie::kCieLambdaStep;
  int row = static_cast<int>(std::floor(u));
  u -= row;
  const double* table = column == 1 ? atmo_cie::kCieXBar
                        : column == 2 ? atmo_cie::kCieYBar
                                      : atmo_cie::kCieZBar;
  return table[row] * (1.0 - u) + table[row + 1] * u;
}

// model.cc:562-595
void LuminanceFactors(int n, const double* wl, const double* solar,
                      double lambda_power, double* k) {
  k[0] = k[1] = k[2] = 0.0;
  const double lambdas[3] = {kLambdaR, kLambdaG, kLambdaB};
  double solar_rgb[3];
  for (int c = 0; c < 3; ++c) solar_rgb[c] = Interpolate(n, wl, solar, lambdas[c]);
  const double* m = atmo_cie::kXyzToSrgb;
  for (int lambda = kLambdaMin; lambda < kLambdaMax; ++lambda) {
    double x = CieValue(lambda, 1), y = CieValue(lambda, 2),
           z = CieValue(lambda, 3);
    double irr = Interpolate(n, wl, solar, lambda);
    for (int c = 0; c < 3; ++c) {
      double bar = m[3 * c] * x + m[3 * c + 1] * y + m[3 * c + 2] * z;
      k[c] += bar * irr / solar_rgb[c] * std::pow(lambda / lambdas[c], lambda_power);
    }
  }
  for (int c = 0; c < 3; ++c) k[c] *= atmo_cie::kMaxLuminousEfficacy;
}

// model.cc:644-666: lengths scaled by the length unit, profiles padded with
// null layers at the front up to two layers.
void LowerProfile(int n, const or_layer* layers, double unit, or_layer* out) {
  or_layer padded[2] = {};
  int start = 2 - std::min(n, 2);
  for (int i = 0; i < std::min(n, 2); ++i) padded[start + i] = layers[i];
  for (int i = 0; i < 2; ++i) {
    out[i].width = padded[i].width / unit;
    out[i].exp_term = padded[i].exp_term;
    out[i].exp_scale = padded[i].exp_scale * unit;
    out[i].linear_term = padded[i].linear_term * unit;
    out[i].constant_term = padded[i].constant_term;
  }
}

void Lower(const or_model_params& p, const double* lambdas, Atmo* a) {
  const int n = p.num_wavelengths;
  const double unit = p.length_unit_in_meters;
  for (int c = 0; c < 3; ++c) {
    a->solar_irradiance[c] = Interpolate(n, p.wavelengths, p.solar_irradiance, lambdas[c]);
    a->rayleigh_scattering[c] =
        Interpolate(n, p.wavelengths, p.rayleigh_scattering, lambdas[c]) * unit;
    a->mie_scattering[c] =
        Interpolate(n, p.wavelengths, p.mie_scattering, lambdas[c]) * unit;
    a->mie_extinction[c] =
        Interpolate(n, p.wavelengths, p.mie_extinction, lambdas[c]) * unit;
    a->absorption_extinction[c] =
        Interpolate(n, p.wavelengths, p.absorption_extinction, lambdas[c]) * unit;
    a->ground_albedo[c] = Interpolate(n, p.wavelengths, p.ground_albedo, lambdas[c]);
  }
  a->sun_angular_radius = p.sun_angular_radius;
  a->bottom_radius = p.bottom_radius / unit;
  a->top_radius = p.top_radius / unit;
  LowerProfile(p.num_rayleigh_layers, p.rayleigh_density, unit, a->rayleigh_density);
  LowerProfile(p.num_mie_layers, p.mie_density, unit, a->mie_density);
  LowerProfile(p.num_absorption_layers, p.absorption_density, unit,
               a->absorption_density);
  a->mie_phase_function_g = p.mie_phase_function_g;
  a->mu_s_min = std::cos(p.max_sun_zenith_angle);
}

int NumIterations(const or_model_params& p) {
  if (p.num_precomputed_wavelengths <= 3) return 1;
  return static_cast<int>((p.num_precomputed_wavelengths + 2) / 3);
}

// model.cc:907-943
void Iteration(const or_model_params& p, int i, double* lambdas, double* L) {
  if (p.num_precomputed_wavelengths <= 3) {
    lambdas[0] = kLambdaR;
    lambdas[1] = kLambdaG;
    lambdas[2] = kLambdaB;
    for (int k = 0; k < 9; ++k) L[k] = (k % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  int iters = NumIterations(p);
  double dlambda = (kLambdaMax - kLambdaMin) / (3.0 * iters);
  for (int k = 0; k < 3; ++k) lambdas[k] = kLambdaMin + (3 * i + k + 0.5) * dlambda;
  const double* m = atmo_cie::kXyzToSrgb;
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col) {
      double l = lambdas[col];
      // The reference rounds each coefficient to float (model.cc:934): the
      // GL programs receive luminance_from_radiance as a float mat3.
      L[3 * row + col] = static_cast<double>(static_cast<float>(
          (m[3 * row] * CieValue(l, 1) + m[3 * row + 1] * CieValue(l, 2) +
           m[3 * row + 2] * CieValue(l, 3)) *
          dlambda));
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------

extern "C" {

void or_set_threads(int n) { g_threads = n; }
int or_get_threads(void) { return ThreadCount(); }

double or_distance_to_top(const or_atmosphere* a, double r, double mu) {
  return DistanceToTopAtmosphereBoundary(*a, r, mu);
}
double or_distance_to_bottom(const or_atmosphere* a, double r, double mu) {
  return DistanceToBottomAtmosphereBoundary(*a, r, mu);
}
int or_ray_intersects_ground(const or_atmosphere* a, double r, double mu) {
  return RayIntersectsGround(*a, r, mu) ? 1 : 0;
}
double or_distance_to_nearest(const or_atmosphere* a, double r, double mu,
                              int ground) {
  return DistanceToNearestAtmosphereBoundary(*a, r, mu, ground != 0);
}
double or_profile_density(const or_layer* two_layers, double altitude) {
  return GetProfileDensity(two_layers, altitude);
}
double or_optical_length(const or_atmosphere* a, const or_layer* two_layers,
                         double r, double mu) {
  return ComputeOpticalLengthToTopAtmosphereBoundary(*a, two_layers, r, mu);
}
void or_transmittance_to_top(const or_atmosphere* a, double r, double mu,
                             double* out3) {
  Put3(ComputeTransmittanceToTopAtmosphereBoundary(*a, r, mu), out3);
}
double or_tex_coord_from_unit_range(double x, int n) {
  return GetTextureCoordFromUnitRange(x, n);
}
double or_unit_range_from_tex_coord(double u, int n) {
  return GetUnitRangeFromTextureCoord(u, n);
}
void or_transmittance_uv_from_rmu(const or_atmosphere* a, const or_sizes* s,
                                  double r, double mu, double* out_uv) {
  GetTransmittanceTextureUvFromRMu(*a, *s, r, mu, &out_uv[0], &out_uv[1]);
}
void or_rmu_from_transmittance_uv(const or_atmosphere* a, const or_sizes* s,
                                  double u, double v, double* r, double* mu) {
  GetRMuFromTransmittanceTextureUv(*a, *s, u, v, r, mu);
}
void or_get_transmittance_to_top(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, double r, double mu,
                                 double* out3) {
  Put3(GetTransmittanceToTopAtmosphereBoundary(*a, *s, t_lut, r, mu), out3);
}
void or_get_transmittance(const or_atmosphere* a, const or_sizes* s,
                          const double* t_lut, double r, double mu, double d,
                          int ground, double* out3) {
  Put3(GetTransmittance(*a, *s, t_lut, r, mu, d, ground != 0), out3);
}
void or_get_transmittance_to_sun(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, double r, double mu_s,
                                 double* out3) {
  Put3(GetTransmittanceToSun(*a, *s, t_lut, r, mu_s), out3);
}
void or_single_scattering_integrand(const or_atmosphere* a, const or_sizes* s,
                                    const double* t_lut, double r, double mu,
                                    double mu_s, double nu, double d,
                                    int ground, double* rayleigh3,
                                    double* mie3) {
  C4 ray, mie;
  ComputeSingleScatteringIntegrand(*a, *s, t_lut, r, mu, mu_s, nu, d,
                                   ground != 0, &ray, &mie);
  Put3(ray, rayleigh3);
  Put3(mie, mie3);
}
void or_single_scattering(const or_atmosphere* a, const or_sizes* s,
                          const double* t_lut, double r, double mu,
                          double mu_s, double nu, int ground,
                          double* rayleigh3, double* mie3) {
  C4 ray, mie;
  ComputeSingleScattering(*a, *s, t_lut, r, mu, mu_s, nu, ground != 0, &ray,
                          &mie);
  Put3(ray, rayleigh3);
  Put3(mie, mie3);
}
double or_rayleigh_phase(double nu) { return RayleighPhaseFunction(nu); }
double or_mie_phase(double g, double nu) { return MiePhaseFunction(g, nu); }
void or_scattering_uvwz_from_rmumusnu(const or_atmosphere* a,
                                      const or_sizes* s, double r, double mu,
                                      double mu_s, double nu, int ground,
                                      double* out) {
  Uvwz t = GetScatteringTextureUvwzFromRMuMuSNu(*a, *s, r, mu, mu_s, nu,
                                                ground != 0);
  out[0] = t.u_nu;
  out[1] = t.u_mu_s;
  out[2] = t.u_mu;
  out[3] = t.u_r;
}
void or_rmumusnu_from_scattering_uvwz(const or_atmosphere* a,
                                      const or_sizes* s, const double* uvwz,
                                      double* out, int* ground) {
  Uvwz t{uvwz[0], uvwz[1], uvwz[2], uvwz[3]};
  bool g;
  GetRMuMuSNuFromScatteringTextureUvwz(*a, *s, t, &out[0], &out[1], &out[2],
                                       &out[3], &g);
  *ground = g ? 1 : 0;
}
void or_rmumusnu_from_scattering_texel(const or_atmosphere* a,
                                       const or_sizes* s, int x, int y, int z,
                                       double* out, int* ground) {
  bool g;
  GetRMuMuSNuFromScatteringTextureFragCoord(*a, *s, x, y, z, &out[0], &out[1],
                                            &out[2], &out[3], &g);
  *ground = g ? 1 : 0;
}
void or_get_scattering(const or_atmosphere* a, const or_sizes* s,
                       const double* lut, double r, double mu, double mu_s,
                       double nu, int ground, double* out4) {
  C4 v = GetScattering(*a, *s, lut, r, mu, mu_s, nu, ground != 0);
  for (int c = 0; c < 4; ++c) out4[c] = v[c];
}
void or_scattering_density(const or_atmosphere* a, const or_sizes* s,
                           const double* t_lut, const double* single_rayleigh,
                           const double* single_mie,
                           const double* multiple_scattering,
                           const double* irradiance, double r, double mu,
                           double mu_s, double nu, int order, double* out3) {
  Put3(ComputeScatteringDensity(*a, *s, t_lut, single_rayleigh, single_mie,
                                multiple_scattering, irradiance, r, mu, mu_s,
                                nu, order),
       out3);
}
void or_multiple_scattering(const or_atmosphere* a, const or_sizes* s,
                            const double* t_lut, const double* density_lut,
                            double r, double mu, double mu_s, double nu,
                            int ground, double* out3) {
  Put3(ComputeMultipleScattering(*a, *s, t_lut, density_lut, r, mu, mu_s, nu,
                                 ground != 0),
       out3);
}
void or_direct_irradiance(const or_atmosphere* a, const or_sizes* s,
                          const double* t_lut, double r, double mu_s,
                          double* out3) {
  Put3(ComputeDirectIrradiance(*a, *s, t_lut, r, mu_s), out3);
}
void or_indirect_irradiance(const or_atmosphere* a, const or_sizes* s,
                            const double* single_rayleigh,
                            const double* single_mie,
                            const double* multiple_scattering, double r,
                            double mu_s, int order, double* out3) {
  Put3(ComputeIndirectIrradiance(*a, *s, single_rayleigh, single_mie,
                                 multiple_scattering, r, mu_s, order),
       out3);
}
void or_irradiance_uv_from_rmus(const or_atmosphere* a, const or_sizes* s,
                                double r, double mu_s, double* out_uv) {
  GetIrradianceTextureUvFromRMuS(*a, *s, r, mu_s, &out_uv[0], &out_uv[1]);
}
void or_rmus_from_irradiance_uv(const or_atmosphere* a, const or_sizes* s,
                                double u, double v, double* r, double* mu_s) {
  GetRMuSFromIrradianceTextureUv(*a, *s, u, v, r, mu_s);
}
void or_get_irradiance(const or_atmosphere* a, const or_sizes* s,
                       const double* e_lut, double r, double mu_s,
                       double* out3) {
  Put3(GetIrradiance(*a, *s, e_lut, r, mu_s), out3);
}

void or_pass_transmittance(const or_atmosphere* a, const or_sizes* s,
                           double* t_out) {
  const int tw = s->transmittance_width;
  ParallelFor(s->transmittance_height, [&](int j) {
    for (int i = 0; i < tw; ++i)
      Store(t_out, static_cast<size_t>(j) * tw + i,
            ComputeTransmittanceToTopAtmosphereBoundaryTexture(*a, *s, i, j));
  });
}

void or_pass_direct_irradiance(const or_atmosphere* a, const or_sizes* s,
                               const double* t_lut, double* out) {
  const int ew = s->irradiance_width, eh = s->irradiance_height;
  ParallelFor(eh, [&](int j) {
    for (int i = 0; i < ew; ++i) {
      double r, mu_s;
      GetRMuSFromIrradianceTextureUv(*a, *s, (i + 0.5) / ew, (j + 0.5) / eh,
                                     &r, &mu_s);
      Store(out, static_cast<size_t>(j) * ew + i,
            ComputeDirectIrradiance(*a, *s, t_lut, r, mu_s));
    }
  });
}

void or_pass_single_scattering(const or_atmosphere* a, const or_sizes* s,
                               const double* t_lut, double* ray_out,
                               double* mie_out) {
  const int sw = ScatteringWidth(*s);
  ParallelFor(s->scattering_r, [&](int z) {
    for (int y = 0; y < s->scattering_mu; ++y)
      for (int x = 0; x < sw; ++x) {
        double r, mu, mu_s, nu;
        bool ground;
        GetRMuMuSNuFromScatteringTextureFragCoord(*a, *s, x, y, z, &r, &mu,
                                                  &mu_s, &nu, &ground);
        C4 ray, mie;
        ComputeSingleScattering(*a, *s, t_lut, r, mu, mu_s, nu, ground, &ray,
                                &mie);
        size_t t = (static_cast<size_t>(z) * s->scattering_mu + y) * sw + x;
        Store(ray_out, t, ray);
        Store(mie_out, t, mie);
      }
  });
}

void or_pass_scattering_density(const or_atmosphere* a, const or_sizes* s,
                                const double* t_lut, const double* dr,
                                const double* dm, const double* dms,
                                const double* de, int order, double* out) {
  const int sw = ScatteringWidth(*s);
  ParallelFor(s->scattering_r, [&](int z) {
    for (int y = 0; y < s->scattering_mu; ++y)
      for (int x = 0; x < sw; ++x) {
        double r, mu, mu_s, nu;
        bool ground;
        GetRMuMuSNuFromScatteringTextureFragCoord(*a, *s, x, y, z, &r, &mu,
                                                  &mu_s, &nu, &ground);
        size_t t = (static_cast<size_t>(z) * s->scattering_mu + y) * sw + x;
        Store(out, t,
              ComputeScatteringDensity(*a, *s, t_lut, dr, dm, dms, de, r, mu,
                                       mu_s, nu, order));
      }
  });
}

void or_pass_indirect_irradiance(const or_atmosphere* a, const or_sizes* s,
                                 const double* dr, const double* dm,
                                 const double* dms, int order, double* out) {
  const int ew = s->irradiance_width, eh = s->irradiance_height;
  ParallelFor(eh, [&](int j) {
    for (int i = 0; i < ew; ++i) {
      double r, mu_s;
      GetRMuSFromIrradianceTextureUv(*a, *s, (i + 0.5) / ew, (j + 0.5) / eh,
                                     &r, &mu_s);
      Store(out, static_cast<size_t>(j) * ew + i,
            ComputeIndirectIrradiance(*a, *s, dr, dm, dms, r, mu_s, order));
    }
  });
}

void or_pass_multiple_scattering(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, const double* dj,
                                 double* out, double* nu_out) {
  const int sw = ScatteringWidth(*s);
  ParallelFor(s->scattering_r, [&](int z) {
    for (int y = 0; y < s->scattering_mu; ++y)
      for (int x = 0; x < sw; ++x) {
        double r, mu, mu_s, nu;
        bool ground;
        GetRMuMuSNuFromScatteringTextureFragCoord(*a, *s, x, y, z, &r, &mu,
                                                  &mu_s, &nu, &ground);
        size_t t = (static_cast<size_t>(z) * s->scattering_mu + y) * sw + x;
        Store(out, t,
              ComputeMultipleScattering(*a, *s, t_lut, dj, r, mu, mu_s, nu,
                                        ground));
        if (nu_out != nullptr) nu_out[t] = nu;
      }
  });
}

void or_texels_single_scattering(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, int n, const int* idx,
                                 double* ray_out, double* mie_out) {
  ParallelFor(n, [&](int i) {
    int x, y, z;
    TexelXYZ(*s, static_cast<size_t>(idx[i]), &x, &y, &z);
    double r, mu, mu_s, nu;
    bool ground;
    GetRMuMuSNuFromScatteringTextureFragCoord(*a, *s, x, y, z, &r, &mu, &mu_s,
                                              &nu, &ground);
    C4 ray, mie;
    ComputeSingleScattering(*a, *s, t_lut, r, mu, mu_s, nu, ground, &ray,
                            &mie);
    Store(ray_out, i, ray);
    Store(mie_out, i, mie);
  });
}

void or_texels_scattering_density(const or_atmosphere* a, const or_sizes* s,
                                  const double* t_lut, const double* dr,
                                  const double* dm, const double* dms,
                                  const double* de, int order, int n,
                                  const int* idx, double* out) {
  ParallelFor(n, [&](int i) {
    int x, y, z;
    TexelXYZ(*s, static_cast<size_t>(idx[i]), &x, &y, &z);
    double r, mu, mu_s, nu;
    bool ground;
    GetRMuMuSNuFromScatteringTextureFragCoord(*a, *s, x, y, z, &r, &mu, &mu_s,
                                              &nu, &ground);
    Store(out, i,
          ComputeScatteringDensity(*a, *s, t_lut, dr, dm, dms, de, r, mu, mu_s,
                                   nu, order));
  });
}

void or_texels_multiple_scattering(const or_atmosphere* a, const or_sizes* s,
                                   const double* t_lut, const double* dj,
                                   int n, const int* idx, double* out) {
  ParallelFor(n, [&](int i) {
    int x, y, z;
    TexelXYZ(*s, static_cast<size_t>(idx[i]), &x, &y, &z);
    double r, mu, mu_s, nu;
    bool ground;
    GetRMuMuSNuFromScatteringTextureFragCoord(*a, *s, x, y, z, &r, &mu, &mu_s,
                                              &nu, &ground);
    Store(out, i,
          ComputeMultipleScattering(*a, *s, t_lut, dj, r, mu, mu_s, nu,
                                    ground));
  });
}

double or_interpolate(int n, const double* wavelengths, const double* values,
                      double lambda) {
  return Interpolate(n, wavelengths, values, lambda);
}
double or_cie_table_value(double lambda, int column) {
  return CieValue(lambda, column);
}
void or_luminance_factors(int n, const double* wavelengths,
                          const double* solar, double lambda_power,
                          double* k3) {
  LuminanceFactors(n, wavelengths, solar, lambda_power, k3);
}

// model.cc:1020-1040
void or_convert_spectrum_to_linear_srgb(int n, const double* wavelengths,
                                        const double* spectrum, double* rgb) {
  double x = 0.0, y = 0.0, z = 0.0;
  for (int lambda = kLambdaMin; lambda < kLambdaMax; ++lambda) {
    double v = Interpolate(n, wavelengths, spectrum, lambda);
    x += CieValue(lambda, 1) * v;
    y += CieValue(lambda, 2) * v;
    z += CieValue(lambda, 3) * v;
  }
  const double* m = atmo_cie::kXyzToSrgb;
  for (int c = 0; c < 3; ++c)
    rgb[c] = atmo_cie::kMaxLuminousEfficacy *
             (m[3 * c] * x + m[3 * c + 1] * y + m[3 * c + 2] * z);
}

void or_lower(const or_model_params* p, const double* lambdas3,
              or_atmosphere* out) {
  Lower(*p, lambdas3, out);
}
int or_num_iterations(const or_model_params* p) { return NumIterations(*p); }
void or_iteration(const or_model_params* p, int i, double* lambdas3,
                  double* matrix9) {
  Iteration(*p, i, lambdas3, matrix9);
}

void or_precompute(const or_atmosphere* a, const or_sizes* s,
                   const double* matrix9, int blend,
                   unsigned num_scattering_orders, int combined,
                   double* transmittance, double* scattering,
                   double* single_mie, double* irradiance) {
  Precompute(*a, *s, matrix9, blend != 0, num_scattering_orders, combined != 0,
             transmittance, scattering, single_mie, irradiance);
}

// model.cc:866-975
int or_model_init(const or_model_params* p, const or_sizes* s,
                  unsigned orders, double* transmittance, double* scattering,
                  double* single_mie, double* irradiance) {
  if (p == nullptr || s == nullptr || p->num_wavelengths < 1) return 1;
  bool combined = p->combine_scattering_textures != 0;
  int iters = NumIterations(*p);
  for (int i = 0; i < iters; ++i) {
    double lambdas[3], L[9];
    Iteration(*p, i, lambdas, L);
    Atmo a;
    Lower(*p, lambdas, &a);
    Precompute(a, *s, L, i > 0, orders, combined, transmittance, scattering,
               single_mie, irradiance);
  }
  if (p->num_precomputed_wavelengths > 3) {
    // model.cc:951-963: the transmittance is recomputed at 680/550/440 nm.
    const double rgb[3] = {kLambdaR, kLambdaG, kLambdaB};
    Atmo a;
    Lower(*p, rgb, &a);
    or_pass_transmittance(&a, s, transmittance);
  }
  return 0;
}

// Model::Init for every order count 1..orders in one pass: out_*[n - 1]
// receive the LUTs of Init(n). Each iteration is computed on its own
// (blend off) and the per-order states are summed over iterations, which is
// what blending the iterations into one LUT computes.
int or_model_init_orders(const or_model_params* p, const or_sizes* s,
                         unsigned orders, double* transmittance,
                         double* const* scattering, double* const* single_mie,
                         double* const* irradiance) {
  if (p == nullptr || s == nullptr || p->num_wavelengths < 1 || orders < 1) return 1;
  const bool combined = p->combine_scattering_textures != 0;
  const size_t n_s = ScatteringTexels(*s) * 4;
  const size_t n_e = static_cast<size_t>(s->irradiance_width) * s->irradiance_height * 4;
  for (unsigned n = 0; n < orders; ++n) {
    std::fill(scattering[n], scattering[n] + n_s, 0.0);
    if (!combined) std::fill(single_mie[n], single_mie[n] + n_s, 0.0);
    std::fill(irradiance[n], irradiance[n] + n_e, 0.0);
  }
  std::vector<double> S(n_s), M(n_s), E(n_e);
  const int iters = NumIterations(*p);
  for (int i = 0; i < iters; ++i) {
    double lambdas[3], L[9];
    Iteration(*p, i, lambdas, L);
    Atmo a;
    Lower(*p, lambdas, &a);
    Precompute(a, *s, L, false, orders, combined, transmittance, S.data(), M.data(),
               E.data(), [&](unsigned n) {
                 for (size_t k = 0; k < n_s; ++k) scattering[n - 1][k] += S[k];
                 if (!combined)
                   for (size_t k = 0; k < n_s; ++k) single_mie[n - 1][k] += M[k];
                 for (size_t k = 0; k < n_e; ++k) irradiance[n - 1][k] += E[k];
               });
  }
  if (p->num_precomputed_wavelengths > 3) {
    const double rgb[3] = {kLambdaR, kLambdaG, kLambdaB};
    Atmo a;
    Lower(*p, rgb, &a);
    or_pass_transmittance(&a, s, transmittance);
  }
  return 0;
}

void or_sky_radiance(const or_atmosphere* a, const or_sizes* s,
                     const double* t_lut, const double* scattering,
                     const double* single_mie, int combined, int n,
                     const double* camera, const double* view_ray,
                     const double* shadow_length, const double* sun_direction,
                     double* radiance_out, double* transmittance_out) {
  ParallelFor(n, [&](int i) {
    V3 c{camera[3 * i], camera[3 * i + 1], camera[3 * i + 2]};
    V3 v{view_ray[3 * i], view_ray[3 * i + 1], view_ray[3 * i + 2]};
    V3 sd{sun_direction[3 * i], sun_direction[3 * i + 1],
          sun_direction[3 * i + 2]};
    C4 t;
    C4 l = GetSkyRadiance(*a, *s, t_lut, scattering, single_mie, combined != 0,
                          c, v, shadow_length[i], sd, &t);
    Put3(l, radiance_out + 3 * i);
    Put3(t, transmittance_out + 3 * i);
  });
}

void or_sky_radiance_to_point(const or_atmosphere* a, const or_sizes* s,
                              const double* t_lut, const double* scattering,
                              const double* single_mie, int combined, int n,
                              const double* camera, const double* point,
                              const double* shadow_length,
                              const double* sun_direction,
                              double* radiance_out,
                              double* transmittance_out) {
  ParallelFor(n, [&](int i) {
    V3 c{camera[3 * i], camera[3 * i + 1], camera[3 * i + 2]};
    V3 p{point[3 * i], point[3 * i + 1], point[3 * i + 2]};
    V3 sd{sun_direction[3 * i], sun_direction[3 * i + 1],
          sun_direction[3 * i + 2]};
    C4 t;
    C4 l = GetSkyRadianceToPoint(*a, *s, t_lut, scattering, single_mie,
                                 combined != 0, c, p, shadow_length[i], sd, &t);
    Put3(l, radiance_out + 3 * i);
    Put3(t, transmittance_out + 3 * i);
  });
}

void or_sun_and_sky_irradiance(const or_atmosphere* a, const or_sizes* s,
                               const double* t_lut, const double* e_lut,
                               int n, const double* point,
                               const double* normal,
                               const double* sun_direction,
                               double* sun_out, double* sky_out) {
  ParallelFor(n, [&](int i) {
    V3 p{point[3 * i], point[3 * i + 1], point[3 * i + 2]};
    V3 nn{normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]};
    V3 sd{sun_direction[3 * i], sun_direction[3 * i + 1],
          sun_direction[3 * i + 2]};
    C4 sky;
    C4 sun = GetSunAndSkyIrradiance(*a, *s, t_lut, e_lut, p, nn, sd, &sky);
    Put3(sun, sun_out + 3 * i);
    Put3(sky, sky_out + 3 * i);
  });
}

}  // extern "C"
