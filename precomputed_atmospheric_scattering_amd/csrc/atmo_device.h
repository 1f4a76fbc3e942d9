// Device-side parameter block and the scalar mathematics shared by the
// precompute and lookup kernels (gfx950).
//
// Precision plan. Everything that sets up a texel or a ray-march sample (the
// LUT-coordinate mappings, ray/sphere intersections, ground flags, the
// transmittance LUT coordinates) runs in FP64: those quantities are shared by
// many inner-loop samples, so the cost is amortised, and they are where FP32
// cancellation shifts LUT coordinates near the horizon and near the ground.
// The inner quadrature loops (phase functions, LUT interpolation, sums) run in
// FP32 with FP32 accumulation.
#ifndef ATMO_DEVICE_H_
#define ATMO_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#define ATMO_HD __host__ __device__ inline

// LUT loads and stores through the global address space. The LUT pointers
// are read from the job table, so without this the compiler cannot tell
// they are global and emits flat accesses, which also wait on LDS traffic.
template <class T>
ATMO_HD T gload(const T* p, int64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(1))) T*)p)[i];
#else
  return p[i];
#endif
}
template <class T>
ATMO_HD void gstore(T* p, int64_t i, const T& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  ((__attribute__((address_space(1))) T*)p)[i] = v;
#else
  p[i] = v;
#endif
}
// Load at an unsigned 32-bit element index (LUTs hold at most 2^28 texels,
// so the byte offset fits 32 bits): with a uniform base the compiler uses
// the scalar-base + 32-bit vector-offset form, no 64-bit address per lane.
template <class T>
ATMO_HD T gload32(const T* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(1))) char* b =
      (const __attribute__((address_space(1))) char*)p;
  return *(const __attribute__((address_space(1))) T*)(b + i * static_cast<uint32_t>(sizeof(T)));
#else
  return p[i];
#endif
}

// Load at an unsigned 32-bit byte offset from a uniform base (the same
// scalar-base + vector-offset form, with the offset computed by the caller).
template <class T>
ATMO_HD T gload_at(const T* p, uint32_t byte_offset) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(1))) char* b =
      (const __attribute__((address_space(1))) char*)p;
  return *(const __attribute__((address_space(1))) T*)(b + byte_offset);
#else
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + byte_offset);
#endif
}

namespace atmo {

constexpr double kPi = 3.14159265358979323846;

struct LayerD {
  double width, exp_term, exp_scale, linear_term, constant_term;
};

// One precompute job: an atmosphere lowered at three wavelengths (one
// Model::Precompute call of atmosphere/model.cc:1048-1215), with its LUTs.
struct Job {
  // AtmosphereParameters (atmosphere/definitions.glsl:205-255), FP64, lengths
  // in the model's length unit.
  double solar[3];
  double rayleigh[3];
  double mie_sca[3];
  double mie_ext[3];
  double absorption[3];
  double albedo[3];
  double sun_angular_radius;
  double rb, rt;
  double mu_s_min;
  double g;
  LayerD prof_r[2], prof_m[2], prof_a[2];
  // Derived constants (host, FP64): H = sqrt(rt^2 - rb^2) and the A of the
  // mu_s mapping (functions.glsl:819-821).
  double H;
  double A;
  // luminance_from_radiance, row-major (atmosphere/model.cc:909, 939-943).
  float lum[9];
  // SKY/SUN_SPECTRAL_RADIANCE_TO_LUMINANCE (atmosphere/model.cc:668-686,
  // 735-742), used by the luminance lookups.
  float sky_k[3];
  float sun_k[3];
  int combined;  // COMBINED_SCATTERING_TEXTURES: no separate S_Mie output
  int orders;    // number of scattering orders of this job
  // Tile shapes (indices into kSdTile and kMsTile of kernels.h): host-side
  // block counts and launch grouping; the kernels are instantiated per shape.
  int sd_tile, ms_tile;
  // Order in which the multiple-scattering blocks walk the packed (r, mu)
  // rows (ms_row): 0 r-major (the LUT's own order), 1 mu-major, 2 ray
  // order, in which row p of the walk is ms_rows[p] = (z << 16) | y (the
  // model's rows sorted by the impact parameter of their view ray).
  int ms_row_order;
  const int* ms_rows;
  // LUT sizes.
  int tw, th;
  int nr, nmu, nmus, nnu;
  int ew, eh;
  // LUTs (RGBA FP32, GL layout, x fastest).
  float4* T;
  float4* S;
  float4* SM;
  float4* E;
  float4* dE;
  float4* dR;  // delta Rayleigh; delta multiple scattering from order 2 on
  float4* dM;
  // Delta scattering density, read only by the multiple-scattering pass:
  // RGB float texels (float3, 12 bytes; the reference's GL_RGB delta
  // textures), or for FP16-storage jobs (half_luts) three 21-bit floats
  // (uint2, 8 bytes, pack_f21x3).
  void* dJ;
  // FP16 storage (half_precision): the scattering LUTs are accumulated in
  // FP32 in S / SM during Init, then stored as 4 x FP16 texels times a
  // per-channel power-of-two scale: S16[t].c = S[t].c * scale[c],
  // scales[0..3] for S16, scales[4..7] for SM16.
  // Illuminance mode (several Precompute iterations): S and SM accumulate
  // in FP64 (4 doubles per texel) over the iterations and orders, and are
  // rounded to FP32 once at the end. Luminance channels can be small
  // differences of the iterations' contributions (the XYZ -> sRGB matrix has
  // negative entries), and FP32 partial sums lose them.
  // Each iteration job writes its own S64, SM64 and (through E) irradiance
  // accumulator; the final job sums the niter iterations' accumulators,
  // which start at S64, SM64 and E_it, into S, SM and E.
  int acc64;
  int niter;
  // Optical depths of the transmittance texels (4 doubles per texel: the
  // Rayleigh, Mie and absorption density integrals), shared by the jobs of
  // one model: they depend on the geometry and density profiles only, not
  // on the wavelengths. Null: the transmittance pass integrates them itself.
  double* OD;
  double* S64;
  double* SM64;
  float4* E_it;
  int half_luts;
  uint2* S16;
  uint2* SM16;
  // The lookup copy of S and SM (s_quad_taps): 128-byte blocks of quads of
  // 4 (S, SM) texel pairs, 8 float4 (FP32) or 4 uint4 (FP16) per quad; null
  // if absent.
  void* Q;
  float* scales;    // 8 floats, written by the conversion kernel
  unsigned* maxbits;  // 8 words: per-channel max |value| (float bits)
};

// ---------------------------------------------------------------------------
// FP64 scalar functions (atmosphere/functions.glsl).
// ---------------------------------------------------------------------------

ATMO_HD double clamp_d(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
ATMO_HD double clamp_cos(double mu) { return clamp_d(mu, -1.0, 1.0); }
ATMO_HD double safe_sqrt(double a) { return sqrt(a > 0.0 ? a : 0.0); }
ATMO_HD double clamp_radius(const Job& j, double r) {
  return clamp_d(r, j.rb, j.rt);
}

// functions.glsl:207-214
ATMO_HD double dist_top(const Job& j, double r, double mu) {
  double disc = r * r * (mu * mu - 1.0) + j.rt * j.rt;
  double d = -r * mu + safe_sqrt(disc);
  return d > 0.0 ? d : 0.0;
}
// functions.glsl:222-229
ATMO_HD double dist_bottom(const Job& j, double r, double mu) {
  double disc = r * r * (mu * mu - 1.0) + j.rb * j.rb;
  double d = -r * mu - safe_sqrt(disc);
  return d > 0.0 ? d : 0.0;
}
// functions.glsl:240-246
ATMO_HD bool hits_ground(const Job& j, double r, double mu) {
  return mu < 0.0 && r * r * (mu * mu - 1.0) + j.rb * j.rb >= 0.0;
}
// functions.glsl:680-687
ATMO_HD double dist_nearest(const Job& j, double r, double mu, bool ground) {
  return ground ? dist_bottom(j, r, mu) : dist_top(j, r, mu);
}
// functions.glsl:263-273. A null exponential term is skipped (its product
// with a finite exp is 0).
ATMO_HD double layer_density(const LayerD& l, double h) {
  double v = l.linear_term * h + l.constant_term;
  if (l.exp_term != 0.0) v += l.exp_term * exp(l.exp_scale * h);
  return clamp_d(v, 0.0, 1.0);
}
ATMO_HD double profile_density(const LayerD* p, double h) {
  return h < p[0].width ? layer_density(p[0], h) : layer_density(p[1], h);
}

// GL LINEAR + CLAMP_TO_EDGE taps along one axis of n texels, from the
// texel-space coordinate x = u * n - 1/2 (GL 3.3 core, section 3.8.11).
struct Taps {
  int i0, i1;
  float w;
};
ATMO_HD Taps axis_taps(double x, int n) {
  // Out-of-range and NaN coordinates are pinned to [-1, n] first, so the
  // integer taps below are always in bounds.
  if (!(x >= -1.0)) x = -1.0;
  if (x > static_cast<double>(n)) x = static_cast<double>(n);
  double f = floor(x);
  Taps t;
  t.w = static_cast<float>(x - f);
  int i = static_cast<int>(f);
  t.i0 = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
  t.i1 = i + 1 < 0 ? 0 : (i + 1 > n - 1 ? n - 1 : i + 1);
  return t;
}
// functions.glsl:342-348, composed with the GL texel-space coordinate.
ATMO_HD double unit_to_texel(double x, int n) {
  double u = 0.5 / n + x * (1.0 - 1.0 / n);
  return u * n - 0.5;
}
ATMO_HD double unit_from_tex(double u, int n) {
  return (u - 0.5 / n) / (1.0 - 1.0 / n);
}

// ---------------------------------------------------------------------------
// Transmittance LUT (functions.glsl:402-519).
// ---------------------------------------------------------------------------

// functions.glsl:402-421, as GL texel-space (x, y).
ATMO_HD void t_texel_coords(const Job& j, double r, double mu, double* x,
                            double* y) {
  double rho = safe_sqrt(r * r - j.rb * j.rb);
  double d = dist_top(j, r, mu);
  double d_min = j.rt - r;
  double d_max = rho + j.H;
  *x = unit_to_texel((d - d_min) / (d_max - d_min), j.tw);
  *y = unit_to_texel(rho / j.H, j.th);
}

// functions.glsl:427-447
ATMO_HD void t_rmu_from_uv(const Job& j, double u, double v, double* r,
                           double* mu) {
  double x_mu = unit_from_tex(u, j.tw);
  double x_r = unit_from_tex(v, j.th);
  double rho = j.H * x_r;
  *r = sqrt(rho * rho + j.rb * j.rb);
  double d_min = j.rt - *r;
  double d_max = rho + j.H;
  double d = d_min + x_mu * (d_max - d_min);
  double m = d == 0.0 ? 1.0 : (j.H * j.H - rho * rho - d * d) / (2.0 * *r * d);
  *mu = clamp_cos(m);
}

ATMO_HD float3 make3(float4 v) { return make_float3(v.x, v.y, v.z); }

// A texel of 4 FP16 values stored in 8 bytes.
ATMO_HD float4 unpack_half4(uint2 u) {
  _Float16 h[4];
  __builtin_memcpy(h, &u, sizeof(h));
  return make_float4(static_cast<float>(h[0]), static_cast<float>(h[1]),
                     static_cast<float>(h[2]), static_cast<float>(h[3]));
}
// Delta J of FP16-storage models: three 21-bit floats in 64 bits. The
// reference keeps its delta textures as GL_RGB16F in half-precision mode
// (atmosphere/model.cc:872-891); this keeps the FP32 sign and 8-bit
// exponent of each channel and rounds its mantissa to nearest (even) at 12
// bits: relative error at most 2^-13 at any magnitude (plain FP16 would lose
// the small texels of high orders, the night side, and channels far below
// the texel's largest), in 8 bytes instead of 12. The layout puts each
// channel where one or two instructions recover its float bits (the
// multiple-scattering pass decodes four texels per texel and sample):
//   x = c0 << 11 | c2 >> 10        c0 in x[31:11], c2[20:10] in x[10:0]
//   y = c2 << 22 | c1 << 1         c2[9:0] in y[31:22], c1 in y[21:1]
// so c0 = x masked, c1 = y << 10, and c2 = the 64-bit word (x high, y low)
// shifted right by 11 (one funnel shift), masked.
ATMO_HD uint32_t f21_bits(float v) {
  uint32_t b;
  __builtin_memcpy(&b, &v, sizeof(b));
  b += 0x3FFu + ((b >> 11) & 1u);
  return b >> 11;
}
ATMO_HD uint2 pack_f21x3(float r, float g, float b) {
  const uint32_t c0 = f21_bits(r), c1 = f21_bits(g), c2 = f21_bits(b);
  return make_uint2((c0 << 11) | (c2 >> 10), (c2 << 22) | (c1 << 1));
}
ATMO_HD float3 unpack_f21x3(uint2 u) {
  const uint32_t b0 = u.x & 0xFFFFF800u;
  const uint32_t b1 = u.y << 10;
  const uint32_t b2 =
      static_cast<uint32_t>(((static_cast<uint64_t>(u.x) << 32) | u.y) >> 11) & 0xFFFFF800u;
  float3 v;
  __builtin_memcpy(&v.x, &b0, sizeof(float));
  __builtin_memcpy(&v.y, &b1, sizeof(float));
  __builtin_memcpy(&v.z, &b2, sizeof(float));
  return v;
}

ATMO_HD uint2 pack_half4(float4 v) {
  _Float16 h[4] = {static_cast<_Float16>(v.x), static_cast<_Float16>(v.y),
                   static_cast<_Float16>(v.z), static_cast<_Float16>(v.w)};
  uint2 u;
  __builtin_memcpy(&u, h, sizeof(u));
  return u;
}

// Bilinear read of a 2D RGBA LUT at GL texel-space (x, y).
ATMO_HD float3 bilinear(const float4* lut, int w, int h, double x, double y) {
  Taps tx = axis_taps(x, w), ty = axis_taps(y, h);
  float4 a = gload(lut, ty.i0 * w + tx.i0), b = gload(lut, ty.i0 * w + tx.i1);
  float4 c = gload(lut, ty.i1 * w + tx.i0), d = gload(lut, ty.i1 * w + tx.i1);
  float wx1 = tx.w, wx0 = 1.0f - tx.w, wy1 = ty.w, wy0 = 1.0f - ty.w;
  return make_float3(wy0 * (wx0 * a.x + wx1 * b.x) + wy1 * (wx0 * c.x + wx1 * d.x),
                     wy0 * (wx0 * a.y + wx1 * b.y) + wy1 * (wx0 * c.y + wx1 * d.y),
                     wy0 * (wx0 * a.z + wx1 * b.z) + wy1 * (wx0 * c.z + wx1 * d.z));
}

// functions.glsl:473-480
ATMO_HD float3 t_lookup(const Job& j, const float4* T, double r, double mu) {
  double x, y;
  t_texel_coords(j, r, mu, &x, &y);
  return bilinear(T, j.tw, j.th, x, y);
}

ATMO_HD float min1(float v) { return v < 1.0f ? v : 1.0f; }

// Clamp to [0, hi]; NaN maps to 0. Used before every float -> index
// conversion so that indices stay in bounds whatever the input.
ATMO_HD float clamp_index_coord(float t, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  // One v_med3_f32 (the median of t, 0 and hi >= 0); with a NaN operand it
  // returns the minimum of the other two, 0. fminf(fmaxf()) takes three
  // instructions here: hi is canonicalized before the min.
  return __builtin_amdgcn_fmed3f(t, 0.0f, hi);
#else
  // fmaxf returns the non-NaN operand.
  return fminf(fmaxf(t, 0.0f), hi);
#endif
}

// functions.glsl:493-519
ATMO_HD float3 transmittance(const Job& j, const float4* T, double r,
                             double mu, double d, bool ground) {
  double r_d = clamp_radius(j, sqrt(d * d + 2.0 * r * mu * d + r * r));
  double mu_d = clamp_cos((r * mu + d) / r_d);
  float3 num, den;
  if (ground) {
    num = t_lookup(j, T, r_d, -mu_d);
    den = t_lookup(j, T, r, -mu);
  } else {
    num = t_lookup(j, T, r, mu);
    den = t_lookup(j, T, r_d, mu_d);
  }
  return make_float3(min1(num.x / den.x), min1(num.y / den.y),
                     min1(num.z / den.z));
}

ATMO_HD double smoothstep_d(double e0, double e1, double x) {
  double t = clamp_d((x - e0) / (e1 - e0), 0.0, 1.0);
  return t * t * (3.0 - 2.0 * t);
}


// ---------------------------------------------------------------------------
// Scattering LUT (functions.glsl:773-976).
// ---------------------------------------------------------------------------

// Texel-space coordinates of the packed 4D lookup: z (r), y (mu), the mu_s
// coordinate within a nu slice (x_in), and the nu slice coordinate.
struct SCoords {
  double z, y, x_in, nu;
};

// functions.glsl:781-787 (u_r), as GL texel-space z.
ATMO_HD double s_r_texel(const Job& j, double r) {
  double rho = safe_sqrt(r * r - j.rb * j.rb);
  return unit_to_texel(rho / j.H, j.nr);
}

// functions.glsl:789-812 (u_mu), as GL texel-space y. The ground and sky
// halves each map to MU/2 texels.
ATMO_HD double s_mu_texel(const Job& j, double r, double mu, bool ground) {
  double rho = safe_sqrt(r * r - j.rb * j.rb);
  double r_mu = r * mu;
  double disc = r_mu * r_mu - r * r + j.rb * j.rb;
  const int half = j.nmu / 2;
  double u_mu;
  if (ground) {
    double d = -r_mu - safe_sqrt(disc);
    double d_min = r - j.rb;
    double d_max = rho;
    double x = d_max == d_min ? 0.0 : (d - d_min) / (d_max - d_min);
    u_mu = 0.5 - 0.5 * (0.5 / half + x * (1.0 - 1.0 / half));
  } else {
    double d = -r_mu + safe_sqrt(disc + j.H * j.H);
    double d_min = j.rt - r;
    double d_max = rho + j.H;
    double x = (d - d_min) / (d_max - d_min);
    u_mu = 0.5 + 0.5 * (0.5 / half + x * (1.0 - 1.0 / half));
  }
  return u_mu * j.nmu - 0.5;
}

// functions.glsl:814-827 (u_mu_s), as texel-space x within a nu slice.
ATMO_HD double s_mus_texel(const Job& j, double mu_s) {
  double d = dist_top(j, j.rb, mu_s);
  double d_min = j.rt - j.rb;
  double d_max = j.H;
  double a = (d - d_min) / (d_max - d_min);
  double x_mu_s = (1.0 - a / j.A > 0.0 ? 1.0 - a / j.A : 0.0) / (1.0 + a);
  return unit_to_texel(x_mu_s, j.nmus);
}

// functions.glsl:773-831, with u_nu expressed as the nu slice coordinate
// u_nu * (NU - 1) of functions.glsl:967.
ATMO_HD SCoords s_coords(const Job& j, double r, double mu, double mu_s,
                         double nu, bool ground) {
  SCoords c;
  c.z = s_r_texel(j, r);
  c.y = s_mu_texel(j, r, mu, ground);
  c.x_in = s_mus_texel(j, mu_s);
  c.nu = (nu + 1.0) / 2.0 * (j.nnu - 1);
  return c;
}


// FP64-accumulating variants for the lookups: texels are FP32, weights and
// sums FP64, so that differences of two nearby lookups (GetSkyRadianceToPoint)
// keep the precision of the texel values.
struct D4 {
  double x, y, z, w;
};

// Texel sources of the lookups: FP32 texels, or FP16 texels with per-channel
// inverse scales. The scales are powers of two, so a lookup sums the stored
// values and scales its sums once (scale()): multiplying by a power of two
// commutes with rounding, and the result is the sum of the scaled texels bit
// for bit (the physical values are far from overflow and underflow).
struct F32Texels {
  const float4* p;
  ATMO_HD float4 operator[](int64_t i) const { return gload(p, i); }
  ATMO_HD float4 scale() const { return make_float4(1.0f, 1.0f, 1.0f, 1.0f); }
};
struct F16Texels {
  const uint2* p;
  float4 inv;
  ATMO_HD float4 operator[](int64_t i) const { return unpack_half4(gload(p, i)); }
  ATMO_HD float4 scale() const { return inv; }
};
ATMO_HD D4 scaled(D4 a, float4 s) {
  return D4{a.x * static_cast<double>(s.x), a.y * static_cast<double>(s.y),
            a.z * static_cast<double>(s.z), a.w * static_cast<double>(s.w)};
}
ATMO_HD float4 scaled(float4 a, float4 s) {
  return make_float4(a.x * s.x, a.y * s.y, a.z * s.z, a.w * s.w);
}

// The 16 taps of GetScattering (functions.glsl:958-976) with the GL
// texture() of each nu slice, a trilinear (mu_s, mu, r) filter: f(texel
// index, FP64 weight) in a fixed order. A slice index past the last one only
// ever gets weight 0 (the reference's x > 1 coordinate is clamped by the
// sampler); it is clamped here too.
// Real: the type of the tap weights (double: exact FP64 weights; float:
// the FP64 axis weights rounded, products in FP32).
template <class Real = double, class F>
ATMO_HD void s_taps(const Job& j, const SCoords& c, F&& f) {
  double cnu = c.nu;
  if (!(cnu >= -1.0)) cnu = -1.0;
  if (cnu > static_cast<double>(j.nnu)) cnu = static_cast<double>(j.nnu);
  const double fl = floor(cnu);
  const double ws = cnu - fl;
  int s0 = static_cast<int>(fl);
  int s1 = s0 + 1;
  s0 = s0 < 0 ? 0 : (s0 > j.nnu - 1 ? j.nnu - 1 : s0);
  s1 = s1 < 0 ? 0 : (s1 > j.nnu - 1 ? j.nnu - 1 : s1);
  double xz = c.z, xy = c.y, xx = c.x_in;
  Taps tz = axis_taps(xz, j.nr), ty = axis_taps(xy, j.nmu), tx = axis_taps(xx, j.nmus);
  // Exact FP64 weights (the Taps weights are rounded to float).
  auto frac = [](double x, int n) {
    if (!(x >= -1.0)) x = -1.0;
    if (x > static_cast<double>(n)) x = static_cast<double>(n);
    return x - floor(x);
  };
  const double fz = frac(xz, j.nr), fy = frac(xy, j.nmu), fx = frac(xx, j.nmus);
  const int w = j.nnu * j.nmus;
  const int zs[2] = {tz.i0, tz.i1};
  const int ys[2] = {ty.i0, ty.i1};
  const int xs[2] = {tx.i0, tx.i1};
  const Real wz[2] = {static_cast<Real>(1.0 - fz), static_cast<Real>(fz)};
  const Real wy[2] = {static_cast<Real>(1.0 - fy), static_cast<Real>(fy)};
  const Real wx[2] = {static_cast<Real>(1.0 - fx), static_cast<Real>(fx)};
  const int ss[2] = {s0, s1};
  const Real wsl[2] = {static_cast<Real>(1.0 - ws), static_cast<Real>(ws)};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int e = 0; e < 2; ++e)
        for (int q = 0; q < 2; ++q)
          f((static_cast<int64_t>(zs[a]) * j.nmu + ys[b]) * w + ss[q] * j.nmus + xs[e],
            wz[a] * wy[b] * wx[e] * wsl[q]);
}

ATMO_HD void acc_d4(D4* acc, double wt, float4 v) {
  acc->x = fma(wt, static_cast<double>(v.x), acc->x);
  acc->y = fma(wt, static_cast<double>(v.y), acc->y);
  acc->z = fma(wt, static_cast<double>(v.z), acc->z);
  acc->w = fma(wt, static_cast<double>(v.w), acc->w);
}

template <class Texels>
ATMO_HD D4 s_lookup_d(const Job& j, const Texels& lut, const SCoords& c) {
  D4 acc{0.0, 0.0, 0.0, 0.0};
  s_taps(j, c, [&](int64_t i, double wt) { acc_d4(&acc, wt, lut[i]); });
  return scaled(acc, lut.scale());
}

// The lookup copy of S and SM (Job::Q), in 128-byte blocks. For each (r,
// mu) row and each anchor (s, i) in [-1, NU - 1] x [-1, MU_S - 1], a quad
// holds the four texels that a bilinear (nu slice, mu_s) tap set anchored
// there reads, in the tap order of s_taps: (s, i), (s + 1, i), (s, i + 1),
// (s + 1, i + 1), each index clamped to the LUT as the taps clamp it, each
// texel an S and SM pair.
//   FP32 models: a block is one quad (8 float4), for each mu texel y.
//   FP16 models: a block is the quads of rows clamp(ya) and clamp(ya + 1)
//   (8 x 4 FP16 x 4), for each mu anchor ya in [-1, MU - 1]: the two mu
//   taps of a lookup.
// A scattering lookup (16 taps) then reads 4 (FP32) or 2 (FP16) blocks
// instead of 32 scattered texels, with the same texels, weights and sums.
// Random gathers cost per cache line, whatever part of it is used, so a
// block is one 128-byte line in either format.
ATMO_HD int lookup_copy_rows(int nmu, bool half) { return half ? nmu + 1 : nmu; }
ATMO_HD int64_t lookup_copy_blocks(int nr, int nmu, int nmus, int nnu, bool half) {
  return static_cast<int64_t>(nr) * lookup_copy_rows(nmu, half) * (nnu + 1) * (nmus + 1);
}
// Block of layer z, row ry (y, or ya + 1 for FP16 models) and anchor (s, i).
ATMO_HD int64_t lookup_copy_block(const Job& j, int z, int ry, int s, int i) {
  return ((static_cast<int64_t>(z) * lookup_copy_rows(j.nmu, j.half_luts != 0) + ry) *
              (j.nnu + 1) +
          (s + 1)) *
             (j.nmus + 1) +
         (i + 1);
}
// Anchor and fraction along one axis: the floor of the texel-space
// coordinate clamped to [-1, n], and its fraction; x = n takes anchor n - 1
// with fraction 1 (both of its texels are n - 1, whose weight is then 1 as
// in the taps, and the other 0).
struct QuadAxis {
  int a;
  double w;
};
ATMO_HD QuadAxis quad_axis(double x, int n) {
  if (!(x >= -1.0)) x = -1.0;
  if (x > static_cast<double>(n)) x = static_cast<double>(n);
  const double f = floor(x);
  QuadAxis q{static_cast<int>(f), x - f};
  if (q.a > n - 1) {
    q.a = n - 1;
    q.w = 1.0;
  }
  return q;
}
// The 4 quads of a lookup: f(quad slot, weights of its 4 texels), the
// weights the products of s_taps in the same order. The slot of a quad is
// its index in units of quads: the block for FP32 models, 2 x block + (mu
// tap) for FP16 ones.
template <class Real, class F>
ATMO_HD void s_quad_taps(const Job& j, const SCoords& c, F&& f) {
  const QuadAxis qs = quad_axis(c.nu, j.nnu), qx = quad_axis(c.x_in, j.nmus);
  double xz = c.z, xy = c.y;
  Taps tz = axis_taps(xz, j.nr), ty = axis_taps(xy, j.nmu);
  auto frac = [](double x, int n) {
    if (!(x >= -1.0)) x = -1.0;
    if (x > static_cast<double>(n)) x = static_cast<double>(n);
    return x - floor(x);
  };
  const double fz = frac(xz, j.nr), fy = frac(xy, j.nmu);
  const int zs[2] = {tz.i0, tz.i1};
  const int ys[2] = {ty.i0, ty.i1};
  // The mu anchor of the FP16 blocks: its rows clamp(ya) and clamp(ya + 1)
  // are ty.i0 and ty.i1.
  const int ya = quad_axis(xy, j.nmu).a;
  const bool half = j.half_luts != 0;
  const Real wz[2] = {static_cast<Real>(1.0 - fz), static_cast<Real>(fz)};
  const Real wy[2] = {static_cast<Real>(1.0 - fy), static_cast<Real>(fy)};
  const Real wx[2] = {static_cast<Real>(1.0 - qx.w), static_cast<Real>(qx.w)};
  const Real wsl[2] = {static_cast<Real>(1.0 - qs.w), static_cast<Real>(qs.w)};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      Real w[4];
      for (int e = 0; e < 2; ++e)
        for (int q = 0; q < 2; ++q) w[2 * e + q] = wz[a] * wy[b] * wx[e] * wsl[q];
      const int64_t slot =
          half ? 2 * lookup_copy_block(j, zs[a], ya + 1, qs.a, qx.a) + b
               : lookup_copy_block(j, zs[a], ys[b], qs.a, qx.a);
      f(slot, w);
    }
}
struct F32Quads {
  const float4* q;  // 8 float4 per quad: (S, SM) of its 4 texels
  ATMO_HD void load(int64_t quad, int k, float4* s, float4* m) const {
    *s = gload(q, 8 * quad + 2 * k);
    *m = gload(q, 8 * quad + 2 * k + 1);
  }
  ATMO_HD float4 scale_s() const { return make_float4(1.0f, 1.0f, 1.0f, 1.0f); }
  ATMO_HD float4 scale_m() const { return make_float4(1.0f, 1.0f, 1.0f, 1.0f); }
};
// FP16 quads: stored values, scaled after the sums as for F16Texels.
struct F16Quads {
  const uint4* q;  // 4 uint4 per quad: S and SM of a texel as 4 x FP16 each
  float4 inv_s, inv_m;
  ATMO_HD void load(int64_t quad, int k, float4* s, float4* m) const {
    const uint4 u = gload(q, 4 * quad + k);
    *s = unpack_half4(make_uint2(u.x, u.y));
    *m = unpack_half4(make_uint2(u.z, u.w));
  }
  ATMO_HD float4 scale_s() const { return inv_s; }
  ATMO_HD float4 scale_m() const { return inv_m; }
};
template <class Quads>
ATMO_HD void s_lookup_quads_d(const Job& j, const Quads& lut, const SCoords& c, D4* s,
                              D4* m) {
  D4 as{0.0, 0.0, 0.0, 0.0}, am{0.0, 0.0, 0.0, 0.0};
  s_quad_taps<double>(j, c, [&](int64_t quad, const double* w) {
    for (int k = 0; k < 4; ++k) {
      float4 vs, vm;
      lut.load(quad, k, &vs, &vm);
      acc_d4(&as, w[k], vs);
      acc_d4(&am, w[k], vm);
    }
  });
  *s = scaled(as, lut.scale_s());
  *m = scaled(am, lut.scale_m());
}

// The same lookups with FP32 weights and sums, for lookups whose result is
// not a difference of two lookups (GetSkyRadiance): relative error about
// 1e-7 instead of 1e-16, and half the instructions.
ATMO_HD void acc_f4(float4* acc, float wt, float4 v) {
  acc->x = fmaf(wt, v.x, acc->x);
  acc->y = fmaf(wt, v.y, acc->y);
  acc->z = fmaf(wt, v.z, acc->z);
  acc->w = fmaf(wt, v.w, acc->w);
}
ATMO_HD D4 to_d4(float4 v) { return D4{v.x, v.y, v.z, v.w}; }

template <class Texels>
ATMO_HD D4 s_lookup_f(const Job& j, const Texels& lut, const SCoords& c) {
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  s_taps<float>(j, c, [&](int64_t i, float wt) { acc_f4(&acc, wt, lut[i]); });
  return to_d4(scaled(acc, lut.scale()));
}

template <class Quads>
ATMO_HD void s_lookup_quads_f(const Job& j, const Quads& lut, const SCoords& c, D4* s,
                              D4* m) {
  float4 as = make_float4(0.0f, 0.0f, 0.0f, 0.0f), am = as;
  s_quad_taps<float>(j, c, [&](int64_t quad, const float* w) {
    for (int k = 0; k < 4; ++k) {
      float4 vs, vm;
      lut.load(quad, k, &vs, &vm);
      acc_f4(&as, w[k], vs);
      acc_f4(&am, w[k], vm);
    }
  });
  *s = to_d4(scaled(as, lut.scale_s()));
  *m = to_d4(scaled(am, lut.scale_m()));
}

// Real: the type of the weights and sums (double: exact FP64 weights;
// float: for lookups whose result is not a difference of two lookups).
template <class Real = double>
ATMO_HD D4 bilinear_r(const float4* lut, int w, int h, double x, double y) {
  Taps tx = axis_taps(x, w), ty = axis_taps(y, h);
  if (!(x >= -1.0)) x = -1.0;
  if (x > static_cast<double>(w)) x = static_cast<double>(w);
  if (!(y >= -1.0)) y = -1.0;
  if (y > static_cast<double>(h)) y = static_cast<double>(h);
  const double fx = x - floor(x), fy = y - floor(y);
  const float4 a = gload(lut, ty.i0 * w + tx.i0), b = gload(lut, ty.i0 * w + tx.i1);
  const float4 c = gload(lut, ty.i1 * w + tx.i0), d = gload(lut, ty.i1 * w + tx.i1);
  const Real wx1 = static_cast<Real>(fx), wx0 = static_cast<Real>(1.0 - fx);
  const Real wy1 = static_cast<Real>(fy), wy0 = static_cast<Real>(1.0 - fy);
  D4 o;
  o.x = wy0 * (wx0 * a.x + wx1 * b.x) + wy1 * (wx0 * c.x + wx1 * d.x);
  o.y = wy0 * (wx0 * a.y + wx1 * b.y) + wy1 * (wx0 * c.y + wx1 * d.y);
  o.z = wy0 * (wx0 * a.z + wx1 * b.z) + wy1 * (wx0 * c.z + wx1 * d.z);
  o.w = 0.0;
  return o;
}
ATMO_HD D4 bilinear_d(const float4* lut, int w, int h, double x, double y) {
  return bilinear_r<double>(lut, w, h, x, y);
}

template <class Real = double>
ATMO_HD D4 t_lookup_r(const Job& j, const float4* T, double r, double mu) {
  double x, y;
  t_texel_coords(j, r, mu, &x, &y);
  return bilinear_r<Real>(T, j.tw, j.th, x, y);
}
ATMO_HD D4 t_lookup_d(const Job& j, const float4* T, double r, double mu) {
  return t_lookup_r<double>(j, T, r, mu);
}

// functions.glsl:493-519 in FP64 arithmetic.
ATMO_HD D4 transmittance_d(const Job& j, const float4* T, double r, double mu,
                           double d, bool ground) {
  const double r_d = clamp_radius(j, sqrt(d * d + 2.0 * r * mu * d + r * r));
  const double mu_d = clamp_cos((r * mu + d) / r_d);
  D4 num, den;
  if (ground) {
    num = t_lookup_d(j, T, r_d, -mu_d);
    den = t_lookup_d(j, T, r, -mu);
  } else {
    num = t_lookup_d(j, T, r, mu);
    den = t_lookup_d(j, T, r_d, mu_d);
  }
  auto m1 = [](double v) { return v < 1.0 ? v : 1.0; };
  return D4{m1(num.x / den.x), m1(num.y / den.y), m1(num.z / den.z), 0.0};
}

// functions.glsl:552-563 in FP64 arithmetic (Real: the lookup's weights and
// sums).
template <class Real = double>
ATMO_HD D4 transmittance_to_sun_d(const Job& j, const float4* T, double r,
                                  double mu_s) {
  const double sin_h = j.rb / r;
  const double cos_h = -sqrt(1.0 - sin_h * sin_h > 0.0 ? 1.0 - sin_h * sin_h : 0.0);
  const double f = smoothstep_d(-sin_h * j.sun_angular_radius,
                                sin_h * j.sun_angular_radius, mu_s - cos_h);
  D4 t = t_lookup_r<Real>(j, T, r, mu_s);
  return D4{t.x * f, t.y * f, t.z * f, 0.0};
}

// Inverse mappings of a scattering texel (functions.glsl:837-890 and
// :905-926, frag_coord at the texel centre), split per axis.

// r of layer z; rho = sqrt(r^2 - rb^2).
ATMO_HD void s_layer_r(const Job& j, int z, double* r, double* rho) {
  *rho = j.H * unit_from_tex((z + 0.5) / j.nr, j.nr);
  *r = sqrt(*rho * *rho + j.rb * j.rb);
}

// mu and ground flag of row y at radius r.
ATMO_HD void s_row_mu(const Job& j, int y, double r, double rho, double* mu,
                      bool* ground) {
  const double u_mu = (y + 0.5) / j.nmu;
  const int half = j.nmu / 2;
  if (u_mu < 0.5) {
    double d_min = r - j.rb;
    double d_max = rho;
    double d = d_min + (d_max - d_min) * unit_from_tex(1.0 - 2.0 * u_mu, half);
    *mu = d == 0.0 ? -1.0 : clamp_cos(-(rho * rho + d * d) / (2.0 * r * d));
    *ground = true;
  } else {
    double d_min = j.rt - r;
    double d_max = rho + j.H;
    double d = d_min + (d_max - d_min) * unit_from_tex(2.0 * u_mu - 1.0, half);
    *mu = d == 0.0 ? 1.0
                   : clamp_cos((j.H * j.H - rho * rho - d * d) / (2.0 * r * d));
    *ground = false;
  }
}

// mu_s of column i (within a nu slice).
ATMO_HD double s_col_mus(const Job& j, int i) {
  double x_mu_s = unit_from_tex((i + 0.5) / j.nmus, j.nmus);
  double d_min = j.rt - j.rb;
  double d_max = j.H;
  double a = (j.A - x_mu_s * j.A) / (1.0 + x_mu_s * j.A);
  double d = d_min + (a < j.A ? a : j.A) * (d_max - d_min);
  return d == 0.0 ? 1.0 : clamp_cos((j.H * j.H - d * d) / (2.0 * j.rb * d));
}

// nu of slice k (sampled endpoint-exact over NU - 1), clamped to the range
// allowed by mu and mu_s.
ATMO_HD double s_slice_nu(const Job& j, int k, double mu, double mu_s) {
  double n = clamp_cos(static_cast<double>(k) / (j.nnu - 1) * 2.0 - 1.0);
  double half_range = sqrt((1.0 - mu * mu) * (1.0 - mu_s * mu_s));
  return clamp_d(n, mu * mu_s - half_range, mu * mu_s + half_range);
}


// ---------------------------------------------------------------------------
// Irradiance LUT (functions.glsl:1524-1601).
// ---------------------------------------------------------------------------

ATMO_HD void e_rmus_from_texel(const Job& j, int i, int k, double* r,
                               double* mu_s) {
  double x_mu_s = unit_from_tex((i + 0.5) / j.ew, j.ew);
  double x_r = unit_from_tex((k + 0.5) / j.eh, j.eh);
  *r = j.rb + x_r * (j.rt - j.rb);
  *mu_s = clamp_cos(2.0 * x_mu_s - 1.0);
}


template <class Real = double>
ATMO_HD D4 e_lookup_d(const Job& j, const float4* E, double r, double mu_s) {
  double x_r = (r - j.rb) / (j.rt - j.rb);
  double x_mu_s = mu_s * 0.5 + 0.5;
  return bilinear_r<Real>(E, j.ew, j.eh, unit_to_texel(x_mu_s, j.ew),
                          unit_to_texel(x_r, j.eh));
}

// functions.glsl:739-747
ATMO_HD double rayleigh_phase_d(double nu) {
  return 3.0 / (16.0 * kPi) * (1.0 + nu * nu);
}
ATMO_HD double mie_phase_d(double g, double nu) {
  const double k = 3.0 / (8.0 * kPi) * (1.0 - g * g) / (2.0 + g * g);
  const double y = 1.0 + g * g - 2.0 * g * nu;
  return k * (1.0 + nu * nu) / (y * sqrt(y));
}

// D4 texels of the FP64 accumulators (4 doubles per texel, 32 bytes).
ATMO_HD D4 load_d4(const double* p, int64_t t) {
  const double2 a = gload(reinterpret_cast<const double2*>(p), 2 * t);
  const double2 b = gload(reinterpret_cast<const double2*>(p), 2 * t + 1);
  return D4{a.x, a.y, b.x, b.y};
}
ATMO_HD void store_d4(double* p, int64_t t, D4 v) {
  gstore(reinterpret_cast<double2*>(p), 2 * t, make_double2(v.x, v.y));
  gstore(reinterpret_cast<double2*>(p), 2 * t + 1, make_double2(v.z, v.w));
}
ATMO_HD D4 add_d4(D4 a, D4 b) { return D4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
// luminance_from_radiance * v in FP64 (the products of FP32 factors are
// exact in FP64).
ATMO_HD D4 apply_lum_d(const float* L, const double* v) {
  const double x = v[0], y = v[1], z = v[2];
  return D4{L[0] * x + L[1] * y + L[2] * z, L[3] * x + L[4] * y + L[5] * z,
            L[6] * x + L[7] * y + L[8] * z, 0.0};
}
ATMO_HD D4 apply_lum_d(const float* L, float3 v) {
  const double w[3] = {v.x, v.y, v.z};
  return apply_lum_d(L, w);
}

ATMO_HD float3 apply_lum(const float* L, float3 v) {
  return make_float3(L[0] * v.x + L[1] * v.y + L[2] * v.z,
                     L[3] * v.x + L[4] * v.y + L[5] * v.z,
                     L[6] * v.x + L[7] * v.y + L[8] * v.z);
}

}  // namespace atmo

#endif  // ATMO_DEVICE_H_
