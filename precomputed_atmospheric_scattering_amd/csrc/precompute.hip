// Precompute kernels for gfx950: the six passes of Model::Precompute
// (atmosphere/model.cc:1048-1215) as HIP kernels.
//
// Every kernel runs over the blocks of one or more jobs (atmosphere x
// wavelength triple): block b belongs to the job whose block range contains
// it (first_block prefix sums), so one launch serves a single model or a
// batch of models with different LUT sizes. Blocks are 256 threads.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "atmo_device.h"
#include "kernels.h"

namespace atmo {
namespace {

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round-robin
// by workgroup id. This permutation of a job's nb blocks gives the blocks of
// one XCD (same id mod 8) a contiguous range of the job's work, so that the
// blocks resident on an XCD at a time work on neighbouring rows and share
// their L2 lines.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_contiguous(int lb, int nb) {
  const int q = nb / kXcds, rem = nb % kXcds;
  const int g = lb % kXcds;
  return g * q + (g < rem ? g : rem) + lb / kXcds;
}

// Block b's job: the i with first_block[i] <= b < first_block[i + 1]. The
// search is uniform across the block.
__device__ __forceinline__ int find_job(const int* __restrict__ first_block,
                                        int njobs, int b) {
  int lo = 0, hi = njobs;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (first_block[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ float rcp_f(float x) {
  return __builtin_amdgcn_rcpf(x);
}
__device__ __forceinline__ float rsq_f(float x) {
  return __builtin_amdgcn_rsqf(x);
}
__device__ __forceinline__ float sqrt_f(float x) {
  return __builtin_amdgcn_sqrtf(x);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}

// cos and sin of (m + 1/2) * pi / 16 for m = 0..15, evaluated at compile time
// (Taylor series in FP64, converged to the last bit for |x| <= pi).
constexpr double cx_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 24; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 24; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
#define ATMO_PHI(m) ((m + 0.5) * 3.14159265358979323846 / 16.0)
__device__ constexpr float kCosPhi[16] = {
    float(cx_cos(ATMO_PHI(0))),  float(cx_cos(ATMO_PHI(1))),
    float(cx_cos(ATMO_PHI(2))),  float(cx_cos(ATMO_PHI(3))),
    float(cx_cos(ATMO_PHI(4))),  float(cx_cos(ATMO_PHI(5))),
    float(cx_cos(ATMO_PHI(6))),  float(cx_cos(ATMO_PHI(7))),
    float(cx_cos(ATMO_PHI(8))),  float(cx_cos(ATMO_PHI(9))),
    float(cx_cos(ATMO_PHI(10))), float(cx_cos(ATMO_PHI(11))),
    float(cx_cos(ATMO_PHI(12))), float(cx_cos(ATMO_PHI(13))),
    float(cx_cos(ATMO_PHI(14))), float(cx_cos(ATMO_PHI(15)))};
__device__ constexpr float kSinPhi[16] = {
    float(cx_sin(ATMO_PHI(0))),  float(cx_sin(ATMO_PHI(1))),
    float(cx_sin(ATMO_PHI(2))),  float(cx_sin(ATMO_PHI(3))),
    float(cx_sin(ATMO_PHI(4))),  float(cx_sin(ATMO_PHI(5))),
    float(cx_sin(ATMO_PHI(6))),  float(cx_sin(ATMO_PHI(7))),
    float(cx_sin(ATMO_PHI(8))),  float(cx_sin(ATMO_PHI(9))),
    float(cx_sin(ATMO_PHI(10))), float(cx_sin(ATMO_PHI(11))),
    float(cx_sin(ATMO_PHI(12))), float(cx_sin(ATMO_PHI(13))),
    float(cx_sin(ATMO_PHI(14))), float(cx_sin(ATMO_PHI(15)))};
#undef ATMO_PHI

// ---------------------------------------------------------------------------
// Transmittance (functions.glsl:275-320, :454-463). FP64 throughout: the
// 501-sample trapezoid of each texel is split over the 64 lanes of a wave and
// reduced with cross-lane shuffles, 4 texels per block. The three optical
// depths do not depend on the wavelengths, so a model with several
// Precompute iterations integrates them once (k_optical_depth, into the
// job's OD) and each of its transmittance LUTs is then one exp per texel
// and channel; a model with one iteration integrates them in
// k_transmittance itself.
// ---------------------------------------------------------------------------
// The optical depths of transmittance texel 'texel' (sums over the wave's
// lanes, returned on every lane).
__device__ D4 optical_depths(const Job& J, int texel, int lane) {
  const int i = texel % J.tw, k = texel / J.tw;
  double r, mu;
  t_rmu_from_uv(J, (i + 0.5) / J.tw, (k + 0.5) / J.th, &r, &mu);
  constexpr int kIntervals = 500;
  const double dx = dist_top(J, r, mu) / kIntervals;
  double lr = 0.0, lm = 0.0, la = 0.0;
  for (int s = lane; s <= kIntervals; s += 64) {
    const double d = s * dx;
    const double h = sqrt(d * d + 2.0 * r * mu * d + r * r) - J.rb;
    const double w = (s == 0 || s == kIntervals) ? 0.5 : 1.0;
    lr += w * profile_density(J.prof_r, h);
    lm += w * profile_density(J.prof_m, h);
    la += w * profile_density(J.prof_a, h);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lr += __shfl_xor(lr, off);
    lm += __shfl_xor(lm, off);
    la += __shfl_xor(la, off);
  }
  return D4{lr * dx, lm * dx, la * dx, 0.0};
}

__device__ void store_transmittance(const Job& J, int texel, const D4& od) {
  float4 t;
  t.x = static_cast<float>(exp(-(J.rayleigh[0] * od.x + J.mie_ext[0] * od.y +
                                 J.absorption[0] * od.z)));
  t.y = static_cast<float>(exp(-(J.rayleigh[1] * od.x + J.mie_ext[1] * od.y +
                                 J.absorption[1] * od.z)));
  t.z = static_cast<float>(exp(-(J.rayleigh[2] * od.x + J.mie_ext[2] * od.y +
                                 J.absorption[2] * od.z)));
  t.w = 0.0f;
  gstore(J.T, texel, t);
}

// One wave per texel (4 texels per block), into the job's OD.
__global__ __launch_bounds__(kBlock) void k_optical_depth(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs) {
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int lb = blockIdx.x - first_block[ji];
  const int lane = threadIdx.x & 63;
  const int texel = lb * (kBlock / 64) + (threadIdx.x >> 6);
  if (texel >= J.tw * J.th || J.OD == nullptr) return;
  const D4 od = optical_depths(J, texel, lane);
  if (lane == 0) store_d4(J.OD, texel, od);
}

// Jobs with OD: one thread per texel. Others: one wave per texel (4 texels
// per block), integrating the optical depths.
__global__ __launch_bounds__(kBlock) void k_transmittance(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs) {
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int lb = blockIdx.x - first_block[ji];
  if (J.OD != nullptr) {
    const int texel = lb * kBlock + threadIdx.x;
    if (texel < J.tw * J.th) store_transmittance(J, texel, load_d4(J.OD, texel));
    return;
  }
  const int lane = threadIdx.x & 63;
  const int texel = lb * (kBlock / 64) + (threadIdx.x >> 6);
  if (texel >= J.tw * J.th) return;
  const D4 od = optical_depths(J, texel, lane);
  if (lane == 0) store_transmittance(J, texel, od);
}

// ---------------------------------------------------------------------------
// Direct irradiance (functions.glsl:1443-1461, :1558-1567) into delta E; the
// job's irradiance LUT (or its iteration's accumulator) is cleared, the
// direct term being excluded from it (atmosphere/model.cc:132-140).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_direct_irradiance(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs) {
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int texel = (blockIdx.x - first_block[ji]) * kBlock + threadIdx.x;
  if (texel >= J.ew * J.eh) return;
  double r, mu_s;
  e_rmus_from_texel(J, texel % J.ew, texel / J.ew, &r, &mu_s);
  const double alpha = J.sun_angular_radius;
  const double avg_cos =
      mu_s < -alpha ? 0.0
                    : (mu_s > alpha ? mu_s
                                    : (mu_s + alpha) * (mu_s + alpha) / (4.0 * alpha));
  const float3 t = t_lookup(J, J.T, r, mu_s);
  gstore(J.dE, texel,
         make_float4(static_cast<float>(J.solar[0] * avg_cos) * t.x,
                     static_cast<float>(J.solar[1] * avg_cos) * t.y,
                     static_cast<float>(J.solar[2] * avg_cos) * t.z, 0.0f));
  gstore(J.E, texel, make_float4(0.f, 0.f, 0.f, 0.f));
}

// ---------------------------------------------------------------------------
// Single scattering (functions.glsl:650-730, :933-945) and the output packing
// of atmosphere/model.cc:142-157.
//
// A block covers up to 256 texels of one packed row (fixed r layer and mu
// row). Along a row only mu_s and nu change, so every quantity of the 51
// ray-march samples that depends on (r, mu, d_i) alone - the view
// transmittance, the densities, the transmittance-LUT row of the sample
// radius - is computed once per block in FP64 and kept in LDS. Per texel and
// sample what remains is the sun-side transmittance lookup at
// (r_d, mu_s_d) and the visible-disc smoothstep.
// ---------------------------------------------------------------------------
constexpr int kMarch = 51;  // 50 trapezoid intervals

// Integer part and fraction of a texel coordinate already clamped to
// [0, hi]: for t >= 0 truncation is floor, and v_fract_f32 returns
// t - floor(t) exactly (a difference of two floats within a factor 2, or t
// itself below 1), one instruction fewer than floor and subtract.
__device__ __forceinline__ int split_coord(float t, float* frac) {
  *frac = __builtin_amdgcn_fractf(t);
  return static_cast<int>(t);
}

// FP32 path (radiance mode) and FP64 path (illuminance mode, whose S and SM
// accumulate in FP64): per-sample geometry in Real, the r mu_s + d nu term
// and the horizon term of every texel in FP64 in both.
template <class Real>
struct SsSample {
  double d;        // distance along the view ray
  double inv_r;    // 1 / r_d
  double rho;      // sqrt(r_d^2 - rb^2)
  Real r;          // r_d
  Real c;          // rt^2 - r_d^2
  Real dmin;       // rt - r_d
  Real inv_range;  // 1 / (rho + H - dmin)
  Real e;          // sin(theta_h) * sun angular radius
  Real inv_span;   // 1 / (2 e)
  Real wy;
  int t0, t1;      // first texels of the transmittance LUT rows of r_d
  Real fr[3];      // trapezoid weight * view transmittance * Rayleigh density
  Real fm[3];      // same with the Mie density
};

// functions.glsl:402-421 at (r_d, mu_s_d): DistanceToTopAtmosphereBoundary
// in the cancellation-free form, the LUT column coordinate, and the bilinear
// transmittance lookup (functions.glsl:473-480).
__device__ __forceinline__ void tsun_lookup(const Job& J, const SsSample<float>& S,
                                            float mus, float o[3]) {
  const float tt = S.r * mus;
  const float sq = sqrt_f(fmaf(tt, tt, S.c));
  const float dd = tt > 0.0f ? S.c * rcp_f(sq + tt) : sq - tt;
  const float xm = (dd - S.dmin) * S.inv_range;
  const float ut = clamp_index_coord(xm * static_cast<float>(J.tw - 1),
                                     static_cast<float>(J.tw - 1));
  float w;
  const int i0 = split_coord(ut, &w);
  const int i1 = min(i0 + 1, J.tw - 1);
  const float4* __restrict__ T = J.T;
  const uint32_t r0 = static_cast<uint32_t>(S.t0), r1 = static_cast<uint32_t>(S.t1);
  const float4 a = gload32(T, r0 + i0), b = gload32(T, r0 + i1);
  const float4 c = gload32(T, r1 + i0), d = gload32(T, r1 + i1);
  const float wy1 = S.wy, wy0 = 1.0f - S.wy;
  o[0] = fmaf(wy1, fmaf(w, d.x - c.x, c.x), wy0 * fmaf(w, b.x - a.x, a.x));
  o[1] = fmaf(wy1, fmaf(w, d.y - c.y, c.y), wy0 * fmaf(w, b.y - a.y, a.y));
  o[2] = fmaf(wy1, fmaf(w, d.z - c.z, c.z), wy0 * fmaf(w, b.z - a.z, a.z));
}

// In FP64 the distance needs no cancellation-free form: sq - tt loses at most
// about 1e-16 sq = 6e-13 km, 1e-13 texel of the LUT coordinate, and saves a
// division and a divergent branch per sample.
__device__ __forceinline__ void tsun_lookup(const Job& J, const SsSample<double>& S,
                                            double mus, double o[3]) {
  const double tt = S.r * mus;
  const double sq = sqrt(fma(tt, tt, S.c));
  const double dd = sq - tt;
  double ut = (dd - S.dmin) * S.inv_range * (J.tw - 1);
  ut = fmin(fmax(ut, 0.0), static_cast<double>(J.tw - 1));
  // ut >= 0: truncation is floor, and v_fract_f64 gives ut - floor(ut)
  // exactly (as split_coord in FP32).
  const double w = __builtin_amdgcn_fract(ut);
  const int i0 = static_cast<int>(ut);
  const int i1 = min(i0 + 1, J.tw - 1);
  const float4* __restrict__ T = J.T;
  const uint32_t r0 = static_cast<uint32_t>(S.t0), r1 = static_cast<uint32_t>(S.t1);
  const float4 a = gload32(T, r0 + i0), b = gload32(T, r0 + i1);
  const float4 c = gload32(T, r1 + i0), d = gload32(T, r1 + i1);
  const double wy1 = S.wy, wy0 = 1.0 - S.wy;
  o[0] = fma(wy1, fma(w, double(d.x) - c.x, double(c.x)), wy0 * fma(w, double(b.x) - a.x, double(a.x)));
  o[1] = fma(wy1, fma(w, double(d.y) - c.y, double(c.y)), wy0 * fma(w, double(b.y) - a.y, double(a.y)));
  o[2] = fma(wy1, fma(w, double(d.z) - c.z, double(c.z)), wy0 * fma(w, double(b.z) - a.z, double(a.z)));
}

// The outputs of one single-scattering texel: the delta Rayleigh and Mie
// LUTs, and the scattering accumulators (FP64 in illuminance mode).
template <bool kF64, class Real>
__device__ __forceinline__ void ss_store(const Job& J, int64_t t_idx, double dx,
                                         const Real (&ar)[3], const Real (&am)[3]) {
  double ray[3], mie[3];
  for (int c = 0; c < 3; ++c) {
    ray[c] = static_cast<double>(ar[c]) * (dx * J.solar[c] * J.rayleigh[c]);
    mie[c] = static_cast<double>(am[c]) * (dx * J.solar[c] * J.mie_sca[c]);
  }
  gstore(J.dR, t_idx, make_float4(static_cast<float>(ray[0]), static_cast<float>(ray[1]),
                                  static_cast<float>(ray[2]), 0.0f));
  gstore(J.dM, t_idx, make_float4(static_cast<float>(mie[0]), static_cast<float>(mie[1]),
                                  static_cast<float>(mie[2]), 0.0f));
  if (kF64) {
    const D4 lr = apply_lum_d(J.lum, ray), lm = apply_lum_d(J.lum, mie);
    store_d4(J.S64, t_idx, D4{lr.x, lr.y, lr.z, lm.x});
    if (!J.combined) store_d4(J.SM64, t_idx, D4{lm.x, lm.y, lm.z, 0.0});
    return;
  }
  const float3 rayf = make_float3(static_cast<float>(ray[0]), static_cast<float>(ray[1]),
                                  static_cast<float>(ray[2]));
  const float3 mief = make_float3(static_cast<float>(mie[0]), static_cast<float>(mie[1]),
                                  static_cast<float>(mie[2]));
  const float3 lr = apply_lum(J.lum, rayf);
  const float3 lm = apply_lum(J.lum, mief);
  gstore(J.S, t_idx, make_float4(lr.x, lr.y, lr.z, lm.x));
  if (!J.combined) gstore(J.SM, t_idx, make_float4(lm.x, lm.y, lm.z, 0.0f));
}

// Texels per thread of the single-scattering pass for packed rows of W
// texels: 2 once a row fills two blocks (texels x and x + kBlock of a row
// then share each sample's transmittance rows), else 1.
__host__ __device__ inline int ss_texels_per_thread(int w) {
  return w >= 4 * kBlock ? 4 : (w >= 2 * kBlock ? 2 : 1);
}

// kF64: the kernel serves the jobs with FP64 accumulators (illuminance
// mode) and computes the sun-side lookups and the sums in FP64; otherwise it
// serves the others. kTpt: texels per thread, ss_texels_per_thread(W) of the
// job: texels x + k kBlock (k < kTpt) of one row, whose sun lookups read the
// same two transmittance rows per sample, so a wave keeps kTpt times the
// gathers in flight and the later texels often find their lines in L1. The
// instances are launched over the same blocks when a launch holds several
// kinds; each returns at once on the others' jobs.
template <bool kF64, int kTpt>
__global__ __launch_bounds__(kBlock) void k_single_scattering(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs) {
  using Real = typename std::conditional<kF64, double, float>::type;
  __shared__ SsSample<Real> smp[kMarch];
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  if ((J.acc64 != 0) != kF64) return;
  const int W = J.nnu * J.nmus;
  if (ss_texels_per_thread(W) != kTpt) return;
  const int lb = blockIdx.x - first_block[ji];
  const int tiles = (W + kBlock * kTpt - 1) / (kBlock * kTpt);
  const int tile = lb % tiles;
  const int row = lb / tiles;  // z * nmu + y
  const int y = row % J.nmu, z = row / J.nmu;
  const int tid = threadIdx.x;

  double r, rho, mu;
  bool ground;
  s_layer_r(J, z, &r, &rho);
  s_row_mu(J, y, r, rho, &mu, &ground);
  const double dx = dist_nearest(J, r, mu, ground) / (kMarch - 1);

  if (tid < kMarch) {
    SsSample<Real>& S = smp[tid];
    const double d = tid * dx;
    const double r_d = clamp_radius(J, sqrt(d * d + 2.0 * r * mu * d + r * r));
    double tv[3];
    if (kF64) {
      const D4 t = transmittance_d(J, J.T, r, mu, d, ground);
      tv[0] = t.x;
      tv[1] = t.y;
      tv[2] = t.z;
    } else {
      const float3 t = transmittance(J, J.T, r, mu, d, ground);
      tv[0] = t.x;
      tv[1] = t.y;
      tv[2] = t.z;
    }
    const double rho_d = safe_sqrt(r_d * r_d - J.rb * J.rb);
    S.d = d;
    S.inv_r = 1.0 / r_d;
    S.rho = rho_d;
    S.r = static_cast<Real>(r_d);
    S.c = static_cast<Real>(J.rt * J.rt - r_d * r_d);
    S.dmin = static_cast<Real>(J.rt - r_d);
    S.inv_range = static_cast<Real>(1.0 / (rho_d + J.H - (J.rt - r_d)));
    const double e = J.rb / r_d * J.sun_angular_radius;
    S.e = static_cast<Real>(e);
    S.inv_span = static_cast<Real>(0.5 / e);
    const double ty = unit_to_texel(rho_d / J.H, J.th);
    const Taps tt = axis_taps(ty, J.th);
    S.t0 = tt.i0 * J.tw;
    S.t1 = tt.i1 * J.tw;
    if (kF64) {
      // The exact weight (Taps rounds it to float).
      const double cy = fmin(fmax(ty, -1.0), static_cast<double>(J.th));
      S.wy = static_cast<Real>(cy - floor(cy));
    } else {
      S.wy = static_cast<Real>(tt.w);
    }
    const double w = (tid == 0 || tid == kMarch - 1) ? 0.5 : 1.0;
    const double dr = w * profile_density(J.prof_r, r_d - J.rb);
    const double dm = w * profile_density(J.prof_m, r_d - J.rb);
    for (int c = 0; c < 3; ++c) {
      S.fr[c] = static_cast<Real>(dr * tv[c]);
      S.fm[c] = static_cast<Real>(dm * tv[c]);
    }
  }
  __syncthreads();

  double mu_s[kTpt], nu[kTpt], rmus[kTpt];
  bool on[kTpt];
  bool any = false;
#pragma unroll
  for (int j = 0; j < kTpt; ++j) {
    const int x = tile * kBlock * kTpt + j * kBlock + tid;
    on[j] = x < W;
    any = any || on[j];
    mu_s[j] = on[j] ? s_col_mus(J, x % J.nmus) : 0.0;
    nu[j] = on[j] ? s_slice_nu(J, x / J.nmus, mu, mu_s[j]) : 0.0;
    rmus[j] = r * mu_s[j];
  }
  if (!any) return;
  Real ar[kTpt][3], am[kTpt][3];
#pragma unroll
  for (int j = 0; j < kTpt; ++j)
    for (int c = 0; c < 3; ++c) ar[j][c] = am[j][c] = Real(0);
  for (int s = 0; s < kMarch; ++s) {
    const SsSample<Real>& S = smp[s];
#pragma unroll
    for (int j = 0; j < kTpt; ++j) {
      if (!on[j]) continue;
      const double t = fma(S.d, nu[j], rmus[j]);  // r * mu_s + d * nu = r_d * mu_s_d
      const Real xs = static_cast<Real>((t + S.rho) * S.inv_r);  // mu_s_d - cos(theta_h)
      // Clamps as max/min, 2 instructions each (the operands are finite; the
      // ternary form, which also keeps NaNs, compiles to compares and
      // selects).
      const Real mus = fmin(fmax(static_cast<Real>(t * S.inv_r), Real(-1)), Real(1));
      const Real u = fmin(fmax((xs + S.e) * S.inv_span, Real(0)), Real(1));
      const Real vis = u * u * fma(Real(-2), u, Real(3));
      Real ts[3];
      tsun_lookup(J, S, mus, ts);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const Real f = ts[c] * vis;
        ar[j][c] = fma(f, S.fr[c], ar[j][c]);
        am[j][c] = fma(f, S.fm[c], am[j][c]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kTpt; ++j) {
    if (!on[j]) continue;
    const int x = tile * kBlock * kTpt + j * kBlock + tid;
    ss_store<kF64>(J, static_cast<int64_t>(row) * W + x, dx, ar[j], am[j]);
  }
}

// ---------------------------------------------------------------------------
// Scattering density (functions.glsl:1163-1260, :1348-1367).
//
// A block covers up to 256 texels (mu, nu) of one (r layer, mu_s column)
// plane. All lookups of the 16 x 32 incident directions are made at the
// texel's own r and mu_s (whose LUT coordinates round-trip to the texel
// centres), and for incident zenith angle theta_l at the fixed mu coordinate
// u_mu(r, cos theta_l). The 4D lookup therefore reduces to a linear
// interpolation in nu of a per-theta table of NU values (the two mu rows
// already blended), built once per block in LDS. The 32 azimuths pair up as
// phi and 2 pi - phi, which share the phase function of the outgoing
// direction. Ground-hitting theta_l add the ground-reflected term read from a
// table of delta-irradiance row 0 (r = rb).
// ---------------------------------------------------------------------------
constexpr int kTheta = 16;

// A table of n values v[0..n-1] has n + 1 entries of two float4:
// (v.x, v.y, v.z, dv.x), (dv.y, dv.z, -, -). Entry 0 is a guard holding v[0]
// with dv = 0; entry k >= 1 holds v[k - 1] with dv = v[k] - v[k - 1], and
// dv = 0 for the last. It is read at the coordinate t' = t + 1 of the
// texel-space coordinate t in [0, n - 1]: this reproduces GL's clamp to
// edge for every t in (-2, n) without a clamp instruction. Below 0 the
// guard gives v[0] whatever the weight; above n - 1 the last entry gives
// v[n - 1]. The coordinates are finite dot products of unit vectors mapped
// to [0, n - 1], off by rounding at most; were t' NaN, the float -> int
// conversion gives 0, still in bounds.
__device__ __forceinline__ float3 table_lerp(const float4* __restrict__ tab, float t1) {
  float w;
  const int i = split_coord(t1, &w);
  const float4 e0 = tab[2 * i];
  const float4 e1 = tab[2 * i + 1];
  return make_float3(fmaf(w, e0.w, e0.x), fmaf(w, e1.x, e0.y),
                     fmaf(w, e1.y, e0.z));
}

// Writes table entry k (k >= 1: v[k - 1] and its successor vn; k = 0: the
// guard, vn = v).
__device__ __forceinline__ void put_entry(float4* tab, int k, float3 v, float3 vn) {
  tab[2 * k] = make_float4(v.x, v.y, v.z, vn.x - v.x);
  tab[2 * k + 1] = make_float4(vn.y - v.y, vn.z - v.z, 0.f, 0.f);
}

struct SdTheta {
  float cos_t, sin_t;
  float ga, gb;   // ground normal . sun = ga + gb * nu1
  float gf[3];    // transmittance to ground * albedo / pi * d_omega
  int ground;
};

// One incident zenith angle theta_l for the kChunk outputs (same r, mu,
// mu_s; consecutive nu) of a thread. The phase function of the scattering
// towards the view direction depends on (mu, theta_l, phi) only, so it is
// computed once per azimuth pair for all kChunk outputs.
template <int kChunk, bool kOrder2, bool kGround>
__device__ __forceinline__ void sd_theta(
    const float4* __restrict__ tab_a, const float4* __restrict__ tab_b,
    const float4* __restrict__ etab, float nu_max, float e_max, float A, float P,
    float Q, const float (&Bv)[kChunk], const float (&Cv)[kChunk], float g1,
    float mg2, const float cR[3], const float cM[3], const SdTheta& th,
    float (&acc)[kChunk][3]) {
  // nu table coordinate of the incident direction at azimuth phi:
  // t = (nu1 + 1) / 2 * (NU - 1) with nu1 = A + B cos(phi) +- C sin(phi),
  // as one FMA from per-angle constants; likewise the ground-irradiance
  // table coordinate (mu_g + 1) / 2 * (E_W - 1) with mu_g = ga + gb nu1.
  const float hn = 0.5f * nu_max;
  const float inv_hn = 1.0f / hn;
  // t' = t + 1 (table_lerp): the 1 is folded into the constant term.
  const float tA = fmaf(A, hn, hn + 1.0f);
  float tB[kChunk], tC[kChunk];
#pragma unroll
  for (int j = 0; j < kChunk; ++j) {
    tB[j] = Bv[j] * hn;
    tC[j] = Cv[j] * hn;
  }
  const float he = 0.5f * e_max;
  const float eB = th.gb * inv_hn * he;
  // The ground-irradiance coordinate plus 1, as a function of t' = t + 1.
  const float eA = fmaf(th.ga - th.gb, he, he) - eB + 1.0f;
  const float nu1_0 = -1.0f - inv_hn;  // nu1 = t' / hn + nu1_0
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const float c = kCosPhi[m], s = kSinPhi[m];
    // Phase function of the scattering towards -omega, shared by phi and
    // 2 pi - phi and by the nu outputs.
    const float nu2 = fmaf(Q, c, P);
    const float x2 = fmaf(nu2, nu2, 1.0f);
    const float rs = rsq_f(fmaf(mg2, nu2, g1));
    const float rs3 = rs * rs * rs;
    const float p0 = x2 * fmaf(cM[0], rs3, cR[0]);
    const float p1 = x2 * fmaf(cM[1], rs3, cR[1]);
    const float p2 = x2 * fmaf(cM[2], rs3, cR[2]);
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const float tb = fmaf(tB[j], c, tA);
      float l0, l1, l2;
#pragma unroll
      for (int sgn = 0; sgn < 2; ++sgn) {
        const float t = sgn == 0 ? fmaf(tC[j], s, tb) : fmaf(-tC[j], s, tb);
        float3 v = table_lerp(tab_a, t);
        float x1 = 1.0f;  // order 2: the Rayleigh phase factor 1 + nu1^2
        if (kOrder2) {
          // Order-1 incident radiance: Rayleigh and Mie single scattering
          // times their phase functions of nu1 (constants folded in tables).
          const float nu1 = fmaf(t, inv_hn, nu1_0);
          const float3 vm = table_lerp(tab_b, t);
          const float r1 = rsq_f(fmaf(mg2, nu1, g1));
          const float q = r1 * r1 * r1;
          x1 = fmaf(nu1, nu1, 1.0f);
          v.x = fmaf(vm.x, q, v.x);
          v.y = fmaf(vm.y, q, v.y);
          v.z = fmaf(vm.z, q, v.z);
        }
        if (kGround) {
          if (kOrder2) {
            v.x = x1 * v.x;
            v.y = x1 * v.y;
            v.z = x1 * v.z;
            x1 = 1.0f;
          }
          const float3 e = table_lerp(etab, fmaf(eB, t, eA));
          v.x = fmaf(th.gf[0], e.x, v.x);
          v.y = fmaf(th.gf[1], e.y, v.y);
          v.z = fmaf(th.gf[2], e.z, v.z);
        }
        // l += x1 * v (x1 = 1 unless order 2 without ground term).
        if (sgn == 0) {
          l0 = kOrder2 && !kGround ? x1 * v.x : v.x;
          l1 = kOrder2 && !kGround ? x1 * v.y : v.y;
          l2 = kOrder2 && !kGround ? x1 * v.z : v.z;
        } else if (kOrder2 && !kGround) {
          l0 = fmaf(x1, v.x, l0);
          l1 = fmaf(x1, v.y, l1);
          l2 = fmaf(x1, v.z, l2);
        } else {
          l0 += v.x;
          l1 += v.y;
          l2 += v.z;
        }
      }
      acc[j][0] = fmaf(l0, p0, acc[j][0]);
      acc[j][1] = fmaf(l1, p1, acc[j][1]);
      acc[j][2] = fmaf(l2, p2, acc[j][2]);
    }
  }
}

// Five waves per SIMD: the 4-output shapes then fit 96 VGPRs (101 unbounded,
// four waves), and the pass, which issues a VALU instruction in about 90% of
// the SIMD cycles, runs 2-3% faster on MI355X.
template <int kChunk, int kB>
__global__ __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(5))) void
k_scattering_density(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs, int order) {
  extern __shared__ float4 lds[];
  __shared__ SdTheta th[kTheta];
  __shared__ int s_y0[kTheta], s_y1[kTheta];
  __shared__ float s_wy[kTheta], s_dw[kTheta];
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int lb = blockIdx.x - first_block[ji];
  const int nnu = J.nnu, nmus = J.nmus, nmu = J.nmu;
  const int nch = (nnu + kChunk - 1) / kChunk;
  const int plane = nmu * nch;  // threads per (r, mu_s) plane
  const int tiles = (plane + kB - 1) / kB;
  // The blocks of neighbouring mu_s columns of a layer write the same cache
  // lines of delta J (a plane's texels are W texels apart). The block order
  // walks groups of G consecutive mu_s columns (G = the largest power of two
  // dividing MU_S, at most 16; 16 x 12 B spans a 128 B line and more), layer
  // by layer within a group, and each XCD takes a contiguous range of that
  // order: the blocks writing a line share one L2, where the partial lines
  // merge before they are written back, and each XCD gets planes of many
  // layers (the ground-heavy low layers cost more).
  const int G = min(nmus & -nmus, 16);
  const int lbx = xcd_contiguous(lb, tiles * nmus * J.nr);
  const int tile = lbx % tiles;
  const int pl = lbx / tiles;
  const int im = pl % G, rest = pl / G;
  const int k = rest % J.nr, imus = (rest / J.nr) * G + im;
  const int W = nnu * nmus;
  const bool two = order == 2;
  // Per theta a table of NU + 1 entries (table_lerp); then the
  // ground-irradiance table of E_W + 1 entries.
  const int stride = 2 * (nnu + 1);
  float4* tab_a = lds;
  float4* tab_b = lds + kTheta * stride;
  float4* etab = lds + (two ? 2 : 1) * kTheta * stride;
  const int tid = threadIdx.x;

  double r, rho;
  s_layer_r(J, k, &r, &rho);
  const double mu_s = s_col_mus(J, imus);

  if (tid < kTheta) {
    const double dtheta = kPi / kTheta;
    const double theta = (tid + 0.5) * dtheta;
    const double ct = cos(theta), st = sin(theta);
    const bool g = hits_ground(J, r, ct);
    const Taps ty = axis_taps(s_mu_texel(J, r, ct, g), nmu);
    s_y0[tid] = ty.i0;
    s_y1[tid] = ty.i1;
    s_wy[tid] = ty.w;
    const double dw = dtheta * (kPi / kTheta) * st;
    s_dw[tid] = static_cast<float>(dw);
    SdTheta& t = th[tid];
    t.cos_t = static_cast<float>(ct);
    t.sin_t = static_cast<float>(st);
    t.ground = g ? 1 : 0;
    if (g) {
      const double dg = dist_bottom(J, r, ct);
      const float3 tg = transmittance(J, J.T, r, ct, dg, true);
      const double len = sqrt(r * r + 2.0 * r * dg * ct + dg * dg);
      t.ga = static_cast<float>(r * mu_s / len);
      t.gb = static_cast<float>(dg / len);
      t.gf[0] = static_cast<float>(tg.x * J.albedo[0] / kPi * dw);
      t.gf[1] = static_cast<float>(tg.y * J.albedo[1] / kPi * dw);
      t.gf[2] = static_cast<float>(tg.z * J.albedo[2] / kPi * dw);
    } else {
      t.ga = t.gb = 0.0f;
      t.gf[0] = t.gf[1] = t.gf[2] = 0.0f;
    }
  }
  __syncthreads();

  // Per-theta nu tables of the order-(n-1) scattering at (r, mu_s), mu rows
  // blended, times d_omega (and the phase-function constants for order 1).
  const float kR = static_cast<float>(3.0 / (16.0 * kPi));
  const double gd = J.g;
  const float kM = static_cast<float>(3.0 / (8.0 * kPi) * (1.0 - gd * gd) / (2.0 + gd * gd));
  const int64_t layer = static_cast<int64_t>(k) * nmu;
  for (int e = tid; e < kTheta * nnu; e += kB) {
    const int l = e / nnu, n = e % nnu;
    const int64_t r0 = (layer + s_y0[l]) * W, r1 = (layer + s_y1[l]) * W;
    const int c0 = n * nmus + imus;
    const int c1 = (n + 1 < nnu ? n + 1 : n) * nmus + imus;
    const float wy = s_wy[l], wy0 = 1.0f - wy;
    const float fa = s_dw[l] * (two ? kR : 1.0f);
    const float4* src = J.dR;
    float4 a0 = gload(src, r0 + c0), a1 = gload(src, r1 + c0);
    float4 b0 = gload(src, r0 + c1), b1 = gload(src, r1 + c1);
    float3 v = make_float3((wy0 * a0.x + wy * a1.x) * fa, (wy0 * a0.y + wy * a1.y) * fa,
                           (wy0 * a0.z + wy * a1.z) * fa);
    float3 vn = make_float3((wy0 * b0.x + wy * b1.x) * fa, (wy0 * b0.y + wy * b1.y) * fa,
                            (wy0 * b0.z + wy * b1.z) * fa);
    float4* ta = tab_a + l * stride;
    put_entry(ta, n + 1, v, vn);
    if (n == 0) put_entry(ta, 0, v, v);
    if (two) {
      const float fb = s_dw[l] * kM;
      src = J.dM;
      a0 = gload(src, r0 + c0);
      a1 = gload(src, r1 + c0);
      b0 = gload(src, r0 + c1);
      b1 = gload(src, r1 + c1);
      v = make_float3((wy0 * a0.x + wy * a1.x) * fb, (wy0 * a0.y + wy * a1.y) * fb,
                      (wy0 * a0.z + wy * a1.z) * fb);
      vn = make_float3((wy0 * b0.x + wy * b1.x) * fb, (wy0 * b0.y + wy * b1.y) * fb,
                       (wy0 * b0.z + wy * b1.z) * fb);
      float4* tb = tab_b + l * stride;
      put_entry(tb, n + 1, v, vn);
      if (n == 0) put_entry(tb, 0, v, v);
    }
  }
  // Row 0 of delta irradiance (r = rb): the ground-irradiance lookup of the
  // reference at bottom_radius reads exactly this row.
  for (int e = tid; e < J.ew; e += kB) {
    const float3 v = make3(gload(J.dE, e));
    const float3 vn = make3(gload(J.dE, e + 1 < J.ew ? e + 1 : e));
    put_entry(etab, e + 1, v, vn);
    if (e == 0) put_entry(etab, 0, v, v);
  }
  __syncthreads();

  const int flat = tile * kB + tid;
  if (flat >= plane) return;
  const int ch = flat % nch, y = flat / nch;
  double mu;
  bool ground_unused;
  s_row_mu(J, y, r, rho, &mu, &ground_unused);
  // functions.glsl:1181-1185 for each nu of the chunk. Outputs past the last
  // nu slice get sx = sy = 0 (in-range table reads) and are not stored.
  const double wx = sqrt(1.0 - mu * mu);
  float fsx[kChunk], fsy[kChunk];
#pragma unroll
  for (int jn = 0; jn < kChunk; ++jn) {
    const int nui = ch * kChunk + jn;
    fsx[jn] = fsy[jn] = 0.0f;
    if (nui >= nnu) continue;
    const double nu = s_slice_nu(J, nui, mu, mu_s);
    const double sx = wx == 0.0 ? 0.0 : (nu - mu * mu_s) / wx;
    const double sy = sqrt(fmax(1.0 - sx * sx - mu_s * mu_s, 0.0));
    fsx[jn] = static_cast<float>(sx);
    fsy[jn] = static_cast<float>(sy);
  }
  const float fmus = static_cast<float>(mu_s), fmu = static_cast<float>(mu),
              fwx = static_cast<float>(wx);
  const double h = r - J.rb;
  const double dr = profile_density(J.prof_r, h);
  const double dm = profile_density(J.prof_m, h);
  float cR[3], cM[3];
  for (int c = 0; c < 3; ++c) {
    cR[c] = static_cast<float>(J.rayleigh[c] * dr * (3.0 / (16.0 * kPi)));
    cM[c] = static_cast<float>(J.mie_sca[c] * dm * (3.0 / (8.0 * kPi)) *
                               (1.0 - gd * gd) / (2.0 + gd * gd));
  }
  const float g1 = static_cast<float>(1.0 + gd * gd);
  const float mg2 = static_cast<float>(-2.0 * gd);
  const float nu_max = static_cast<float>(nnu - 1);
  const float e_max = static_cast<float>(J.ew - 1);
  float acc[kChunk][3];
#pragma unroll
  for (int jn = 0; jn < kChunk; ++jn) acc[jn][0] = acc[jn][1] = acc[jn][2] = 0.0f;
  for (int l = 0; l < kTheta; ++l) {
    const SdTheta t = th[l];
    const float A = fmus * t.cos_t, P = fmu * t.cos_t, Q = fwx * t.sin_t;
    float Bv[kChunk], Cv[kChunk];
#pragma unroll
    for (int jn = 0; jn < kChunk; ++jn) {
      Bv[jn] = fsx[jn] * t.sin_t;
      Cv[jn] = fsy[jn] * t.sin_t;
    }
    const float4* ta = tab_a + l * stride;
    const float4* tb = tab_b + l * stride;
    const bool g = __builtin_amdgcn_readfirstlane(t.ground) != 0;
    if (two) {
      if (g)
        sd_theta<kChunk, true, true>(ta, tb, etab, nu_max, e_max, A, P, Q, Bv, Cv, g1, mg2, cR, cM,
                             t, acc);
      else
        sd_theta<kChunk, true, false>(ta, tb, etab, nu_max, e_max, A, P, Q, Bv, Cv, g1, mg2, cR,
                              cM, t, acc);
    } else {
      if (g)
        sd_theta<kChunk, false, true>(ta, tb, etab, nu_max, e_max, A, P, Q, Bv, Cv, g1, mg2, cR,
                              cM, t, acc);
      else
        sd_theta<kChunk, false, false>(ta, tb, etab, nu_max, e_max, A, P, Q, Bv, Cv, g1, mg2, cR,
                               cM, t, acc);
    }
  }
#pragma unroll
  for (int jn = 0; jn < kChunk; ++jn) {
    const int nui = ch * kChunk + jn;
    if (nui < nnu) {
      const int64_t t = (layer + y) * W + nui * nmus + imus;
      if (J.half_luts)
        gstore(static_cast<uint2*>(J.dJ), t, pack_f21x3(acc[jn][0], acc[jn][1], acc[jn][2]));
      else
        gstore(static_cast<float3*>(J.dJ), t, make_float3(acc[jn][0], acc[jn][1], acc[jn][2]));
    }
  }
}

// ---------------------------------------------------------------------------
// Indirect irradiance (functions.glsl:1477-1511, :1573-1586) and the packing
// of atmosphere/model.cc:176-190. One wave per irradiance texel: the wave
// builds, for each of the 16 zenith angles, the nu table of the scattering at
// (r, mu = cos theta, mu_s) (trilinear in r, mu, mu_s), then each lane takes
// one of the 64 azimuths, sums over theta, and the wave reduces. One wave per
// block, so the tables of a texel need at most 4 * 16 * (NU + 1) float4 of
// LDS (table_lerp's layout).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_indirect_irradiance(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs, int order) {
  extern __shared__ float4 lds[];
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int lane = threadIdx.x;
  const int texel = blockIdx.x - first_block[ji];
  const int nnu = J.nnu, nmus = J.nmus, nmu = J.nmu;
  const bool one = order == 1;
  const int stride = 2 * (nnu + 1);
  const int per_table = kTheta * stride;
  float4* tab_a = lds;
  float4* tab_b = tab_a + per_table;
  const bool active = texel < J.ew * J.eh;
  double r = J.rb, mu_s = 1.0;
  if (active) e_rmus_from_texel(J, texel % J.ew, texel / J.ew, &r, &mu_s);
  const int W = nnu * nmus;
  const double dtheta = kPi / 32.0, dphi = kPi / 32.0;
  if (active) {
    const Taps tz = axis_taps(s_r_texel(J, r), J.nr);
    const Taps tx = axis_taps(s_mus_texel(J, mu_s), nmus);
    const float kR = static_cast<float>(3.0 / (16.0 * kPi));
    const double gd = J.g;
    const float kM = static_cast<float>(3.0 / (8.0 * kPi) * (1.0 - gd * gd) / (2.0 + gd * gd));
    for (int e = lane; e < kTheta * nnu; e += 64) {
      const int jt = e / nnu, n = e % nnu;
      const double theta = (jt + 0.5) * dtheta;
      const double ct = cos(theta);
      const Taps ty = axis_taps(s_mu_texel(J, r, ct, false), nmu);
      const float f = static_cast<float>(ct * dtheta * dphi * sin(theta));
      const int zs[2] = {tz.i0, tz.i1}, ys[2] = {ty.i0, ty.i1}, xs[2] = {tx.i0, tx.i1};
      const float wz[2] = {1.f - tz.w, tz.w}, wy[2] = {1.f - ty.w, ty.w},
                  wx[2] = {1.f - tx.w, tx.w};
      const int n1 = n + 1 < nnu ? n + 1 : n;
      for (int which = 0; which < (one ? 2 : 1); ++which) {
        const float4* src = which == 0 ? J.dR : J.dM;
        const float fk = f * (one ? (which == 0 ? kR : kM) : 1.0f);
        float3 v = make_float3(0.f, 0.f, 0.f), vn = make_float3(0.f, 0.f, 0.f);
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
              const float w = wz[a] * wy[b] * wx[c];
              const int64_t base = (static_cast<int64_t>(zs[a]) * nmu + ys[b]) * W + xs[c];
              const float4 p = gload(src, base + n * nmus), q = gload(src, base + n1 * nmus);
              v.x += w * p.x;
              v.y += w * p.y;
              v.z += w * p.z;
              vn.x += w * q.x;
              vn.y += w * q.y;
              vn.z += w * q.z;
            }
        v.x *= fk;
        v.y *= fk;
        v.z *= fk;
        vn.x *= fk;
        vn.y *= fk;
        vn.z *= fk;
        float4* tab = (which == 0 ? tab_a : tab_b) + jt * stride;
        put_entry(tab, n + 1, v, vn);
        if (n == 0) put_entry(tab, 0, v, v);
      }
    }
  }
  __syncthreads();
  if (!active) return;
  const double phi = (lane + 0.5) * dphi;
  const float wsx = static_cast<float>(sqrt(1.0 - mu_s * mu_s) * cos(phi));
  const float fmus = static_cast<float>(mu_s);
  const float nu_max = static_cast<float>(nnu - 1);
  const float hn = 0.5f * nu_max;
  const float hn1 = hn + 1.0f;  // table_lerp reads at t + 1
  const float g1 = static_cast<float>(1.0 + J.g * J.g);
  const float mg2 = static_cast<float>(-2.0 * J.g);
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  for (int jt = 0; jt < kTheta; ++jt) {
    const double theta = (jt + 0.5) * dtheta;
    const float ct = static_cast<float>(cos(theta)), st = static_cast<float>(sin(theta));
    const float nu = fmaf(st, wsx, ct * fmus);  // omega . omega_s
    float3 v = table_lerp(tab_a + jt * stride, fmaf(nu, hn, hn1));
    if (one) {
      const float3 vm = table_lerp(tab_b + jt * stride, fmaf(nu, hn, hn1));
      const float r1 = rsq_f(fmaf(mg2, nu, g1));
      const float q = r1 * r1 * r1;
      const float x1 = fmaf(nu, nu, 1.0f);
      a0 = fmaf(x1, fmaf(vm.x, q, v.x), a0);
      a1 = fmaf(x1, fmaf(vm.y, q, v.y), a1);
      a2 = fmaf(x1, fmaf(vm.z, q, v.z), a2);
    } else {
      a0 += v.x;
      a1 += v.y;
      a2 += v.z;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off);
    a1 += __shfl_xor(a1, off);
    a2 += __shfl_xor(a2, off);
  }
  if (lane == 0) {
    const float3 de = make_float3(a0, a1, a2);
    gstore(J.dE, texel, make_float4(de.x, de.y, de.z, 0.0f));
    const float3 le = apply_lum(J.lum, de);
    float4 e = gload(J.E, texel);
    e.x += le.x;
    e.y += le.y;
    e.z += le.z;
    gstore(J.E, texel, e);
  }
}

// ---------------------------------------------------------------------------
// Multiple scattering (functions.glsl:1285-1330, :1369-1383) and the packing
// of atmosphere/model.cc:192-208.
//
// Along one packed row (r layer, mu row) of the output only mu_s and nu
// change, so the ray-march samples' radius r_i, cosine mu_i - hence the LUT
// (r, mu) taps and weights - and the view transmittance are the same for the
// whole row and are computed once per row (FP64). The 16-tap lookup of the
// scattering density at sample i is then bilinear in (r, mu) with weights
// shared by the row, followed by bilinear in (mu_s, nu) per texel. The
// kernel below blends, for each sample, the four (r, mu) source rows into
// one row in LDS with coalesced loads (every texel of a row reads the same
// four rows), and each texel then reads its 4 (mu_s, nu) taps from LDS:
// 4 coalesced global loads per texel and sample instead of 16 gathers.
// Each of the kB threads holds kTpt texels; a block takes
// ms_rows_per_block(W, kTpt * kB) whole rows (several short rows, or one row
// of up to kTpt * kB texels). Rows longer than kMsMaxRow take the per-texel 16-tap gather of
// k_multiple_scattering_gather. Both kernels are launched over the same
// blocks and each returns at once on the other's jobs, so the path of a job
// depends on its own row length only.
// ---------------------------------------------------------------------------
// The mu_s texel coordinate of the scattering LUT (functions.glsl:814-827)
// in FP32, clamped to [0, MU_S - 1]: with d the distance to the top boundary
// from the ground at mu_s, a = (d - d_min) / (d_max - d_min) and
// x = max(1 - a / A, 0) / (1 + a). d = N / D, with (N, D) = (H^2, sq + t)
// for t = r_b mu_s > 0 (cancellation-free) and (sq - t, 1) otherwise, so
// multiplying the fraction by D > 0 gives x = max(D - u / A, 0) / (D + u)
// with u = a D = (N - d_min D) / (d_max - d_min): one reciprocal instead of
// two.
struct MusMap {
  float rb, H2, dmin, inv_range, invA, xmax;
  __device__ explicit MusMap(const Job& J)
      : rb(static_cast<float>(J.rb)),
        H2(static_cast<float>(J.H * J.H)),
        dmin(static_cast<float>(J.rt - J.rb)),
        inv_range(static_cast<float>(1.0 / (J.H - (J.rt - J.rb)))),
        invA(static_cast<float>(1.0 / J.A)),
        xmax(static_cast<float>(J.nmus - 1)) {}
  __device__ __forceinline__ float texel(float mus) const {
    const float t = rb * mus;
    const float sq = sqrt_f(fmaf(t, t, H2));
    const bool up = t > 0.0f;
    const float N = up ? H2 : sq - t;
    const float D = up ? sq + t : 1.0f;
    const float u = fmaf(-dmin, D, N) * inv_range;
    const float x = fmaxf(fmaf(-u, invA, D), 0.0f) * rcp_f(D + u);
    return clamp_index_coord(x * xmax, xmax);
  }
};

// Packed row p of a job's block order: r-major (p = z * MU + y, the LUT's
// own order), mu-major (p = y * R + z), or ray order (the model's table
// ms_rows of (z << 16) | y, rows sorted by the impact parameter of their
// view ray). The texels and their values are the same in every order; the
// order only changes which source rows of the scattering density the blocks
// that run together read.
// Returns (z, y). One division, and selects of values: writing the two
// results through pointers in two branches compiles to stores through a
// selected address, which puts them in scratch memory.
__device__ inline int2 ms_row(const Job& J, int p) {
  if (J.ms_row_order == 2) {
    const int v = gload(J.ms_rows, p);
    return make_int2(v >> 16, v & 0xffff);
  }
  const bool mm = J.ms_row_order != 0;
  const int d = mm ? J.nr : J.nmu;
  const int q = p / d, rem = p - q * d;
  return mm ? make_int2(rem, q) : make_int2(q, rem);
}

struct MsRowSample {
  double d;      // distance along the view ray
  double inv_r;  // 1 / r_i
  // Byte offsets of the (r, mu) source rows in delta J (12-byte RGB float
  // or 8-byte 3 x 21-bit texels; a LUT of at most 2^28 texels spans less
  // than 2^32 bytes).
  uint32_t o00, o01, o10, o11;
  float w00, w01, w10, w11;
  float f[3];    // trapezoid weight * dx * view transmittance
  float pad;
};

// kF16: the kernel serves the jobs whose delta J is 3 x 21-bit floats
// (FP16-storage models); otherwise the FP32 ones. Both are launched over the
// same blocks when a launch holds both kinds; each returns at once on the
// other's jobs.
// One texel per thread (rows up to 256 texels): eight waves per SIMD (64
// VGPRs, 32 bytes of spills) measured 2% faster than six at the reference
// sizes; the two-texel shapes keep the compiler's allocation.
template <int kTpt, bool kF16, int kB>
__global__ __launch_bounds__(kB, kTpt == 1 ? 8 : 1) void k_multiple_scattering(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs, int max_rows) {
  using Src = typename std::conditional<kF16, uint2, float3>::type;
  // Dynamic LDS: two buffers of kB * kTpt blended texels, then the
  // samples of max_rows rows (the launch's largest rows per block).
  // Blended texels are RGB (12 bytes), so that more blocks fit a CU's LDS.
  extern __shared__ float4 ms_lds[];
  float3* const buf = reinterpret_cast<float3*>(ms_lds);
  MsRowSample(*const smp)[kMarch] = reinterpret_cast<MsRowSample(*)[kMarch]>(
      reinterpret_cast<char*>(ms_lds) + 2 * kB * kTpt * sizeof(float3));
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  if ((J.half_luts != 0) != kF16) return;
  const int lb = blockIdx.x - first_block[ji];
  const int nnu = J.nnu, nmus = J.nmus, nmu = J.nmu;
  const int W = nnu * nmus;
  const int rpb = ms_rows_per_block(W, kTpt * kB);
  // Gather-path job (uniform per block). The launcher groups jobs by tile
  // and sizes max_rows over them; the tests only keep LDS indexing in
  // bounds should it not.
  if (rpb <= 0 || rpb > max_rows) return;
  const int nrows = J.nr * nmu;
  const int row0 = xcd_contiguous(lb, (nrows + rpb - 1) / rpb) * rpb;
  const int tid = threadIdx.x;

  // Per-row sample geometry: thread t < rpb * kMarch sets up sample
  // t % kMarch of row slot t / kMarch.
  for (int e = tid; e < rpb * kMarch; e += kB) {
    const int slot = e / kMarch, i = e % kMarch;
    const int row = row0 + slot;
    if (row >= nrows) continue;
    const int2 zy = ms_row(J, row);
    const int z = zy.x, y = zy.y;
    double r, rho, mu;
    bool ground;
    s_layer_r(J, z, &r, &rho);
    s_row_mu(J, y, r, rho, &mu, &ground);
    const double dx = dist_nearest(J, r, mu, ground) / (kMarch - 1);
    const double d = i * dx;
    const double r_i = clamp_radius(J, sqrt(d * d + 2.0 * r * mu * d + r * r));
    const double mu_i = clamp_cos((r * mu + d) / r_i);
    const Taps tz = axis_taps(s_r_texel(J, r_i), J.nr);
    const Taps ty = axis_taps(s_mu_texel(J, r_i, mu_i, ground), nmu);
    const float3 tv = transmittance(J, J.T, r, mu, d, ground);
    const float w = static_cast<float>(((i == 0 || i == kMarch - 1) ? 0.5 : 1.0) * dx);
    MsRowSample& S = smp[slot][i];
    S.d = d;
    S.inv_r = 1.0 / r_i;
    const uint32_t row_bytes = static_cast<uint32_t>(W) * sizeof(Src);
    S.o00 = static_cast<uint32_t>(tz.i0 * nmu + ty.i0) * row_bytes;
    S.o01 = static_cast<uint32_t>(tz.i0 * nmu + ty.i1) * row_bytes;
    S.o10 = static_cast<uint32_t>(tz.i1 * nmu + ty.i0) * row_bytes;
    S.o11 = static_cast<uint32_t>(tz.i1 * nmu + ty.i1) * row_bytes;
    const float wz1 = tz.w, wz0 = 1.0f - tz.w, wy1 = ty.w, wy0 = 1.0f - ty.w;
    S.w00 = wz0 * wy0;
    S.w01 = wz0 * wy1;
    S.w10 = wz1 * wy0;
    S.w11 = wz1 * wy1;
    S.f[0] = w * tv.x;
    S.f[1] = w * tv.y;
    S.f[2] = w * tv.z;
  }

  // Per-texel constants: the texel's nu taps (fixed for all samples,
  // functions.glsl:967-973) and r mu_s.
  // c0_of, c1_of: LDS byte offsets of the texel's two nu slices in a
  // blended row (the slice's first texel).
  int slot_of[kTpt], x_of[kTpt];
  uint32_t c0_of[kTpt], c1_of[kTpt];
  float ws_of[kTpt];
  double nu_of[kTpt], rmus_of[kTpt];
  float acc[kTpt][3];
#pragma unroll
  for (int j = 0; j < kTpt; ++j) {
    const int f = tid + j * kB;
    const int slot = f / W;
    slot_of[j] = (slot < rpb && row0 + slot < nrows) ? slot : -1;
    x_of[j] = f - slot * W;
    acc[j][0] = acc[j][1] = acc[j][2] = 0.0f;
    c0_of[j] = c1_of[j] = 0u;
    ws_of[j] = 0.0f;
    nu_of[j] = rmus_of[j] = 0.0;
    if (slot_of[j] < 0) continue;
    const int2 zy = ms_row(J, row0 + slot);
    const int z = zy.x, y = zy.y;
    double r, rho, mu;
    bool ground;
    s_layer_r(J, z, &r, &rho);
    s_row_mu(J, y, r, rho, &mu, &ground);
    const int x = x_of[j];
    const double mu_s = s_col_mus(J, x % nmus);
    const double nu = s_slice_nu(J, x / nmus, mu, mu_s);
    double tcx = (nu + 1.0) / 2.0 * (nnu - 1);
    if (!(tcx >= -1.0)) tcx = -1.0;
    if (tcx > nnu) tcx = nnu;
    const double fl = floor(tcx);
    int s0 = static_cast<int>(fl);
    int s1 = s0 + 1;
    s0 = s0 < 0 ? 0 : (s0 > nnu - 1 ? nnu - 1 : s0);
    s1 = s1 < 0 ? 0 : (s1 > nnu - 1 ? nnu - 1 : s1);
    ws_of[j] = static_cast<float>(tcx - fl);
    c0_of[j] = static_cast<uint32_t>(slot * W + s0 * nmus) * sizeof(float3);
    c1_of[j] = static_cast<uint32_t>(slot * W + s1 * nmus) * sizeof(float3);
    nu_of[j] = nu;
    rmus_of[j] = r * mu_s;
  }
  __syncthreads();

  const MusMap mus_map(J);
  const Src* __restrict__ D = static_cast<const Src*>(J.dJ);
  uint32_t x_bytes[kTpt];
#pragma unroll
  for (int j = 0; j < kTpt; ++j)
    x_bytes[j] = static_cast<uint32_t>(x_of[j]) * static_cast<uint32_t>(sizeof(Src));
  // Source texels of the next sample, loaded one sample ahead so that the
  // loads are in flight during the barrier and the gather of the current one.
  Src src[kTpt][4];
  auto fetch = [&](int i) {
#pragma unroll
    for (int j = 0; j < kTpt; ++j) {
      // Idle lanes (slot -1) load row slot 0's texel 0, in bounds, so that
      // the loads need no branch.
      const bool on = slot_of[j] >= 0;
      const MsRowSample& S = smp[on ? slot_of[j] : 0][i];
      const uint32_t xb = on ? x_bytes[j] : 0u;
      src[j][0] = gload_at(D, S.o00 + xb);
      src[j][1] = gload_at(D, S.o01 + xb);
      src[j][2] = gload_at(D, S.o10 + xb);
      src[j][3] = gload_at(D, S.o11 + xb);
    }
  };
  fetch(0);
  for (int i = 0; i < kMarch; ++i) {
    float3* B = buf + (i & 1) * (kB * kTpt);
    // Blend the four (r, mu) source rows of sample i into LDS.
#pragma unroll
    for (int j = 0; j < kTpt; ++j) {
      if (slot_of[j] < 0) continue;
      const MsRowSample& S = smp[slot_of[j]][i];
      float3 a, b, c, e;
      const float w00 = S.w00, w01 = S.w01, w10 = S.w10, w11 = S.w11;
      if (kF16) {
        const uint2* u = reinterpret_cast<const uint2*>(src[j]);
        a = unpack_f21x3(u[0]);
        b = unpack_f21x3(u[1]);
        c = unpack_f21x3(u[2]);
        e = unpack_f21x3(u[3]);
      } else {
        const float3* f = reinterpret_cast<const float3*>(src[j]);
        a = f[0];
        b = f[1];
        c = f[2];
        e = f[3];
      }
      float3 v;
      v.x = fmaf(w11, e.x, fmaf(w10, c.x, fmaf(w01, b.x, w00 * a.x)));
      v.y = fmaf(w11, e.y, fmaf(w10, c.y, fmaf(w01, b.y, w00 * a.y)));
      v.z = fmaf(w11, e.z, fmaf(w10, c.z, fmaf(w01, b.z, w00 * a.z)));
      B[tid + j * kB] = v;
    }
    if (i + 1 < kMarch) fetch(i + 1);
    // One barrier per sample: buffer i & 1 is written again at sample i + 2,
    // after every thread has passed the barrier of sample i + 1 and so has
    // finished reading it.
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTpt; ++j) {
      if (slot_of[j] < 0) continue;
      const MsRowSample& S = smp[slot_of[j]][i];
      // mu_s at the sample, (r mu_s + d nu) / r_i (functions.glsl:1311).
      const float mus =
          clampf(static_cast<float>(fma(S.d, nu_of[j], rmus_of[j]) * S.inv_r), -1.0f, 1.0f);
      const float xin = mus_map.texel(mus);
      float wx;
      const int i0 = split_coord(xin, &wx);
      const int i1 = min(i0 + 1, nmus - 1);
      // Byte offsets of the mu_s taps: 24-bit multiplies (i0, i1 < MU_S).
      const uint32_t b0 = __umul24(static_cast<uint32_t>(i0), sizeof(float3));
      const uint32_t b1 = __umul24(static_cast<uint32_t>(i1), sizeof(float3));
      const char* const Bc = reinterpret_cast<const char*>(B);
      const float3 p = *reinterpret_cast<const float3*>(Bc + c0_of[j] + b0);
      const float3 q = *reinterpret_cast<const float3*>(Bc + c0_of[j] + b1);
      const float3 u = *reinterpret_cast<const float3*>(Bc + c1_of[j] + b0);
      const float3 v = *reinterpret_cast<const float3*>(Bc + c1_of[j] + b1);
      const float ws = ws_of[j];
      const float lo0 = fmaf(wx, q.x - p.x, p.x), hi0 = fmaf(wx, v.x - u.x, u.x);
      const float lo1 = fmaf(wx, q.y - p.y, p.y), hi1 = fmaf(wx, v.y - u.y, u.y);
      const float lo2 = fmaf(wx, q.z - p.z, p.z), hi2 = fmaf(wx, v.z - u.z, u.z);
      acc[j][0] = fmaf(fmaf(ws, hi0 - lo0, lo0), S.f[0], acc[j][0]);
      acc[j][1] = fmaf(fmaf(ws, hi1 - lo1, lo1), S.f[1], acc[j][1]);
      acc[j][2] = fmaf(fmaf(ws, hi2 - lo2, lo2), S.f[2], acc[j][2]);
    }
  }

#pragma unroll
  for (int j = 0; j < kTpt; ++j) {
    if (slot_of[j] < 0) continue;
    const int2 zy = ms_row(J, row0 + slot_of[j]);
    const int z = zy.x, y = zy.y;
    const int64_t t_idx = (static_cast<int64_t>(z) * nmu + y) * W + x_of[j];
    gstore(J.dR, t_idx, make_float4(acc[j][0], acc[j][1], acc[j][2], 0.0f));
    const double nu = nu_of[j];
    if (J.acc64) {
      const double k = 1.0 / ((3.0 / (16.0 * kPi)) * (1.0 + nu * nu));
      const D4 l = apply_lum_d(J.lum, make_float3(acc[j][0], acc[j][1], acc[j][2]));
      D4 sv = load_d4(J.S64, t_idx);
      sv.x += l.x * k;
      sv.y += l.y * k;
      sv.z += l.z * k;
      store_d4(J.S64, t_idx, sv);
      continue;
    }
    const float3 l = apply_lum(J.lum, make_float3(acc[j][0], acc[j][1], acc[j][2]));
    const float inv_pr = static_cast<float>(1.0 / ((3.0 / (16.0 * kPi)) * (1.0 + nu * nu)));
    float4 sv = gload(J.S, t_idx);
    sv.x += l.x * inv_pr;
    sv.y += l.y * inv_pr;
    sv.z += l.z * inv_pr;
    gstore(J.S, t_idx, sv);
  }
}

// Rows longer than kMsMaxRow: blocks of 256 texels of one row; per texel
// and sample a 16-tap gather of the scattering density.
struct MsSample {
  double d;
  double inv_r;
  int z0, z1, y0, y1;
  float wz, wy;
  float f[3];  // trapezoid weight * dx * view transmittance
};

__global__ __launch_bounds__(kBlock) void k_multiple_scattering_gather(
    const Job* __restrict__ jobs, const int* __restrict__ first_block,
    int njobs) {
  __shared__ MsSample smp[kMarch];
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int lb = blockIdx.x - first_block[ji];
  const int nnu = J.nnu, nmus = J.nmus, nmu = J.nmu;
  const int W = nnu * nmus;
  if (W <= kMsMaxRow) return;  // LDS-path job
  const int tiles = (W + kBlock - 1) / kBlock;
  const int tile = lb % tiles;
  const int2 zy = ms_row(J, lb / tiles);
  const int z = zy.x, y = zy.y;
  const int tid = threadIdx.x;

  double r, rho, mu;
  bool ground;
  s_layer_r(J, z, &r, &rho);
  s_row_mu(J, y, r, rho, &mu, &ground);
  const double dx = dist_nearest(J, r, mu, ground) / (kMarch - 1);
  if (tid < kMarch) {
    MsSample& S = smp[tid];
    const double d = tid * dx;
    const double r_i = clamp_radius(J, sqrt(d * d + 2.0 * r * mu * d + r * r));
    const double mu_i = clamp_cos((r * mu + d) / r_i);
    const Taps tz = axis_taps(s_r_texel(J, r_i), J.nr);
    const Taps ty = axis_taps(s_mu_texel(J, r_i, mu_i, ground), nmu);
    const float3 tv = transmittance(J, J.T, r, mu, d, ground);
    const float w = static_cast<float>(((tid == 0 || tid == kMarch - 1) ? 0.5 : 1.0) * dx);
    S.d = d;
    S.inv_r = 1.0 / r_i;
    S.z0 = tz.i0;
    S.z1 = tz.i1;
    S.wz = tz.w;
    S.y0 = ty.i0;
    S.y1 = ty.i1;
    S.wy = ty.w;
    S.f[0] = w * tv.x;
    S.f[1] = w * tv.y;
    S.f[2] = w * tv.z;
  }
  __syncthreads();

  const int x = tile * kBlock + tid;
  if (x >= W) return;
  const int nui = x / nmus;
  const double mu_s = s_col_mus(J, x % nmus);
  const double nu = s_slice_nu(J, nui, mu, mu_s);
  // nu slice taps (functions.glsl:967-973).
  double tcx = (nu + 1.0) / 2.0 * (nnu - 1);
  if (!(tcx >= -1.0)) tcx = -1.0;
  if (tcx > nnu) tcx = nnu;
  const double fl = floor(tcx);
  const float ws = static_cast<float>(tcx - fl);
  int s0 = static_cast<int>(fl);
  int s1 = s0 + 1;
  s0 = s0 < 0 ? 0 : (s0 > nnu - 1 ? nnu - 1 : s0);
  s1 = s1 < 0 ? 0 : (s1 > nnu - 1 ? nnu - 1 : s1);
  const int cs0 = s0 * nmus, cs1 = s1 * nmus;
  const double rmus = r * mu_s;
  const MusMap mus_map(J);
  const void* const D = J.dJ;
  const bool f16 = J.half_luts != 0;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  for (int s = 0; s < kMarch; ++s) {
    const MsSample& S = smp[s];
    const float mus = clampf(static_cast<float>(fma(S.d, nu, rmus) * S.inv_r), -1.0f, 1.0f);
    const float xin = mus_map.texel(mus);
    float wx;
    const int i0 = split_coord(xin, &wx);
    const int i1 = min(i0 + 1, nmus - 1);
    const float wz1 = S.wz, wz0 = 1.0f - S.wz, wy1 = S.wy, wy0 = 1.0f - S.wy;
    const float wrow[4] = {wz0 * wy0, wz0 * wy1, wz1 * wy0, wz1 * wy1};
    const int64_t rows[4] = {(static_cast<int64_t>(S.z0) * nmu + S.y0) * W,
                             (static_cast<int64_t>(S.z0) * nmu + S.y1) * W,
                             (static_cast<int64_t>(S.z1) * nmu + S.y0) * W,
                             (static_cast<int64_t>(S.z1) * nmu + S.y1) * W};
    const float wc[4] = {(1.0f - ws) * (1.0f - wx), (1.0f - ws) * wx, ws * (1.0f - wx),
                         ws * wx};
    const int cols[4] = {cs0 + i0, cs0 + i1, cs1 + i0, cs1 + i1};
    float j0 = 0.f, j1 = 0.f, j2 = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float r0 = 0.f, r1 = 0.f, r2 = 0.f;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int64_t t = rows[q] + cols[c];
        const float3 v = f16 ? unpack_f21x3(gload(static_cast<const uint2*>(D), t))
                             : gload(static_cast<const float3*>(D), t);
        r0 = fmaf(wc[c], v.x, r0);
        r1 = fmaf(wc[c], v.y, r1);
        r2 = fmaf(wc[c], v.z, r2);
      }
      j0 = fmaf(wrow[q], r0, j0);
      j1 = fmaf(wrow[q], r1, j1);
      j2 = fmaf(wrow[q], r2, j2);
    }
    a0 = fmaf(j0, S.f[0], a0);
    a1 = fmaf(j1, S.f[1], a1);
    a2 = fmaf(j2, S.f[2], a2);
  }
  const int64_t t_idx = (static_cast<int64_t>(z) * nmu + y) * W + x;
  gstore(J.dR, t_idx, make_float4(a0, a1, a2, 0.0f));
  if (J.acc64) {
    const double k = 1.0 / ((3.0 / (16.0 * kPi)) * (1.0 + nu * nu));
    const D4 l = apply_lum_d(J.lum, make_float3(a0, a1, a2));
    D4 sv = load_d4(J.S64, t_idx);
    sv.x += l.x * k;
    sv.y += l.y * k;
    sv.z += l.z * k;
    store_d4(J.S64, t_idx, sv);
    return;
  }
  const float3 l = apply_lum(J.lum, make_float3(a0, a1, a2));
  const float inv_pr = static_cast<float>(1.0 / ((3.0 / (16.0 * kPi)) * (1.0 + nu * nu)));
  float4 sv = gload(J.S, t_idx);
  sv.x += l.x * inv_pr;
  sv.y += l.y * inv_pr;
  sv.z += l.z * inv_pr;
  gstore(J.S, t_idx, sv);
}

// ---------------------------------------------------------------------------
// Illuminance mode: sum the iterations' FP64 accumulators into S and SM
// (FP32; in FP16 mode these are the accumulators that k_to_half then
// converts) and their irradiance accumulators into E.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_acc64_to_f32(const Job* __restrict__ jobs,
                                                         const int* __restrict__ first_block,
                                                         int njobs) {
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int64_t n = static_cast<int64_t>(J.nr) * J.nmu * J.nmus * J.nnu;
  const int64_t ne = static_cast<int64_t>(J.ew) * J.eh;
  const int64_t t = static_cast<int64_t>(blockIdx.x - first_block[ji]) * kBlock + threadIdx.x;
  // The iterations' accumulators are summed in iteration order, in FP64.
  if (t < n) {
    D4 s{0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < J.niter; ++k) s = add_d4(s, load_d4(J.S64, k * n + t));
    gstore(J.S, t, make_float4(static_cast<float>(s.x), static_cast<float>(s.y),
                               static_cast<float>(s.z), static_cast<float>(s.w)));
    if (!J.combined) {
      D4 m{0.0, 0.0, 0.0, 0.0};
      for (int k = 0; k < J.niter; ++k) m = add_d4(m, load_d4(J.SM64, k * n + t));
      gstore(J.SM, t, make_float4(static_cast<float>(m.x), static_cast<float>(m.y),
                                  static_cast<float>(m.z), 0.0f));
    }
  }
  if (t < ne) {
    double e[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < J.niter; ++k) {
      const float4 v = gload(J.E_it, k * ne + t);
      e[0] += v.x;
      e[1] += v.y;
      e[2] += v.z;
    }
    gstore(J.E, t, make_float4(static_cast<float>(e[0]), static_cast<float>(e[1]),
                               static_cast<float>(e[2]), 0.0f));
  }
}

// ---------------------------------------------------------------------------
// The lookup copy of S and SM (ATMO_OPTION_INTERLEAVED_LOOKUPS), in 128-byte
// blocks of quads (lookup_copy_block): one thread per 16-byte element, so
// that each store instruction writes whole lines. An element is a float4 of
// S or SM (FP32 models: 8 per quad) or a texel's S and SM as 4 FP16 each
// (FP16 models: 4 per quad, 2 quads per block).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_interleave(const Job* __restrict__ jobs,
                                                       const int* __restrict__ first_block,
                                                       int njobs) {
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const bool half = J.half_luts != 0;
  const int64_t n = 8 * lookup_copy_blocks(J.nr, J.nmu, J.nmus, J.nnu, half);
  const int64_t u = static_cast<int64_t>(blockIdx.x - first_block[ji]) * kBlock + threadIdx.x;
  if (u >= n || J.Q == nullptr || J.combined) return;
  // Block qd = ((z * rows + ry) * (NU + 1) + s + 1) * (MU_S + 1) + i + 1.
  const int64_t qd = u >> 3;
  const int part = static_cast<int>(u & 7);
  const int ia = static_cast<int>(qd % (J.nmus + 1)) - 1;
  const int64_t rest = qd / (J.nmus + 1);
  const int sa = static_cast<int>(rest % (J.nnu + 1)) - 1;
  const int64_t zr = rest / (J.nnu + 1);
  const int rows = lookup_copy_rows(J.nmu, half);
  const int z = static_cast<int>(zr / rows), ry = static_cast<int>(zr % rows);
  // Tap k = 2 e + q of the quad: (s + q, i + e).
  const int k = half ? (part & 3) : (part >> 1);
  const int y = half ? min(max(ry - 1 + (part >> 2), 0), J.nmu - 1) : ry;
  const int i = min(max(ia + (k >> 1), 0), J.nmus - 1);
  const int sl = min(max(sa + (k & 1), 0), J.nnu - 1);
  const int64_t t = (static_cast<int64_t>(z) * J.nmu + y) * (J.nnu * J.nmus) + sl * J.nmus + i;
  if (half) {
    const uint2 a = gload(J.S16, t), b = gload(J.SM16, t);
    gstore(static_cast<uint4*>(J.Q), u, make_uint4(a.x, a.y, b.x, b.y));
  } else {
    gstore(static_cast<float4*>(J.Q), u, (part & 1) ? gload(J.SM, t) : gload(J.S, t));
  }
}

// ---------------------------------------------------------------------------
// FP16 storage of the scattering LUTs (half_precision): per-channel maximum
// of the FP32 accumulators, then conversion with power-of-two scales that map
// each channel's maximum into [2^13, 2^14), well inside the FP16 range, so
// that FP16 keeps its 11-bit relative precision down to ~4e-9 of the maximum.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_clear_max(const Job* __restrict__ jobs,
                                                  const int* __restrict__ first_block,
                                                  int njobs) {
  const int ji = find_job(first_block, njobs, blockIdx.x);
  if (threadIdx.x < 8) jobs[ji].maxbits[threadIdx.x] = 0u;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}

// Launched over blocks_half_conversion blocks per job (one per 256 texels,
// the grid of k_to_half); block b reduces texels [b * kMaxSpan, (b + 1) *
// kMaxSpan) and the rest exit at once. Each thread keeps 64 texels' maxima,
// the block combines its waves in LDS and one lane per channel updates the
// job's maxima: 7 atomics per 16K texels.
constexpr int kMaxPerThread = 64;
constexpr int kMaxSpan = kBlock * kMaxPerThread;
__global__ __launch_bounds__(kBlock) void k_lut_max(const Job* __restrict__ jobs,
                                                    const int* __restrict__ first_block,
                                                    int njobs) {
  __shared__ float part[kBlock / 64][8];
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int64_t n = static_cast<int64_t>(J.nr) * J.nmu * J.nmus * J.nnu;
  const int64_t base = static_cast<int64_t>(blockIdx.x - first_block[ji]) * kMaxSpan;
  if (base >= n) return;  // uniform per block
  // |value| as non-negative float bits compare like unsigned integers.
  float v[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < kMaxPerThread; ++k) {
    const int64_t t = base + k * kBlock + threadIdx.x;
    if (t >= n) break;
    const float4 s = gload(J.S, t);
    v[0] = fmaxf(v[0], fabsf(s.x));
    v[1] = fmaxf(v[1], fabsf(s.y));
    v[2] = fmaxf(v[2], fabsf(s.z));
    v[3] = fmaxf(v[3], fabsf(s.w));
    if (!J.combined) {
      const float4 m = gload(J.SM, t);
      v[4] = fmaxf(v[4], fabsf(m.x));
      v[5] = fmaxf(v[5], fabsf(m.y));
      v[6] = fmaxf(v[6], fabsf(m.z));
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    const float w = wave_max(v[c]);
    if (lane == 0) part[wave][c] = w;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    float w = 0.0f;
    for (int q = 0; q < kBlock / 64; ++q) w = fmaxf(w, part[q][threadIdx.x]);
    if (w > 0.0f) atomicMax(&J.maxbits[threadIdx.x], __float_as_uint(w));
  }
}

// The scale maps the channel maximum mx into [2^13, 2^14). The exponent is
// capped at 126, so that channels whose maximum is tiny (below 2^-112,
// denormals included) keep a finite scale whose inverse is a normal float.
__device__ __forceinline__ float pow2_scale(unsigned bits) {
  const float mx = __uint_as_float(bits);
  if (!(mx > 0.0f) || isinf(mx)) return 1.0f;
  int e;
  frexpf(mx, &e);  // mx = f * 2^e, f in [0.5, 1)
  return ldexpf(1.0f, min(14 - e, 126));
}

__global__ __launch_bounds__(kBlock) void k_to_half(const Job* __restrict__ jobs,
                                                    const int* __restrict__ first_block,
                                                    int njobs) {
  const int ji = find_job(first_block, njobs, blockIdx.x);
  const Job& J = jobs[ji];
  const int lb = blockIdx.x - first_block[ji];
  float sc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) sc[c] = pow2_scale(J.maxbits[c]);
  if (lb == 0 && threadIdx.x < 8) J.scales[threadIdx.x] = sc[threadIdx.x];
  const int64_t n = static_cast<int64_t>(J.nr) * J.nmu * J.nmus * J.nnu;
  const int64_t t = static_cast<int64_t>(lb) * kBlock + threadIdx.x;
  if (t >= n) return;
  const float4 s = gload(J.S, t);
  gstore(J.S16, t, pack_half4(make_float4(s.x * sc[0], s.y * sc[1], s.z * sc[2], s.w * sc[3])));
  if (!J.combined) {
    const float4 m = gload(J.SM, t);
    gstore(J.SM16, t, pack_half4(make_float4(m.x * sc[4], m.y * sc[5], m.z * sc[6], 0.0f)));
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch geometry and launchers.
// ---------------------------------------------------------------------------

int blocks_optical_depth(const Job& j) {
  return (j.tw * j.th + kBlock / 64 - 1) / (kBlock / 64);
}
int blocks_transmittance(const Job& j) {
  if (j.OD != nullptr) return (j.tw * j.th + kBlock - 1) / kBlock;
  return blocks_optical_depth(j);
}
int blocks_direct_irradiance(const Job& j) {
  return (j.ew * j.eh + kBlock - 1) / kBlock;
}
int blocks_single_scattering(const Job& j) {
  const int w = j.nnu * j.nmus;
  const int per_block = kBlock * ss_texels_per_thread(w);
  return ((w + per_block - 1) / per_block) * j.nmu * j.nr;
}
int blocks_scattering_density(const Job& j) {
  const SdTile t = kSdTile[j.sd_tile];
  const int plane = j.nmu * ((j.nnu + t.chunk - 1) / t.chunk);
  return ((plane + t.block - 1) / t.block) * j.nmus * j.nr;
}
int blocks_indirect_irradiance(const Job& j) { return j.ew * j.eh; }
int blocks_multiple_scattering(const Job& j) {
  const int w = j.nnu * j.nmus;
  const int rpb = ms_rows_per_block(w, ms_tile_capacity(j.ms_tile));
  const int rows = j.nmu * j.nr;
  if (rpb == 0) return ((w + kBlock - 1) / kBlock) * rows;
  if (rpb < 0) return 0;  // not a tile of this row length (rejected by the host)
  return (rows + rpb - 1) / rpb;
}
int blocks_combine(const Job& j) {
  const int64_t n = static_cast<int64_t>(j.nr) * j.nmu * j.nmus * j.nnu;
  const int64_t ne = static_cast<int64_t>(j.ew) * j.eh;
  return static_cast<int>(((n > ne ? n : ne) + kBlock - 1) / kBlock);
}
int blocks_lookup_copy(const Job& j) {
  return static_cast<int>(
      (8 * lookup_copy_blocks(j.nr, j.nmu, j.nmus, j.nnu, j.half_luts != 0) + kBlock - 1) /
      kBlock);
}
int blocks_half_conversion(const Job& j) {
  const int64_t n = static_cast<int64_t>(j.nr) * j.nmu * j.nmus * j.nnu;
  return static_cast<int>((n + kBlock - 1) / kBlock);
}

hipError_t launch_optical_depth(const PassLaunch& p, hipStream_t s) {
  hipLaunchKernelGGL(k_optical_depth, dim3(p.total_blocks), dim3(kBlock), 0, s,
                     p.jobs, p.first_block, p.njobs);
  return hipGetLastError();
}
hipError_t launch_transmittance(const PassLaunch& p, hipStream_t s) {
  hipLaunchKernelGGL(k_transmittance, dim3(p.total_blocks), dim3(kBlock), 0, s,
                     p.jobs, p.first_block, p.njobs);
  return hipGetLastError();
}
hipError_t launch_direct_irradiance(const PassLaunch& p, hipStream_t s) {
  hipLaunchKernelGGL(k_direct_irradiance, dim3(p.total_blocks), dim3(kBlock), 0, s,
                     p.jobs, p.first_block, p.njobs);
  return hipGetLastError();
}
template <int kTpt>
void launch_ss_tpt(const PassLaunch& p, hipStream_t s) {
  if (kTpt < ss_texels_per_thread(p.min_w) || kTpt > ss_texels_per_thread(p.max_w)) return;
  const dim3 g(p.total_blocks), b(kBlock);
  if (p.kinds & kJobsF32)
    hipLaunchKernelGGL((k_single_scattering<false, kTpt>), g, b, 0, s, p.jobs, p.first_block,
                       p.njobs);
  if (p.kinds & kJobsAcc64)
    hipLaunchKernelGGL((k_single_scattering<true, kTpt>), g, b, 0, s, p.jobs, p.first_block,
                       p.njobs);
}
hipError_t launch_single_scattering(const PassLaunch& p, hipStream_t s) {
  // One instance per kind of job the launch may hold (accumulator precision,
  // texels per thread from the packed row widths); each skips the others.
  launch_ss_tpt<1>(p, s);
  launch_ss_tpt<2>(p, s);
  launch_ss_tpt<4>(p, s);
  return hipGetLastError();
}
template <int kChunk, int kB>
void launch_sd_tile(const PassLaunch& p, int order, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_scattering_density<kChunk, kB>), dim3(p.total_blocks), dim3(kB), lds,
                     s, p.jobs, p.first_block, p.njobs, order);
}

hipError_t launch_scattering_density(const PassLaunch& p, int order,
                                     hipStream_t s) {
  // Tables of NU + 1 and E_W + 1 entries of two float4 (table_lerp).
  const size_t lds = sizeof(float4) *
                     (static_cast<size_t>((order == 2 ? 4 : 2) * kTheta * (p.max_nnu + 1)) +
                      2 * static_cast<size_t>(p.max_ew + 1));
  static_assert(kSdTiles == 6, "one case per tile shape");
  switch (p.tile) {
    case 0: launch_sd_tile<kSdTile[0].chunk, kSdTile[0].block>(p, order, lds, s); break;
    case 1: launch_sd_tile<kSdTile[1].chunk, kSdTile[1].block>(p, order, lds, s); break;
    case 2: launch_sd_tile<kSdTile[2].chunk, kSdTile[2].block>(p, order, lds, s); break;
    case 3: launch_sd_tile<kSdTile[3].chunk, kSdTile[3].block>(p, order, lds, s); break;
    case 4: launch_sd_tile<kSdTile[4].chunk, kSdTile[4].block>(p, order, lds, s); break;
    case 5: launch_sd_tile<kSdTile[5].chunk, kSdTile[5].block>(p, order, lds, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_indirect_irradiance(const PassLaunch& p, int order,
                                      hipStream_t s) {
  const size_t lds =
      sizeof(float4) * static_cast<size_t>((order == 1 ? 4 : 2) * kTheta * (p.max_nnu + 1));
  hipLaunchKernelGGL(k_indirect_irradiance, dim3(p.total_blocks), dim3(64), lds, s,
                     p.jobs, p.first_block, p.njobs, order);
  return hipGetLastError();
}
hipError_t launch_acc64_to_f32(const PassLaunch& p, hipStream_t s) {
  hipLaunchKernelGGL(k_acc64_to_f32, dim3(p.total_blocks), dim3(kBlock), 0, s, p.jobs,
                     p.first_block, p.njobs);
  return hipGetLastError();
}
hipError_t launch_clear_max(const PassLaunch& p, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_max, dim3(p.total_blocks), dim3(64), 0, s, p.jobs,
                     p.first_block, p.njobs);
  return hipGetLastError();
}
hipError_t launch_interleave(const PassLaunch& p, hipStream_t s) {
  hipLaunchKernelGGL(k_interleave, dim3(p.total_blocks), dim3(kBlock), 0, s, p.jobs,
                     p.first_block, p.njobs);
  return hipGetLastError();
}
hipError_t launch_to_half(const PassLaunch& p, hipStream_t s) {
  hipLaunchKernelGGL(k_lut_max, dim3(p.total_blocks), dim3(kBlock), 0, s, p.jobs,
                     p.first_block, p.njobs);
  hipLaunchKernelGGL(k_to_half, dim3(p.total_blocks), dim3(kBlock), 0, s, p.jobs,
                     p.first_block, p.njobs);
  return hipGetLastError();
}

template <int kTpt, int kB>
void launch_ms_tile(const PassLaunch& p, hipStream_t s) {
  // Rows per block shrink as rows grow: the shortest row sets the launch's
  // largest count.
  const int rows = ms_rows_per_block(p.min_w, kTpt * kB);
  const size_t lds = sizeof(float3) * 2 * kB * kTpt +
                     sizeof(MsRowSample) * kMarch * static_cast<size_t>(rows);
  if (p.dj_kinds & kDjF32)
    hipLaunchKernelGGL((k_multiple_scattering<kTpt, false, kB>), dim3(p.total_blocks),
                       dim3(kB), lds, s, p.jobs, p.first_block, p.njobs, rows);
  if (p.dj_kinds & kDjRgb21)
    hipLaunchKernelGGL((k_multiple_scattering<kTpt, true, kB>), dim3(p.total_blocks),
                       dim3(kB), lds, s, p.jobs, p.first_block, p.njobs, rows);
}

hipError_t launch_multiple_scattering(const PassLaunch& p, hipStream_t s) {
  if (p.tile < 0 || p.tile >= kMsTiles) return hipErrorInvalidValue;
  if (p.min_w <= kMsMaxRow) {
    // Every LDS-path row of the launch must fit the tile.
    if (p.max_lds_w > ms_tile_capacity(p.tile)) return hipErrorInvalidValue;
    static_assert(kMsTiles == 3, "one case per tile shape");
    switch (p.tile) {
      case 0: launch_ms_tile<kMsTile[0].tpt, kMsTile[0].block>(p, s); break;
      case 1: launch_ms_tile<kMsTile[1].tpt, kMsTile[1].block>(p, s); break;
      default: launch_ms_tile<kMsTile[2].tpt, kMsTile[2].block>(p, s); break;
    }
  }
  if (p.max_w > kMsMaxRow)
    hipLaunchKernelGGL(k_multiple_scattering_gather, dim3(p.total_blocks), dim3(kBlock), 0,
                       s, p.jobs, p.first_block, p.njobs);
  return hipGetLastError();
}

}  // namespace atmo
