// C ABI of libatmo (include/atmo.h): model objects, parameter lowering, the
// Model::Init pass sequence, batched graph-captured Init, lookups.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/atmo.h"
#include "atmo_device.h"
#include "cie_1931.h"
#include "kernels.h"
#include "runtime_internal.h"

namespace atmo_rt {

thread_local std::string g_last_error;

atmo_status fail(atmo_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

#define ATMO_HIP(call)                                                       \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(e_ == hipErrorOutOfMemory ? ATMO_ERROR_OUT_OF_MEMORY       \
                                            : ATMO_ERROR_DEVICE,             \
                  std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

constexpr double kLambdaR = 680.0;
constexpr double kLambdaG = 550.0;
constexpr double kLambdaB = 440.0;
constexpr int kLambdaMin = 360;
constexpr int kLambdaMax = 830;

// atmosphere/model.cc:535-552
double interpolate(const std::vector<double>& wl, const std::vector<double>& f,
                   double lambda) {
  if (lambda < wl[0]) return f[0];
  for (size_t i = 0; i + 1 < wl.size(); ++i) {
    if (lambda < wl[i + 1]) {
      double u = (lambda - wl[i]) / (wl[i + 1] - wl[i]);
      return f[i] * (1.0 - u) + f[i + 1] * u;
    }
  }
  return f.back();
}

// atmosphere/model.cc:521-533; column 1, 2, 3 = x_bar, y_bar, z_bar.
double cie_value(double lambda, int column) {
  if (lambda <= kLambdaMin || lambda >= kLambdaMax) return 0.0;
  double u = (lambda - kLambdaMin) / atmo_cie::kCieLambdaStep;
  int row = static_cast<int>(std::floor(u));
  u -= row;
  const double* t = column == 1 ? atmo_cie::kCieXBar
                    : column == 2 ? atmo_cie::kCieYBar
                                  : atmo_cie::kCieZBar;
  return t[row] * (1.0 - u) + t[row + 1] * u;
}

// ComputeSpectralRadianceToLuminanceFactors (atmosphere/model.cc:562-595):
// 1 nm integration of the CIE functions against the solar spectrum relative
// to its value at 680/550/440 nm, times (lambda / lambda_c)^power.
void luminance_factors(const std::vector<double>& wl, const std::vector<double>& solar,
                       double power, double* k) {
  const double lambdas[3] = {kLambdaR, kLambdaG, kLambdaB};
  double solar_c[3];
  for (int c = 0; c < 3; ++c) {
    solar_c[c] = interpolate(wl, solar, lambdas[c]);
    k[c] = 0.0;
  }
  const double* m = atmo_cie::kXyzToSrgb;
  for (int lambda = kLambdaMin; lambda < kLambdaMax; ++lambda) {
    const double x = cie_value(lambda, 1), y = cie_value(lambda, 2),
                 z = cie_value(lambda, 3);
    const double irradiance = interpolate(wl, solar, lambda);
    for (int c = 0; c < 3; ++c) {
      const double bar = m[3 * c] * x + m[3 * c + 1] * y + m[3 * c + 2] * z;
      k[c] += bar * irradiance / solar_c[c] * std::pow(lambda / lambdas[c], power);
    }
  }
  for (int c = 0; c < 3; ++c) k[c] *= atmo_cie::kMaxLuminousEfficacy;
}

int num_iterations(const Spec& s) {
  return s.num_precomputed_wavelengths <= 3
             ? 1
             : static_cast<int>((s.num_precomputed_wavelengths + 2) / 3);
}

// Wavelengths and luminance_from_radiance of iteration i
// (atmosphere/model.cc:907-943).
void iteration(const Spec& s, int i, double* lambdas, float* L) {
  if (s.num_precomputed_wavelengths <= 3) {
    lambdas[0] = kLambdaR;
    lambdas[1] = kLambdaG;
    lambdas[2] = kLambdaB;
    for (int k = 0; k < 9; ++k) L[k] = (k % 4 == 0) ? 1.0f : 0.0f;
    return;
  }
  const int iters = num_iterations(s);
  const double dlambda = (kLambdaMax - kLambdaMin) / (3.0 * iters);
  for (int k = 0; k < 3; ++k) lambdas[k] = kLambdaMin + (3 * i + k + 0.5) * dlambda;
  const double* m = atmo_cie::kXyzToSrgb;
  for (int row = 0; row < 3; ++row)
    for (int col = 0; col < 3; ++col) {
      const double l = lambdas[col];
      L[3 * row + col] = static_cast<float>(
          (m[3 * row] * cie_value(l, 1) + m[3 * row + 1] * cie_value(l, 2) +
           m[3 * row + 2] * cie_value(l, 3)) *
          dlambda);
    }
}

// atmosphere/model.cc:644-666 (layer scaling and front padding).
void lower_profile(const std::vector<atmo_density_layer>& layers, double unit,
                   atmo::LayerD* out) {
  atmo_density_layer padded[2] = {};
  const size_t n = std::min<size_t>(layers.size(), 2);
  for (size_t i = 0; i < n; ++i) padded[2 - n + i] = layers[i];
  for (int i = 0; i < 2; ++i) {
    out[i].width = padded[i].width / unit;
    out[i].exp_term = padded[i].exp_term;
    out[i].exp_scale = padded[i].exp_scale * unit;
    out[i].linear_term = padded[i].linear_term * unit;
    out[i].constant_term = padded[i].constant_term;
  }
}

// AtmosphereParameters at three wavelengths (atmosphere/model.cc:718-734),
// in exact binary values (no text round trip).
void lower(const Spec& s, const double* lambdas, atmo::Job* j) {
  const double unit = s.length_unit;
  for (int c = 0; c < 3; ++c) {
    j->solar[c] = interpolate(s.wavelengths, s.solar, lambdas[c]);
    j->rayleigh[c] = interpolate(s.wavelengths, s.rayleigh, lambdas[c]) * unit;
    j->mie_sca[c] = interpolate(s.wavelengths, s.mie_sca, lambdas[c]) * unit;
    j->mie_ext[c] = interpolate(s.wavelengths, s.mie_ext, lambdas[c]) * unit;
    j->absorption[c] = interpolate(s.wavelengths, s.absorption, lambdas[c]) * unit;
    j->albedo[c] = interpolate(s.wavelengths, s.albedo, lambdas[c]);
  }
  j->sun_angular_radius = s.sun_angular_radius;
  j->rb = s.bottom_radius / unit;
  j->rt = s.top_radius / unit;
  j->mu_s_min = std::cos(s.max_sun_zenith_angle);
  j->g = s.g;
  lower_profile(s.prof_r, unit, j->prof_r);
  lower_profile(s.prof_m, unit, j->prof_m);
  lower_profile(s.prof_a, unit, j->prof_a);
  j->H = std::sqrt(j->rt * j->rt - j->rb * j->rb);
  // functions.glsl:814-821: A from the distance to the top at mu_s_min.
  const double d_min = j->rt - j->rb, d_max = j->H;
  const double D = atmo::dist_top(*j, j->rb, j->mu_s_min);
  j->A = (D - d_min) / (d_max - d_min);
}

void set_sizes(atmo::Job* j, const Sizes& z) {
  j->tw = z.tw;
  j->th = z.th;
  j->nr = z.nr;
  j->nmu = z.nmu;
  j->nmus = z.nmus;
  j->nnu = z.nnu;
  j->ew = z.ew;
  j->eh = z.eh;
}

int pass_blocks(Pass p, const atmo::Job& j) {
  switch (p) {
    case kT: return atmo::blocks_transmittance(j);
    case kE0: return atmo::blocks_direct_irradiance(j);
    case kSS: return atmo::blocks_single_scattering(j);
    case kSD: return atmo::blocks_scattering_density(j);
    case kII: return atmo::blocks_indirect_irradiance(j);
    case kMS: return atmo::blocks_multiple_scattering(j);
    case kHalf: return atmo::blocks_half_conversion(j);
    case kAcc: return atmo::blocks_combine(j);
    case kClear: return 1;
    case kPack: return atmo::blocks_lookup_copy(j);
    case kOD: return atmo::blocks_optical_depth(j);
    default: return 0;
  }
}

}  // namespace atmo_rt

using namespace atmo_rt;

namespace atmo_rt {

namespace {

void clear_temporary_pointers(atmo_model* m) {
  m->T_it = m->E_it = nullptr;
  m->dE_it = m->dR_it = m->dM_it = nullptr;
  m->dE = m->dR = m->dM = nullptr;
  m->dJ_it = m->dJ = nullptr;
  if (m->spec.half_precision) m->S = m->SM = nullptr;
  m->S64 = m->SM64 = nullptr;
  m->OD = nullptr;
}

size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// The layout of the temporaries block; with 'base' null, only its size.
// Illuminance mode (iters > 1): per iteration a transmittance LUT, an
// irradiance accumulator, the four delta LUTs and the FP64 accumulators of S
// and SM, each array of 'iters' copies contiguous, and one array of optical
// depths. FP16 storage adds the FP32 accumulators of S and SM.
size_t carve(const Sizes& z, bool half, bool combined, int iters, char* base,
             atmo_model* m) {
  const bool par = iters > 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base == nullptr ? nullptr : base + off;
    off += round256(bytes);
    return p;
  };
  const size_t P = static_cast<size_t>(iters);
  char* t_it = par ? take(P * z.t_texels() * sizeof(float4)) : nullptr;
  char* e_it = par ? take(P * z.e_texels() * sizeof(float4)) : nullptr;
  char* de = take(P * z.e_texels() * sizeof(float4));
  char* dr = take(P * z.s_texels() * sizeof(float4));
  char* dm = take(P * z.s_texels() * sizeof(float4));
  // Delta J: RGB float, or 3 x 21-bit floats in FP16 mode.
  char* dj = take(P * z.s_texels() * (half ? sizeof(uint2) : sizeof(float3)));
  char* s32 = half ? take(z.s_texels() * sizeof(float4)) : nullptr;
  char* sm32 = half && !combined ? take(z.s_texels() * sizeof(float4)) : nullptr;
  char* s64 = par ? take(P * z.s_texels() * 4 * sizeof(double)) : nullptr;
  char* sm64 = par && !combined ? take(P * z.s_texels() * 4 * sizeof(double)) : nullptr;
  char* od = par ? take(z.t_texels() * 4 * sizeof(double)) : nullptr;
  if (m != nullptr) {
    m->T_it = reinterpret_cast<float4*>(t_it);
    m->E_it = reinterpret_cast<float4*>(e_it);
    m->dE_it = reinterpret_cast<float4*>(de);
    m->dR_it = reinterpret_cast<float4*>(dr);
    m->dM_it = reinterpret_cast<float4*>(dm);
    m->dJ_it = dj;
    // The retained intermediates: the last iteration's delta LUTs.
    m->dE = m->dE_it + (P - 1) * z.e_texels();
    m->dR = m->dR_it + (P - 1) * z.s_texels();
    m->dM = m->dM_it + (P - 1) * z.s_texels();
    m->dJ = dj == nullptr ? nullptr
                          : dj + (P - 1) * z.s_texels() * (half ? sizeof(uint2) : sizeof(float3));
    if (half) {
      m->S = reinterpret_cast<float4*>(s32);
      m->SM = reinterpret_cast<float4*>(sm32);
    }
    m->S64 = reinterpret_cast<double*>(s64);
    m->SM64 = reinterpret_cast<double*>(sm64);
    m->OD = reinterpret_cast<double*>(od);
  }
  return off;
}

}  // namespace

// Temporaries are released once the work that reads them has finished.
void free_temporaries(atmo_model* m) {
  if (m->external_temporaries) {
    drop_external_temporaries(m);
    return;
  }
  if (m->temp_block != nullptr) {
    (void)hipFree(m->temp_block);
    m->temp_block = nullptr;
  }
  clear_temporary_pointers(m);
}

atmo_status alloc_temporaries(atmo_model* m) {
  if (m->temp_block != nullptr || m->external_temporaries) return ATMO_OK;
  ATMO_HIP(hipMalloc(&m->temp_block, temporaries_bytes(m)));
  carve(m->sizes, m->spec.half_precision, m->spec.combined, m->iters, m->temp_block, m);
  return ATMO_OK;
}

size_t temporaries_bytes_for(const Sizes& z, bool half, bool combined, int iters) {
  return carve(z, half, combined, iters, nullptr, nullptr);
}

size_t temporaries_bytes(const atmo_model* m) {
  return temporaries_bytes_for(m->sizes, m->spec.half_precision, m->spec.combined,
                               m->iters);
}

void use_external_temporaries(atmo_model* m, char* base) {
  carve(m->sizes, m->spec.half_precision, m->spec.combined, m->iters, base, m);
  m->external_temporaries = true;
}

void drop_external_temporaries(atmo_model* m) {
  clear_temporary_pointers(m);
  m->external_temporaries = false;
}

// The ray order of a model's packed (r, mu) rows (Job::ms_rows): sky rows
// before ground rows, each sorted by the impact parameter r sqrt(1 - mu^2)
// of the row's view ray, then by decreasing mu; entries (z << 16) | y. It
// depends on the radii and the LUT sizes only.
std::vector<int> ms_ray_order(const atmo::Job& j) {
  struct Row {
    int ground;
    double p, mu;
    int zy;
  };
  std::vector<Row> rows;
  rows.reserve(static_cast<size_t>(j.nr) * j.nmu);
  for (int z = 0; z < j.nr; ++z) {
    double r, rho;
    atmo::s_layer_r(j, z, &r, &rho);
    for (int y = 0; y < j.nmu; ++y) {
      double mu;
      bool ground;
      atmo::s_row_mu(j, y, r, rho, &mu, &ground);
      rows.push_back({ground ? 1 : 0, r * std::sqrt(std::max(0.0, 1.0 - mu * mu)), mu,
                      (z << 16) | y});
    }
  }
  std::stable_sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) {
    if (a.ground != b.ground) return a.ground < b.ground;
    if (a.p != b.p) return a.p < b.p;
    return a.mu > b.mu;
  });
  std::vector<int> out(rows.size());
  for (size_t i = 0; i < rows.size(); ++i) out[i] = rows[i].zy;
  return out;
}

// Job k < iters is Precompute iteration k; k = -1 is a job on the resident
// LUTs (the final transmittance and accumulator sums, the lookups). In
// illuminance mode every iteration has its own transmittance, delta LUTs and
// accumulators.
void fill_job_luts(atmo_model* m, atmo::Job* j, int k) {
  const Sizes& z = m->sizes;
  // (Before the temporaries exist the job table is built for the lookups
  // only; iteration jobs then point at the resident LUTs.)
  const bool par = m->iters > 1 && k >= 0 && m->T_it != nullptr;
  const size_t kk = k >= 0 ? static_cast<size_t>(k) : static_cast<size_t>(m->iters - 1);
  const bool temps = m->dR_it != nullptr;
  j->T = par ? m->T_it + kk * z.t_texels() : m->T;
  j->S = m->S;
  j->SM = m->SM;
  j->E = par ? m->E_it + kk * z.e_texels() : m->E;
  j->E_it = m->E_it;
  j->dE = temps ? m->dE_it + kk * z.e_texels() : nullptr;
  j->dR = temps ? m->dR_it + kk * z.s_texels() : nullptr;
  j->dM = temps ? m->dM_it + kk * z.s_texels() : nullptr;
  j->dJ = temps ? m->dJ_it + kk * z.s_texels() *
                                (m->spec.half_precision ? sizeof(uint2) : sizeof(float3))
               : nullptr;
  j->niter = m->iters;
  j->OD = m->OD;  // every job of the model, the final transmittance's too
  set_sizes(j, m->sizes);
  j->combined = m->spec.combined ? 1 : 0;
  j->half_luts = m->spec.half_precision ? 1 : 0;
  j->S16 = m->S16;
  j->SM16 = m->SM16;
  j->Q = m->Q;
  j->acc64 = accumulates_fp64(m->spec) ? 1 : 0;
  j->S64 = par ? m->S64 + 4 * kk * z.s_texels() : m->S64;
  j->SM64 = par && m->SM64 != nullptr ? m->SM64 + 4 * kk * z.s_texels() : m->SM64;
  j->scales = m->d_scales;
  j->maxbits = m->d_maxbits;
  for (int c = 0; c < 3; ++c) {
    j->sky_k[c] = static_cast<float>(m->sky_k[c]);
    j->sun_k[c] = static_cast<float>(m->sun_k[c]);
  }
  const Tiles t = tiles_for(m->sizes);
  j->sd_tile = t.sd;
  j->ms_tile = t.ms;
  j->ms_row_order = t.ms_row_order;
  j->ms_rows = m->d_ms_rows;
}

// Rebuild the host job table (LUT pointers may have changed) and upload it.
atmo_status upload_jobs(atmo_model* m, unsigned orders, hipStream_t stream) {
  const Spec& s = m->spec;
  const int iters = m->iters;
  m->jobs.assign(iters + 2, atmo::Job{});
  for (int i = 0; i < iters; ++i) {
    double lambdas[3];
    atmo::Job& j = m->jobs[i];
    iteration(s, i, lambdas, j.lum);
    lower(s, lambdas, &j);
    fill_job_luts(m, &j, i);
    j.orders = static_cast<int>(orders);
  }
  // Transmittance at 680/550/440 nm (atmosphere/model.cc:951-963) and the
  // lookup parameters (the shader API header, model.cc:769-772).
  const double rgb[3] = {kLambdaR, kLambdaG, kLambdaB};
  for (int k = iters; k < iters + 2; ++k) {
    atmo::Job& j = m->jobs[k];
    lower(s, rgb, &j);
    fill_job_luts(m, &j, -1);
    for (int q = 0; q < 9; ++q) j.lum[q] = (q % 4 == 0) ? 1.0f : 0.0f;
  }
  ATMO_HIP(hipMemcpyAsync(m->d_jobs, m->jobs.data(),
                          m->jobs.size() * sizeof(atmo::Job),
                          hipMemcpyHostToDevice, stream));
  // Block prefixes of the model's launches (they follow the tiles): every
  // job has the same sizes, so pass p's prefix over the iterations is
  // {0, b, 2b, ...} with b = pass_blocks(p, job).
  m->first_block_host.assign(kPasses * (iters + 1), 0);
  for (int p = 0; p < kPasses; ++p) {
    const int b = pass_blocks(static_cast<Pass>(p), m->jobs[0]);
    for (int i = 0; i <= iters; ++i) m->first_block_host[p * (iters + 1) + i] = i * b;
  }
  ATMO_HIP(hipMemcpyAsync(m->d_first_block, m->first_block_host.data(),
                          m->first_block_host.size() * sizeof(int), hipMemcpyHostToDevice,
                          stream));
  return ATMO_OK;
}

// One launch of pass p over the model's jobs first_job .. first_job + njobs - 1.
atmo::PassLaunch model_launch(const atmo_model* m, int first_job, int njobs, Pass p) {
  atmo::PassLaunch l;
  l.jobs = m->d_jobs + first_job;
  l.first_block = m->d_first_block + p * (m->iters + 1);
  l.njobs = njobs;
  l.total_blocks = m->first_block_host[p * (m->iters + 1) + njobs];
  l.max_nnu = m->sizes.nnu;
  l.max_ew = m->sizes.ew;
  l.max_w = m->sizes.nnu * m->sizes.nmus;
  l.min_w = l.max_w;
  l.max_lds_w = l.max_w <= atmo::kMsMaxRow ? l.max_w : 0;
  l.kinds = accumulates_fp64(m->spec) ? atmo::kJobsAcc64 : atmo::kJobsF32;
  l.dj_kinds = m->spec.half_precision ? atmo::kDjRgb21 : atmo::kDjF32;
  const atmo::Job& j = m->jobs[first_job];
  l.tile = p == kSD ? j.sd_tile : (p == kMS ? j.ms_tile : 0);
  return l;
}

atmo_status build_lookup_copy(atmo_model* m, hipStream_t s) {
  if (m->Q == nullptr) return ATMO_OK;
  return check_launch(atmo::launch_interleave(model_launch(m, m->iters, 1, kPack), s),
                      "interleaved lookup copy");
}

atmo_status check_launch(hipError_t e, const char* what) {
  if (e != hipSuccess)
    return fail(ATMO_ERROR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return ATMO_OK;
}

// Records start/stop events around each launch when 'on' (the model's
// ATMO_OPTION_PASS_TIMING, except in the allocation-free Init).
struct LaunchTimer {
  atmo_model* m;
  hipStream_t s;
  bool on;
  size_t next_event = 0;
  hipEvent_t take() {
    if (next_event == m->event_pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      m->event_pool.push_back(e);
    }
    return m->event_pool[next_event++];
  }
  template <class F>
  atmo_status run(int pass, int order, const char* what, F&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (on) {
      a = take();
      b = take();
      if (a == nullptr || b == nullptr)
        return fail(ATMO_ERROR_DEVICE, "hipEventCreate failed");
      if (hipEventRecord(a, s) != hipSuccess)
        return fail(ATMO_ERROR_DEVICE, "hipEventRecord failed");
    }
    ATMO_TRY(check_launch(launch(), what));
    if (on) {
      if (hipEventRecord(b, s) != hipSuccess)
        return fail(ATMO_ERROR_DEVICE, "hipEventRecord failed");
      m->timed.push_back({pass, order, a, b});
    }
    return ATMO_OK;
  }
};

// Model::Precompute (atmosphere/model.cc:1048-1215) for every iteration at
// once: each launch covers the jobs of all iterations, which write their
// own temporaries and accumulators.
atmo_status enqueue_precompute(atmo_model* m, unsigned orders, LaunchTimer& t) {
  hipStream_t s = t.s;
  const int n = m->iters;
  if (m->OD != nullptr) {
    ATMO_TRY(t.run(kT, 0, "optical depth", [&] {
      return atmo::launch_optical_depth(model_launch(m, 0, 1, kOD), s);
    }));
  }
  ATMO_TRY(t.run(kT, 0, "transmittance", [&] {
    return atmo::launch_transmittance(model_launch(m, 0, n, kT), s);
  }));
  ATMO_TRY(t.run(kE0, 0, "direct irradiance", [&] {
    return atmo::launch_direct_irradiance(model_launch(m, 0, n, kE0), s);
  }));
  ATMO_TRY(t.run(kSS, 1, "single scattering", [&] {
    return atmo::launch_single_scattering(model_launch(m, 0, n, kSS), s);
  }));
  for (unsigned order = 2; order <= orders; ++order) {
    const int o = static_cast<int>(order);
    ATMO_TRY(t.run(kSD, o, "scattering density", [&] {
      return atmo::launch_scattering_density(model_launch(m, 0, n, kSD), o, s);
    }));
    ATMO_TRY(t.run(kII, o - 1, "indirect irradiance", [&] {
      return atmo::launch_indirect_irradiance(model_launch(m, 0, n, kII), o - 1, s);
    }));
    ATMO_TRY(t.run(kMS, o, "multiple scattering", [&] {
      return atmo::launch_multiple_scattering(model_launch(m, 0, n, kMS), s);
    }));
  }
  return ATMO_OK;
}

// Init after its temporaries are in place: the job table, the passes, the
// sums of the iterations' accumulators, the final transmittance, FP16
// conversion and the lookup copy.
atmo_status enqueue_init(atmo_model* m, unsigned orders, hipStream_t s, bool timing) {
  m->last_orders = orders;
  ATMO_TRY(upload_jobs(m, orders, s));
  const int iters = m->iters;
  m->timed.clear();
  LaunchTimer timer{m, s, timing};
  ATMO_TRY(enqueue_precompute(m, orders, timer));
  // Job 'iters' works on the resident LUTs: the sum of the iterations'
  // accumulators, then the transmittance at 680/550/440 nm.
  if (accumulates_fp64(m->spec)) {
    ATMO_TRY(check_launch(atmo::launch_acc64_to_f32(model_launch(m, iters, 1, kAcc), s),
                          "fp64 accumulators"));
  }
  if (m->spec.num_precomputed_wavelengths > 3) {
    ATMO_TRY(timer.run(kT, 0, "transmittance (rgb)", [&] {
      return atmo::launch_transmittance(model_launch(m, iters, 1, kT), s);
    }));
  }
  if (m->spec.half_precision) {
    ATMO_TRY(check_launch(atmo::launch_clear_max(model_launch(m, iters, 1, kClear), s),
                          "fp16 maxima"));
    ATMO_TRY(check_launch(atmo::launch_to_half(model_launch(m, iters, 1, kHalf), s),
                          "fp16 conversion"));
  }
  ATMO_TRY(build_lookup_copy(m, s));
  m->initialized = true;
  return ATMO_OK;
}

// The Init temporaries of atmo_model_init_static_scratch: a file-scope
// device array, one copy per GPU, allocated when the code object is loaded
// and held until the process ends. Every call on a GPU shares that GPU's
// copy. Host code never passes the array's name to a kernel or a copy (it is
// a host-side stand-in): it uses the device address that
// hipGetSymbolAddress returns, fetched once per device with that device
// current and kept in g_scratch_ptr.
constexpr size_t kStaticScratchBytes = static_cast<size_t>(ATMO_STATIC_SCRATCH_BYTES);
static __device__ __align__(256) char scratch[kStaticScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

atmo_status static_scratch(int device, char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  int current = -1;
  ATMO_HIP(hipGetDevice(&current));
  if (current != device)
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                "model was created on device " + std::to_string(device) +
                    ", current device is " + std::to_string(current));
  if (device < 0 || device >= kMaxDevices)
    return fail(ATMO_ERROR_UNSUPPORTED, "device index beyond the static scratch table");
  if (g_scratch_ptr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(ATMO_ERROR_DEVICE,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[device];
  return ATMO_OK;
}

atmo_status validate_layers(int n, const atmo_density_layer* l,
                            const char* what) {
  if (n < 0 || n > 2 || (n > 0 && l == nullptr))
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                std::string(what) + ": at most 2 density layers");
  return ATMO_OK;
}

atmo_status validate_sizes(const atmo_lut_sizes& z) {
  const int v[8] = {z.transmittance_width, z.transmittance_height, z.scattering_r,
                    z.scattering_mu,       z.scattering_mu_s,      z.scattering_nu,
                    z.irradiance_width,    z.irradiance_height};
  for (int x : v)
    if (x < 2 || x > 4096)
      return fail(ATMO_ERROR_INVALID_ARGUMENT, "LUT sizes must be in [2, 4096]");
  // Each half of the mu axis (ground and sky rays) is a texture axis of MU / 2
  // texels whose coordinate mapping divides by 1 - 2 / MU
  // (functions.glsl:346-349, :859-860): at least 2 texels per half.
  if (z.scattering_mu % 2 != 0 || z.scattering_mu < 4)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "scattering_mu must be even and >= 4");
  if (z.scattering_nu > 32)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "scattering_nu must be <= 32");
  // The scattering-density pass keeps a row of the irradiance LUT in LDS
  // next to its nu tables: at NU = 32 and E_W = 512 a block needs 50 KB.
  if (z.irradiance_width > 512)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "irradiance_width must be <= 512");
  const double texels = static_cast<double>(z.scattering_r) * z.scattering_mu *
                        z.scattering_mu_s * z.scattering_nu;
  // At most 2^28 texels (4 GiB in FP32 RGBA): kernels index texels with
  // 32-bit byte offsets.
  if (texels > 268435456.0)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "scattering LUT too large (> 2^28 texels)");
  return ATMO_OK;
}

bool finite_all(const double* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

atmo_status validate_params(const atmo_model_params* p) {
  if (p == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "params is NULL");
  const int n = p->num_wavelengths;
  if (n < 1 || p->wavelengths == nullptr || p->solar_irradiance == nullptr ||
      p->rayleigh_scattering == nullptr || p->mie_scattering == nullptr ||
      p->mie_extinction == nullptr || p->absorption_extinction == nullptr ||
      p->ground_albedo == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "missing spectrum");
  for (int i = 0; i + 1 < n; ++i)
    if (!(p->wavelengths[i] < p->wavelengths[i + 1]))
      return fail(ATMO_ERROR_INVALID_ARGUMENT, "wavelengths must increase");
  const double* spectra[7] = {p->wavelengths,     p->solar_irradiance,
                              p->rayleigh_scattering, p->mie_scattering,
                              p->mie_extinction,  p->absorption_extinction,
                              p->ground_albedo};
  for (const double* v : spectra)
    if (!finite_all(v, n)) return fail(ATMO_ERROR_INVALID_ARGUMENT, "non-finite spectrum");
  ATMO_TRY(validate_layers(p->num_rayleigh_layers, p->rayleigh_density, "rayleigh_density"));
  ATMO_TRY(validate_layers(p->num_mie_layers, p->mie_density, "mie_density"));
  ATMO_TRY(validate_layers(p->num_absorption_layers, p->absorption_density,
                           "absorption_density"));
  const double scalars[6] = {p->sun_angular_radius, p->bottom_radius, p->top_radius,
                             p->mie_phase_function_g, p->max_sun_zenith_angle,
                             p->length_unit_in_meters};
  if (!finite_all(scalars, 6)) return fail(ATMO_ERROR_INVALID_ARGUMENT, "non-finite parameter");
  if (!(p->bottom_radius > 0.0 && p->top_radius > p->bottom_radius))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "need 0 < bottom_radius < top_radius");
  if (!(p->sun_angular_radius > 0.0))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "sun_angular_radius must be > 0");
  if (!(p->length_unit_in_meters > 0.0))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "length_unit_in_meters must be > 0");
  if (!(std::fabs(p->mie_phase_function_g) < 1.0))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "|mie_phase_function_g| must be < 1");
  if (!(std::cos(p->max_sun_zenith_angle) < 1.0))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "max_sun_zenith_angle must be > 0");
  return ATMO_OK;
}

void build_spec(const atmo_model_params* p, Spec* sp) {
  Spec& s = *sp;
  const int n = p->num_wavelengths;
  auto vec = [n](const double* v) { return std::vector<double>(v, v + n); };
  s.wavelengths = vec(p->wavelengths);
  s.solar = vec(p->solar_irradiance);
  s.rayleigh = vec(p->rayleigh_scattering);
  s.mie_sca = vec(p->mie_scattering);
  s.mie_ext = vec(p->mie_extinction);
  s.absorption = vec(p->absorption_extinction);
  s.albedo = vec(p->ground_albedo);
  s.prof_r.assign(p->rayleigh_density, p->rayleigh_density + p->num_rayleigh_layers);
  s.prof_m.assign(p->mie_density, p->mie_density + p->num_mie_layers);
  s.prof_a.assign(p->absorption_density,
                  p->absorption_density + p->num_absorption_layers);
  s.sun_angular_radius = p->sun_angular_radius;
  s.bottom_radius = p->bottom_radius;
  s.top_radius = p->top_radius;
  s.g = p->mie_phase_function_g;
  s.max_sun_zenith_angle = p->max_sun_zenith_angle;
  s.length_unit = p->length_unit_in_meters;
  s.num_precomputed_wavelengths = p->num_precomputed_wavelengths;
  s.combined = p->combine_scattering_textures != 0;
  s.half_precision = p->half_precision != 0;
}

// The radiance API exists only when at most 3 wavelengths are precomputed
// (RADIANCE_API_ENABLED, atmosphere/model.cc:226, 769-772).
atmo_status radiance_api(const atmo_model* m) {
  if (m != nullptr && m->spec.num_precomputed_wavelengths > 3)
    return fail(ATMO_ERROR_UNSUPPORTED,
                "radiance lookups need num_precomputed_wavelengths <= 3; use the "
                "luminance lookups");
  return ATMO_OK;
}

size_t lut_bytes(const atmo_model* m, bool with_single_mie) {
  const Sizes& z = m->sizes;
  const size_t texel = m->spec.half_precision ? sizeof(uint2) : sizeof(float4);
  size_t b = (z.t_texels() + z.e_texels()) * sizeof(float4) + z.s_texels() * texel;
  if (with_single_mie) b += z.s_texels() * texel;
  if (m->Q != nullptr) b += interleaved_bytes(z, m->spec.half_precision, m->spec.combined);
  return b;
}

}  // namespace atmo_rt

using namespace atmo_rt;

extern "C" {

const char* atmo_last_error(void) { return g_last_error.c_str(); }

atmo_status atmo_model_create(const atmo_model_params* p,
                              const atmo_lut_sizes* sizes, atmo_model** out) {
  if (out == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  ATMO_TRY(validate_params(p));
  atmo_lut_sizes z = sizes != nullptr
                         ? *sizes
                         : atmo_lut_sizes{256, 64, 32, 128, 32, 8, 64, 16};
  ATMO_TRY(validate_sizes(z));

  // On any failure below, the partially built model is destroyed (which frees
  // whatever was allocated).
  std::unique_ptr<atmo_model, void (*)(atmo_model*)> m(
      new atmo_model(), [](atmo_model* x) { (void)atmo_model_destroy(x); });
  build_spec(p, &m->spec);
  const Spec& s = m->spec;
  m->sizes = Sizes{z.transmittance_width, z.transmittance_height, z.scattering_r,
                   z.scattering_mu,       z.scattering_mu_s,      z.scattering_nu,
                   z.irradiance_width,    z.irradiance_height};
  // GetSolarRadiance of the GL API (atmosphere/model.cc:227-230), at the
  // three RGB wavelengths.
  const double rgb[3] = {kLambdaR, kLambdaG, kLambdaB};
  for (int c = 0; c < 3; ++c)
    m->solar_radiance[c] = static_cast<float>(
        interpolate(s.wavelengths, s.solar, rgb[c]) /
        (atmo::kPi * s.sun_angular_radius * s.sun_angular_radius));
  // SKY/SUN_SPECTRAL_RADIANCE_TO_LUMINANCE (atmosphere/model.cc:668-686): in
  // illuminance mode the LUTs hold luminance / MAX_LUMINOUS_EFFICACY.
  if (s.num_precomputed_wavelengths > 3) {
    for (int c = 0; c < 3; ++c) m->sky_k[c] = atmo_cie::kMaxLuminousEfficacy;
  } else {
    luminance_factors(s.wavelengths, s.solar, -3.0, m->sky_k);
  }
  luminance_factors(s.wavelengths, s.solar, 0.0, m->sun_k);

  ATMO_HIP(hipGetDevice(&m->device));
  const Sizes& sz = m->sizes;
  ATMO_HIP(hipMalloc(&m->T, sz.t_texels() * sizeof(float4)));
  if (s.half_precision) {
    ATMO_HIP(hipMalloc(&m->S16, sz.s_texels() * sizeof(uint2)));
    if (!s.combined) ATMO_HIP(hipMalloc(&m->SM16, sz.s_texels() * sizeof(uint2)));
    ATMO_HIP(hipMalloc(&m->d_scales, 8 * sizeof(float)));
    ATMO_HIP(hipMalloc(&m->d_maxbits, 8 * sizeof(unsigned)));
  } else {
    ATMO_HIP(hipMalloc(&m->S, sz.s_texels() * sizeof(float4)));
    if (!s.combined) ATMO_HIP(hipMalloc(&m->SM, sz.s_texels() * sizeof(float4)));
  }
  ATMO_HIP(hipMalloc(&m->E, sz.e_texels() * sizeof(float4)));
  if (!s.combined)
    ATMO_HIP(hipMalloc(&m->Q, interleaved_bytes(sz, s.half_precision, s.combined)));
  const int iters = num_iterations(s);
  m->iters = iters;
  ATMO_HIP(hipMalloc(&m->d_jobs, (iters + 2) * sizeof(atmo::Job)));
  ATMO_HIP(hipMalloc(&m->d_first_block, kPasses * (iters + 1) * sizeof(int)));
  ATMO_HIP(hipMalloc(&m->d_ms_rows, static_cast<size_t>(sz.nr) * sz.nmu * sizeof(int)));
  ATMO_TRY(upload_jobs(m.get(), 4, nullptr));
  const std::vector<int> ray_rows = ms_ray_order(m->jobs[0]);
  ATMO_HIP(hipMemcpy(m->d_ms_rows, ray_rows.data(), ray_rows.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  ATMO_HIP(hipStreamSynchronize(nullptr));
  *out = m.release();
  return ATMO_OK;
}

atmo_status atmo_model_destroy(atmo_model* m) {
  if (m == nullptr) return ATMO_OK;
  if (m->live_batches > 0)
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                "model is in a live batch: destroy the batch first");
  (void)hipDeviceSynchronize();
  free_temporaries(m);
  void* bufs[9] = {m->T, m->S, m->SM, m->E, m->S16, m->SM16, m->d_scales, m->d_maxbits, m->Q};
  for (void* b : bufs)
    if (b != nullptr) (void)hipFree(b);
  if (m->d_jobs != nullptr) (void)hipFree(m->d_jobs);
  if (m->d_first_block != nullptr) (void)hipFree(m->d_first_block);
  if (m->d_ms_rows != nullptr) (void)hipFree(m->d_ms_rows);
  for (hipEvent_t e : m->event_pool) (void)hipEventDestroy(e);
  delete m;
  return ATMO_OK;
}

atmo_status atmo_model_set_option(atmo_model* m, atmo_option option, int value) {
  if (m == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "model is NULL");
  switch (option) {
    case ATMO_OPTION_RETAIN_INTERMEDIATES:
      m->retain = value != 0;
      return ATMO_OK;
    case ATMO_OPTION_PASS_TIMING:
      m->timing = value != 0;
      return ATMO_OK;
    case ATMO_OPTION_INTERLEAVED_LOOKUPS: {
      const bool on = value != 0 && !m->spec.combined;
      if (on == (m->Q != nullptr)) return ATMO_OK;
      if (m->live_batches > 0)
        return fail(ATMO_ERROR_INVALID_ARGUMENT,
                    "model is in a live batch, whose launches write the lookup copy");
      // The job table holds Q: wait for work that may use it, switch, and
      // upload the table again.
      ATMO_HIP(hipDeviceSynchronize());
      if (on) {
        ATMO_HIP(hipMalloc(&m->Q, interleaved_bytes(m->sizes, m->spec.half_precision,
                                                    m->spec.combined)));
      } else {
        ATMO_HIP(hipFree(m->Q));
        m->Q = nullptr;
      }
      ATMO_TRY(upload_jobs(m, m->last_orders, nullptr));
      if (on && m->initialized) ATMO_TRY(build_lookup_copy(m, nullptr));
      ATMO_HIP(hipDeviceSynchronize());
      return ATMO_OK;
    }
  }
  return fail(ATMO_ERROR_INVALID_ARGUMENT, "unknown option");
}

atmo_status atmo_model_init(atmo_model* m, unsigned orders, void* stream) {
  if (m == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "model is NULL");
  if (orders < 1) return fail(ATMO_ERROR_INVALID_ARGUMENT, "num_scattering_orders must be >= 1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  ATMO_TRY(alloc_temporaries(m));
  ATMO_TRY(enqueue_init(m, orders, s, m->timing));
  // A live batch holding the model keeps reading and writing its temporaries.
  if (!m->retain && m->live_batches == 0) {
    // Init then returns once the precomputation is complete, as the GL
    // Init does once its draw calls have been issued and the temporaries
    // deleted (atmosphere/model.cc:966-974).
    ATMO_HIP(hipStreamSynchronize(s));
    free_temporaries(m);
  }
  return ATMO_OK;
}

atmo_status atmo_model_init_static_scratch(atmo_model* m, unsigned orders, void* stream) {
  if (m == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "model is NULL");
  if (orders < 1) return fail(ATMO_ERROR_INVALID_ARGUMENT, "num_scattering_orders must be >= 1");
  if (m->live_batches > 0)
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                "model is in a live batch, whose launches use its own temporaries");
  if (m->external_temporaries)
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                "model's temporaries are carved from caller memory (a server model)");
  const size_t need = temporaries_bytes(m);
  if (need > kStaticScratchBytes)
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                "Init temporaries of this model take " + std::to_string(need) +
                    " bytes; the static scratch holds " +
                    std::to_string(kStaticScratchBytes) + " (use atmo_model_init)");
  char* base = nullptr;
  ATMO_TRY(static_scratch(m->device, &base));
  hipStream_t s = static_cast<hipStream_t>(stream);
  // The passes write every temporary before reading it (the first pass of
  // each writes, never accumulates), so the scratch needs no clearing.
  use_external_temporaries(m, base);
  const atmo_status st = enqueue_init(m, orders, s, false);
  drop_external_temporaries(m);
  // Retained temporaries of an earlier Init stay the model's; the job table
  // points at them again at the next Init.
  if (m->temp_block != nullptr)
    carve(m->sizes, m->spec.half_precision, m->spec.combined, m->iters, m->temp_block, m);
  return st;
}

atmo_status atmo_model_pass_timings(atmo_model* m, atmo_pass_timing* out, int capacity,
                                    int* count) {
  if (m == nullptr || count == nullptr || (capacity > 0 && out == nullptr))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  *count = static_cast<int>(m->timed.size());
  for (int i = 0; i < *count && i < capacity; ++i) {
    const atmo_model::Timed& t = m->timed[i];
    ATMO_HIP(hipEventSynchronize(t.stop));
    float ms = 0.0f;
    ATMO_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
    out[i] = {t.pass, t.order, ms};
  }
  return ATMO_OK;
}

atmo_status atmo_model_get_lut(const atmo_model* m, atmo_lut_id id, atmo_lut_view* v) {
  if (m == nullptr || v == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  const Sizes& z = m->sizes;
  std::memset(v, 0, sizeof(*v));
  v->channels = 4;
  v->dtype = ATMO_DTYPE_F32;
  for (int c = 0; c < 4; ++c) v->scale[c] = 1.0f;
  const bool scat = id == ATMO_LUT_SCATTERING || id == ATMO_LUT_SINGLE_MIE_SCATTERING ||
                    id == ATMO_LUT_DELTA_RAYLEIGH || id == ATMO_LUT_DELTA_MIE ||
                    id == ATMO_LUT_DELTA_SCATTERING_DENSITY;
  if (scat) {
    v->width = z.nnu * z.nmus;
    v->height = z.nmu;
    v->depth = z.nr;
  } else if (id == ATMO_LUT_TRANSMITTANCE) {
    v->width = z.tw;
    v->height = z.th;
    v->depth = 1;
  } else {
    v->width = z.ew;
    v->height = z.eh;
    v->depth = 1;
  }
  const bool half = m->spec.half_precision &&
                    (id == ATMO_LUT_SCATTERING || id == ATMO_LUT_SINGLE_MIE_SCATTERING);
  if (half) {
    float sc[8];
    ATMO_HIP(hipDeviceSynchronize());
    ATMO_HIP(hipMemcpy(sc, m->d_scales, sizeof(sc), hipMemcpyDeviceToHost));
    v->dtype = ATMO_DTYPE_F16;
    const int off = id == ATMO_LUT_SCATTERING ? 0 : 4;
    for (int c = 0; c < 4; ++c) v->scale[c] = sc[off + c];
    v->data = id == ATMO_LUT_SCATTERING ? static_cast<void*>(m->S16)
                                        : static_cast<void*>(m->SM16);
    if (v->data == nullptr)
      return fail(ATMO_ERROR_UNSUPPORTED, "this LUT is not held by the model");
    return ATMO_OK;
  }
  switch (id) {
    case ATMO_LUT_TRANSMITTANCE: v->data = m->T; break;
    case ATMO_LUT_SCATTERING: v->data = m->S; break;
    case ATMO_LUT_SINGLE_MIE_SCATTERING: v->data = m->SM; break;
    case ATMO_LUT_IRRADIANCE: v->data = m->E; break;
    case ATMO_LUT_DELTA_IRRADIANCE: v->data = m->dE; break;
    case ATMO_LUT_DELTA_RAYLEIGH: v->data = m->dR; break;
    case ATMO_LUT_DELTA_MIE: v->data = m->dM; break;
    case ATMO_LUT_DELTA_SCATTERING_DENSITY:
      v->data = m->dJ;
      v->channels = 3;  // RGB texels
      if (m->spec.half_precision) v->dtype = ATMO_DTYPE_RGB21;
      break;
    default: return fail(ATMO_ERROR_INVALID_ARGUMENT, "unknown LUT id");
  }
  if (v->data == nullptr)
    return fail(ATMO_ERROR_UNSUPPORTED, "this LUT is not held by the model");
  return ATMO_OK;
}

atmo_status atmo_model_copy_lut(const atmo_model* m, atmo_lut_id id, float* host) {
  atmo_lut_view v;
  ATMO_TRY(atmo_model_get_lut(m, id, &v));
  if (host == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "host is NULL");
  ATMO_HIP(hipDeviceSynchronize());
  const size_t texels = static_cast<size_t>(v.width) * v.height * v.depth;
  if (v.dtype == ATMO_DTYPE_RGB21) {
    std::vector<uint2> raw(texels);
    ATMO_HIP(hipMemcpy(raw.data(), v.data, texels * sizeof(uint2), hipMemcpyDeviceToHost));
    for (size_t t = 0; t < texels; ++t) {
      const float3 f = atmo::unpack_f21x3(raw[t]);
      host[4 * t] = f.x;
      host[4 * t + 1] = f.y;
      host[4 * t + 2] = f.z;
      host[4 * t + 3] = 0.0f;
    }
    return ATMO_OK;
  }
  if (v.dtype == ATMO_DTYPE_F32 && v.channels == 3) {
    std::vector<float> rgb(texels * 3);
    ATMO_HIP(hipMemcpy(rgb.data(), v.data, texels * 3 * sizeof(float),
                       hipMemcpyDeviceToHost));
    for (size_t t = 0; t < texels; ++t) {
      for (int k = 0; k < 3; ++k) host[4 * t + k] = rgb[3 * t + k];
      host[4 * t + 3] = 0.0f;
    }
    return ATMO_OK;
  }
  if (v.dtype == ATMO_DTYPE_F32) {
    ATMO_HIP(hipMemcpy(host, v.data, texels * 4 * sizeof(float), hipMemcpyDeviceToHost));
    return ATMO_OK;
  }
  std::vector<uint2> raw(texels);
  ATMO_HIP(hipMemcpy(raw.data(), v.data, texels * sizeof(uint2), hipMemcpyDeviceToHost));
  for (size_t t = 0; t < texels; ++t) {
    const float4 f = atmo::unpack_half4(raw[t]);
    const float c[4] = {f.x, f.y, f.z, f.w};
    for (int k = 0; k < 4; ++k) host[4 * t + k] = v.scale[k] > 0.0f ? c[k] / v.scale[k] : 0.0f;
  }
  return ATMO_OK;
}

atmo_status atmo_model_footprint(const atmo_model* m, size_t* resident, size_t* init_scratch) {
  if (m == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "model is NULL");
  const Sizes& z = m->sizes;
  if (resident != nullptr) *resident = lut_bytes(m, !m->spec.combined);
  if (init_scratch != nullptr) {
    (void)z;
    *init_scratch = temporaries_bytes(m);
  }
  return ATMO_OK;
}

atmo_status atmo_get_solar_radiance(const atmo_model* m, float* rgb) {
  if (m == nullptr || rgb == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  ATMO_TRY(radiance_api(m));
  for (int c = 0; c < 3; ++c) rgb[c] = m->solar_radiance[c];
  return ATMO_OK;
}

static atmo_status check_query(const atmo_model* m, int64_t n) {
  if (m == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "model is NULL");
  if (n < 0) return fail(ATMO_ERROR_INVALID_ARGUMENT, "n < 0");
  if (!m->initialized) return fail(ATMO_ERROR_NOT_INITIALIZED, "model not initialized");
  return ATMO_OK;
}

static const atmo::Job* query_job(const atmo_model* m) {
  return m->d_jobs + (m->jobs.size() - 1);
}

static atmo_status sky(const atmo_model* m, int64_t n, const float* camera,
                       const float* view_ray, const float* shadow_length,
                       const float* sun_direction, float* radiance, float* transmittance,
                       void* stream, int luminance) {
  ATMO_TRY(check_query(m, n));
  if (n > 0 && (!camera || !view_ray || !shadow_length || !sun_direction || !radiance ||
                !transmittance))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL array");
  return check_launch(atmo::launch_sky_radiance(query_job(m), n, camera, view_ray,
                                                shadow_length, sun_direction, radiance,
                                                transmittance, luminance,
                                                static_cast<hipStream_t>(stream)),
                      "sky radiance");
}

static atmo_status sky_to_point(const atmo_model* m, int64_t n, const float* camera,
                                const float* point, const float* shadow_length,
                                const float* sun_direction, float* radiance,
                                float* transmittance, void* stream, int luminance) {
  ATMO_TRY(check_query(m, n));
  if (n > 0 && (!camera || !point || !shadow_length || !sun_direction || !radiance ||
                !transmittance))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL array");
  return check_launch(atmo::launch_sky_radiance_to_point(
                          query_job(m), n, camera, point, shadow_length, sun_direction,
                          radiance, transmittance, luminance,
                          static_cast<hipStream_t>(stream)),
                      "sky radiance to point");
}

static atmo_status sun_and_sky(const atmo_model* m, int64_t n, const float* point,
                               const float* normal, const float* sun_direction,
                               float* sun_out, float* sky_out, void* stream,
                               int luminance) {
  ATMO_TRY(check_query(m, n));
  if (n > 0 && (!point || !normal || !sun_direction || !sun_out || !sky_out))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL array");
  return check_launch(atmo::launch_sun_and_sky_irradiance(
                          query_job(m), n, point, normal, sun_direction, sun_out, sky_out,
                          luminance, static_cast<hipStream_t>(stream)),
                      "sun and sky irradiance");
}

atmo_status atmo_get_sky_radiance(const atmo_model* m, int64_t n, const float* camera,
                                  const float* view_ray, const float* shadow_length,
                                  const float* sun_direction, float* radiance,
                                  float* transmittance, void* stream) {
  ATMO_TRY(radiance_api(m));
  return sky(m, n, camera, view_ray, shadow_length, sun_direction, radiance,
             transmittance, stream, 0);
}

atmo_status atmo_get_sky_radiance_to_point(const atmo_model* m, int64_t n,
                                           const float* camera, const float* point,
                                           const float* shadow_length,
                                           const float* sun_direction, float* radiance,
                                           float* transmittance, void* stream) {
  ATMO_TRY(radiance_api(m));
  return sky_to_point(m, n, camera, point, shadow_length, sun_direction, radiance,
                      transmittance, stream, 0);
}

atmo_status atmo_get_sun_and_sky_irradiance(const atmo_model* m, int64_t n,
                                            const float* point, const float* normal,
                                            const float* sun_direction,
                                            float* sun_irradiance, float* sky_irradiance,
                                            void* stream) {
  ATMO_TRY(radiance_api(m));
  return sun_and_sky(m, n, point, normal, sun_direction, sun_irradiance, sky_irradiance,
                     stream, 0);
}

atmo_status atmo_get_sky_luminance(const atmo_model* m, int64_t n, const float* camera,
                                   const float* view_ray, const float* shadow_length,
                                   const float* sun_direction, float* luminance,
                                   float* transmittance, void* stream) {
  return sky(m, n, camera, view_ray, shadow_length, sun_direction, luminance,
             transmittance, stream, 1);
}

atmo_status atmo_get_sky_luminance_to_point(const atmo_model* m, int64_t n,
                                            const float* camera, const float* point,
                                            const float* shadow_length,
                                            const float* sun_direction, float* luminance,
                                            float* transmittance, void* stream) {
  return sky_to_point(m, n, camera, point, shadow_length, sun_direction, luminance,
                      transmittance, stream, 1);
}

atmo_status atmo_get_sun_and_sky_illuminance(const atmo_model* m, int64_t n,
                                             const float* point, const float* normal,
                                             const float* sun_direction,
                                             float* sun_illuminance,
                                             float* sky_illuminance, void* stream) {
  return sun_and_sky(m, n, point, normal, sun_direction, sun_illuminance, sky_illuminance,
                     stream, 1);
}

atmo_status atmo_get_solar_luminance(const atmo_model* m, float* rgb) {
  if (m == nullptr || rgb == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  for (int c = 0; c < 3; ++c)
    rgb[c] = static_cast<float>(m->solar_radiance[c] * m->sun_k[c]);
  return ATMO_OK;
}

atmo_status atmo_get_luminance_factors(const atmo_model* m, double* sky_k, double* sun_k) {
  if (m == nullptr || sky_k == nullptr || sun_k == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  for (int c = 0; c < 3; ++c) {
    sky_k[c] = m->sky_k[c];
    sun_k[c] = m->sun_k[c];
  }
  return ATMO_OK;
}

atmo_status atmo_lower_parameters(const atmo_model_params* p, const double* lambdas3,
                                  atmo_atmosphere* out) {
  ATMO_TRY(validate_params(p));
  if (lambdas3 == nullptr || out == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  Spec s;
  build_spec(p, &s);
  atmo::Job j{};
  lower(s, lambdas3, &j);
  for (int c = 0; c < 3; ++c) {
    out->solar_irradiance[c] = j.solar[c];
    out->rayleigh_scattering[c] = j.rayleigh[c];
    out->mie_scattering[c] = j.mie_sca[c];
    out->mie_extinction[c] = j.mie_ext[c];
    out->absorption_extinction[c] = j.absorption[c];
    out->ground_albedo[c] = j.albedo[c];
  }
  out->sun_angular_radius = j.sun_angular_radius;
  out->bottom_radius = j.rb;
  out->top_radius = j.rt;
  out->mie_phase_function_g = j.g;
  out->mu_s_min = j.mu_s_min;
  auto copy_layers = [](const atmo::LayerD* src, atmo_density_layer* dst) {
    for (int i = 0; i < 2; ++i)
      dst[i] = {src[i].width, src[i].exp_term, src[i].exp_scale, src[i].linear_term,
                src[i].constant_term};
  };
  copy_layers(j.prof_r, out->rayleigh_density);
  copy_layers(j.prof_m, out->mie_density);
  copy_layers(j.prof_a, out->absorption_density);
  return ATMO_OK;
}

int atmo_num_iterations(const atmo_model_params* p) {
  if (validate_params(p) != ATMO_OK) return 0;
  Spec s;
  build_spec(p, &s);
  return num_iterations(s);
}

atmo_status atmo_iteration(const atmo_model_params* p, int i, double* lambdas3,
                           float* matrix9) {
  ATMO_TRY(validate_params(p));
  if (lambdas3 == nullptr || matrix9 == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  Spec s;
  build_spec(p, &s);
  if (i < 0 || i >= num_iterations(s))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "iteration out of range");
  iteration(s, i, lambdas3, matrix9);
  return ATMO_OK;
}

atmo_status atmo_convert_spectrum_to_linear_srgb(int n, const double* wl,
                                                 const double* spectrum, double* r,
                                                 double* g, double* b) {
  if (n < 1 || wl == nullptr || spectrum == nullptr || !r || !g || !b)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "bad arguments");
  std::vector<double> w(wl, wl + n), f(spectrum, spectrum + n);
  double x = 0.0, y = 0.0, z = 0.0;
  for (int lambda = kLambdaMin; lambda < kLambdaMax; ++lambda) {
    const double v = interpolate(w, f, lambda);
    x += cie_value(lambda, 1) * v;
    y += cie_value(lambda, 2) * v;
    z += cie_value(lambda, 3) * v;
  }
  const double* m = atmo_cie::kXyzToSrgb;
  const double k = atmo_cie::kMaxLuminousEfficacy;
  *r = k * (m[0] * x + m[1] * y + m[2] * z);
  *g = k * (m[3] * x + m[4] * y + m[5] * z);
  *b = k * (m[6] * x + m[7] * y + m[8] * z);
  return ATMO_OK;
}

atmo_status atmo_generate_sky_queries(uint64_t seed, int64_t n, double bottom_radius,
                                      double top_radius, double max_altitude,
                                      double max_distance, double max_sun_zenith_angle,
                                      float* camera, float* view_ray, float* sun_direction,
                                      float* point, float* normal, float* shadow_length,
                                      void* stream) {
  if (n < 0) return fail(ATMO_ERROR_INVALID_ARGUMENT, "n < 0");
  if (n > 0 && (camera == nullptr || view_ray == nullptr || sun_direction == nullptr ||
                point == nullptr || normal == nullptr || shadow_length == nullptr))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL output array");
  if (!(bottom_radius > 0.0 && top_radius > bottom_radius + 2e-3))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "need 0 < bottom_radius < top_radius");
  if (!(max_altitude >= 0.0 && max_distance >= 0.0 && max_sun_zenith_angle >= 0.0))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "negative range");
  return check_launch(
      atmo::launch_generate_queries(seed, n, bottom_radius, top_radius, max_altitude,
                                    max_distance, max_sun_zenith_angle, camera, view_ray,
                                    sun_direction, point, normal, shadow_length,
                                    static_cast<hipStream_t>(stream)),
      "generate sky queries");
}

atmo_status atmo_malloc(void** ptr, size_t bytes) {
  if (ptr == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "ptr is NULL");
  ATMO_HIP(hipMalloc(ptr, bytes));
  return ATMO_OK;
}
atmo_status atmo_free(void* ptr) {
  ATMO_HIP(hipFree(ptr));
  return ATMO_OK;
}
atmo_status atmo_memcpy_to_device(void* dst, const void* src, size_t bytes, void* stream) {
  ATMO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault,
                          static_cast<hipStream_t>(stream)));
  return ATMO_OK;
}
atmo_status atmo_memcpy_to_host(void* dst, const void* src, size_t bytes, void* stream) {
  ATMO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                          static_cast<hipStream_t>(stream)));
  ATMO_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return ATMO_OK;
}
atmo_status atmo_stream_create(void** stream) {
  if (stream == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "stream is NULL");
  hipStream_t s;
  ATMO_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return ATMO_OK;
}
atmo_status atmo_stream_destroy(void* stream) {
  ATMO_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return ATMO_OK;
}
atmo_status atmo_stream_synchronize(void* stream) {
  ATMO_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return ATMO_OK;
}
atmo_status atmo_device_synchronize(void) {
  ATMO_HIP(hipDeviceSynchronize());
  return ATMO_OK;
}
atmo_status atmo_event_create(void** event) {
  if (event == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "event is NULL");
  hipEvent_t e;
  ATMO_HIP(hipEventCreate(&e));
  *event = e;
  return ATMO_OK;
}
atmo_status atmo_event_destroy(void* event) {
  ATMO_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return ATMO_OK;
}
atmo_status atmo_event_record(void* event, void* stream) {
  ATMO_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return ATMO_OK;
}
atmo_status atmo_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (ms == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "ms is NULL");
  ATMO_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  ATMO_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                               static_cast<hipEvent_t>(stop)));
  return ATMO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched Init: one launch per pass level over every participating job of
// every model, the level sequence captured once into a HIP graph.
// ---------------------------------------------------------------------------

struct atmo_batch {
  std::vector<atmo_model*> models;
  std::vector<unsigned> orders;
  atmo_rt::LevelPlan plan;
  atmo::Job* d_jobs = nullptr;
  int* d_first_block = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
  atmo_rt::Fork fork;  // concurrent per-tile launches of one level
};

namespace atmo_rt {

void plan_levels(atmo_model* const* models, const unsigned* orders, int count,
                 LevelPlan* plan) {
  plan->jobs.clear();
  plan->first_block.clear();
  plan->levels.clear();
  auto add_one = [&](Pass p, int order, int tile,
                     const std::vector<std::pair<int, int>>& sel) {
    if (sel.empty()) return;
    Level l;
    l.pass = p;
    l.order = order;
    l.tile = tile;
    l.job_offset = static_cast<int>(plan->jobs.size());
    l.njobs = static_cast<int>(sel.size());
    l.fb_offset = static_cast<int>(plan->first_block.size());
    l.max_nnu = 0;
    l.max_ew = 0;
    l.max_w = 0;
    l.min_w = std::numeric_limits<int>::max();
    l.max_lds_w = 0;
    l.kinds = 0;
    l.dj_kinds = 0;
    int total = 0;
    for (const auto& mj : sel) {
      const atmo::Job& j = models[mj.first]->jobs[mj.second];
      plan->jobs.push_back(j);
      plan->first_block.push_back(total);
      total += pass_blocks(p, j);
      l.max_nnu = std::max(l.max_nnu, j.nnu);
      l.max_ew = std::max(l.max_ew, j.ew);
      l.max_w = std::max(l.max_w, j.nnu * j.nmus);
      l.min_w = std::min(l.min_w, j.nnu * j.nmus);
      if (j.nnu * j.nmus <= atmo::kMsMaxRow)
        l.max_lds_w = std::max(l.max_lds_w, j.nnu * j.nmus);
      l.kinds |= j.acc64 ? atmo::kJobsAcc64 : atmo::kJobsF32;
      l.dj_kinds |= j.half_luts ? atmo::kDjRgb21 : atmo::kDjF32;
    }
    plan->first_block.push_back(total);
    l.total_blocks = total;
    plan->levels.push_back(l);
  };
  // Scattering density and multiple scattering: one launch per tile shape
  // among the jobs (each kernel instance has one shape).
  auto add_level = [&](Pass p, int order, const std::vector<std::pair<int, int>>& sel) {
    if (p != kSD && p != kMS) {
      add_one(p, order, 0, sel);
      return;
    }
    const int ntiles = p == kSD ? atmo::kSdTiles : atmo::kMsTiles;
    for (int t = 0; t < ntiles; ++t) {
      std::vector<std::pair<int, int>> part;
      for (const auto& mj : sel) {
        const atmo::Job& j = models[mj.first]->jobs[mj.second];
        if ((p == kSD ? j.sd_tile : j.ms_tile) == t) part.push_back(mj);
      }
      add_one(p, order, t, part);
    }
  };
  // Every Precompute iteration of every model is a job of each level: the
  // iterations of one model have their own temporaries and accumulators.
  unsigned max_orders = 0;
  std::vector<std::pair<int, int>> sel;
  std::vector<std::pair<int, int>> od;
  for (int i = 0; i < count; ++i) {
    max_orders = std::max(max_orders, orders[i]);
    for (int it = 0; it < models[i]->iters; ++it) sel.push_back({i, it});
    if (models[i]->jobs[0].OD != nullptr) od.push_back({i, 0});
  }
  // The optical depths, once per model that shares them among its
  // transmittance LUTs.
  add_level(kOD, 0, od);
  add_level(kT, 0, sel);
  add_level(kE0, 0, sel);
  add_level(kSS, 0, sel);
  for (unsigned n = 2; n <= max_orders; ++n) {
    std::vector<std::pair<int, int>> on;
    for (const auto& mj : sel)
      if (orders[mj.first] >= n) on.push_back(mj);
    add_level(kSD, static_cast<int>(n), on);
    add_level(kII, static_cast<int>(n) - 1, on);
    add_level(kMS, 0, on);
  }
  // Job 'iters' of a model works on its resident LUTs.
  std::vector<std::pair<int, int>> rgb, half, acc, pack;
  for (int i = 0; i < count; ++i) {
    const int last = models[i]->iters;
    if (models[i]->spec.num_precomputed_wavelengths > 3) rgb.push_back({i, last});
    if (models[i]->spec.half_precision) half.push_back({i, last});
    if (accumulates_fp64(models[i]->spec)) acc.push_back({i, last});
    if (models[i]->Q != nullptr) pack.push_back({i, last});
  }
  add_level(kAcc, 0, acc);
  add_level(kT, 0, rgb);
  add_level(kClear, 0, half);
  add_level(kHalf, 0, half);
  add_level(kPack, 0, pack);
}

atmo_status create_fork(Fork* f) {
  for (hipStream_t& st : f->side)
    ATMO_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  for (hipEvent_t& e : f->ev) ATMO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return ATMO_OK;
}

void destroy_fork(Fork* f) {
  for (hipStream_t& st : f->side) {
    if (st != nullptr) (void)hipStreamDestroy(st);
    st = nullptr;
  }
  for (hipEvent_t& e : f->ev) {
    if (e != nullptr) (void)hipEventDestroy(e);
    e = nullptr;
  }
}

namespace {

atmo_status launch_level(const Level& l, const atmo::Job* d_jobs, const int* d_first_block,
                         hipStream_t s) {
  atmo::PassLaunch pl;
  pl.jobs = d_jobs + l.job_offset;
  pl.first_block = d_first_block + l.fb_offset;
  pl.njobs = l.njobs;
  pl.total_blocks = l.total_blocks;
  pl.max_nnu = l.max_nnu;
  pl.max_ew = l.max_ew;
  pl.max_w = l.max_w;
  pl.min_w = l.min_w;
  pl.max_lds_w = l.max_lds_w;
  pl.kinds = l.kinds;
  pl.dj_kinds = l.dj_kinds;
  pl.tile = l.tile;
  hipError_t e = hipSuccess;
  switch (l.pass) {
    case kT: e = atmo::launch_transmittance(pl, s); break;
    case kE0: e = atmo::launch_direct_irradiance(pl, s); break;
    case kSS: e = atmo::launch_single_scattering(pl, s); break;
    case kSD: e = atmo::launch_scattering_density(pl, l.order, s); break;
    case kII: e = atmo::launch_indirect_irradiance(pl, l.order, s); break;
    case kMS: e = atmo::launch_multiple_scattering(pl, s); break;
    case kClear: e = atmo::launch_clear_max(pl, s); break;
    case kAcc: e = atmo::launch_acc64_to_f32(pl, s); break;
    case kHalf: e = atmo::launch_to_half(pl, s); break;
    case kPack: e = atmo::launch_interleave(pl, s); break;
    case kOD: e = atmo::launch_optical_depth(pl, s); break;
    default: break;
  }
  return check_launch(e, "batched pass");
}

}  // namespace

atmo_status enqueue_levels(const LevelPlan& plan, const atmo::Job* d_jobs,
                           const int* d_first_block, hipStream_t s, const Fork* fork) {
  const std::vector<Level>& L = plan.levels;
  for (size_t i = 0; i < L.size();) {
    // [i, j): the launches of one level (same pass and order, one per tile
    // shape among its jobs; plan_levels emits them next to each other).
    size_t j = i + 1;
    while (j < L.size() && L[j].pass == L[i].pass && L[j].order == L[i].order &&
           (L[i].pass == kSD || L[i].pass == kMS))
      ++j;
    const int k = static_cast<int>(j - i);
    if (fork == nullptr || k == 1) {
      for (size_t q = i; q < j; ++q) ATMO_TRY(launch_level(L[q], d_jobs, d_first_block, s));
    } else {
      // Fork: side streams start after everything enqueued so far on s,
      // and s continues after all of them.
      ATMO_HIP(hipEventRecord(fork->ev[0], s));
      for (int q = 1; q < k; ++q) {
        hipStream_t side = fork->side[q - 1];
        ATMO_HIP(hipStreamWaitEvent(side, fork->ev[0], 0));
        ATMO_TRY(launch_level(L[i + q], d_jobs, d_first_block, side));
        ATMO_HIP(hipEventRecord(fork->ev[q], side));
      }
      ATMO_TRY(launch_level(L[i], d_jobs, d_first_block, s));
      for (int q = 1; q < k; ++q) ATMO_HIP(hipStreamWaitEvent(s, fork->ev[q], 0));
    }
    i = j;
  }
  return ATMO_OK;
}

}  // namespace atmo_rt

namespace {

void destroy_batch(atmo_batch* b) {
  if (b == nullptr) return;
  if (b->exec != nullptr) (void)hipGraphExecDestroy(b->exec);
  if (b->graph != nullptr) (void)hipGraphDestroy(b->graph);
  if (b->d_jobs != nullptr) (void)hipFree(b->d_jobs);
  if (b->d_first_block != nullptr) (void)hipFree(b->d_first_block);
  atmo_rt::destroy_fork(&b->fork);
  delete b;
}

}  // namespace

extern "C" {

atmo_status atmo_batch_create(atmo_model* const* models, const unsigned* orders,
                              int count, int use_graph, atmo_batch** out) {
  if (out == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (count < 1 || models == nullptr || orders == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "bad batch arguments");
  std::unique_ptr<atmo_batch, void (*)(atmo_batch*)> b(new atmo_batch(), destroy_batch);
  b->use_graph = use_graph != 0;
  for (int i = 0; i < count; ++i) {
    if (models[i] == nullptr || orders[i] < 1)
      return fail(ATMO_ERROR_INVALID_ARGUMENT, "bad model or order in batch");
    for (int k = 0; k < i; ++k)
      if (models[k] == models[i])
        return fail(ATMO_ERROR_INVALID_ARGUMENT, "a model appears twice in a batch");
    if (models[i]->live_batches > 0)
      return fail(ATMO_ERROR_INVALID_ARGUMENT, "a model is already in a live batch");
    if (models[i]->external_temporaries)
      return fail(ATMO_ERROR_INVALID_ARGUMENT, "a model is owned by a server");
    b->models.push_back(models[i]);
    b->orders.push_back(orders[i]);
  }
  // Temporaries and per-model job tables (the lookups use the latter).
  for (int i = 0; i < count; ++i) {
    ATMO_TRY(alloc_temporaries(models[i]));
    models[i]->last_orders = orders[i];
    ATMO_TRY(upload_jobs(models[i], orders[i], nullptr));
  }
  ATMO_HIP(hipStreamSynchronize(nullptr));
  plan_levels(models, orders, count, &b->plan);
  const LevelPlan& plan = b->plan;
  ATMO_HIP(hipMalloc(&b->d_jobs, plan.jobs.size() * sizeof(atmo::Job)));
  ATMO_HIP(hipMalloc(&b->d_first_block, plan.first_block.size() * sizeof(int)));
  ATMO_HIP(hipMemcpy(b->d_jobs, plan.jobs.data(), plan.jobs.size() * sizeof(atmo::Job),
                     hipMemcpyHostToDevice));
  ATMO_HIP(hipMemcpy(b->d_first_block, plan.first_block.data(),
                     plan.first_block.size() * sizeof(int), hipMemcpyHostToDevice));
  ATMO_TRY(create_fork(&b->fork));
  if (b->use_graph) {
    hipStream_t cap;
    ATMO_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    atmo_status st = ATMO_OK;
    if (e == hipSuccess)
      st = enqueue_levels(plan, b->d_jobs, b->d_first_block, cap, &b->fork);
    hipError_t e2 = hipStreamEndCapture(cap, &b->graph);
    (void)hipStreamDestroy(cap);
    if (e != hipSuccess || e2 != hipSuccess || st != ATMO_OK)
      return fail(ATMO_ERROR_DEVICE, std::string("graph capture failed: ") +
                                         hipGetErrorString(e != hipSuccess ? e : e2));
    ATMO_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  for (atmo_model* m : b->models) ++m->live_batches;
  *out = b.release();
  return ATMO_OK;
}

atmo_status atmo_batch_launch(atmo_batch* b, void* stream) {
  if (b == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "batch is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (b->use_graph) {
    ATMO_HIP(hipGraphLaunch(b->exec, s));
  } else {
    ATMO_TRY(enqueue_levels(b->plan, b->d_jobs, b->d_first_block, s, &b->fork));
  }
  for (atmo_model* m : b->models) m->initialized = true;
  return ATMO_OK;
}

int atmo_batch_num_launches(const atmo_batch* b) {
  return b == nullptr ? 0 : static_cast<int>(b->plan.levels.size());
}

void atmo_batch_destroy(atmo_batch* b) {
  if (b == nullptr) return;
  (void)hipDeviceSynchronize();
  for (atmo_model* m : b->models) {
    --m->live_batches;
    if (!m->retain && m->live_batches == 0) free_temporaries(m);
  }
  destroy_batch(b);
}

atmo_status atmo_model_init_batch(atmo_model* const* models, const unsigned* orders,
                                  int count, void* stream) {
  if (count == 0) return ATMO_OK;
  atmo_batch* b = nullptr;
  ATMO_TRY(atmo_batch_create(models, orders, count, 1, &b));
  atmo_status st = atmo_batch_launch(b, stream);
  if (st == ATMO_OK) {
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) st = fail(ATMO_ERROR_DEVICE, hipGetErrorString(e));
  }
  atmo_batch_destroy(b);
  return st;
}

}  // extern "C"
