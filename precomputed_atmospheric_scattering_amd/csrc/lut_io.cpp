// LUT files of libatmo (include/atmo.h, atmo_model_export_luts and friends):
// raw float32 RGBA texels, the format of the reference's WebGL exporter
// (atmosphere/demo/webgl/precompute.cc:63-106), under the file names of the
// reference CPU model's cache (atmosphere/reference/model.cc:74-84, :240-244).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "runtime_internal.h"

using namespace atmo_rt;

namespace {

struct LutFile {
  atmo_lut_id id;
  const char* name;
};
constexpr LutFile kFiles[] = {
    {ATMO_LUT_TRANSMITTANCE, "transmittance.dat"},
    {ATMO_LUT_SCATTERING, "scattering.dat"},
    {ATMO_LUT_SINGLE_MIE_SCATTERING, "single_mie_scattering.dat"},
    {ATMO_LUT_IRRADIANCE, "irradiance.dat"},
};

std::string path_of(const char* dir, const char* name) {
  std::string p(dir);
  if (!p.empty() && p.back() != '/') p.push_back('/');
  return p + name;
}

size_t texels_of(const atmo_model* m, atmo_lut_id id) {
  const Sizes& z = m->sizes;
  switch (id) {
    case ATMO_LUT_TRANSMITTANCE: return z.t_texels();
    case ATMO_LUT_IRRADIANCE: return z.e_texels();
    default: return z.s_texels();
  }
}

bool exported(const atmo_model* m, atmo_lut_id id) {
  return id != ATMO_LUT_SINGLE_MIE_SCATTERING || !m->spec.combined;
}

// The FP16 scale of a channel whose largest magnitude is mx: a power of two
// that maps mx into [2^13, 2^14), with the exponent capped at 126 (the rule
// of k_to_half).
float pow2_scale(float mx) {
  if (!(mx > 0.0f) || std::isinf(mx)) return 1.0f;
  int e;
  std::frexp(mx, &e);
  return std::ldexp(1.0f, std::min(14 - e, 126));
}

uint2 pack_half4(const float* v) {
  _Float16 h[4] = {static_cast<_Float16>(v[0]), static_cast<_Float16>(v[1]),
                   static_cast<_Float16>(v[2]), static_cast<_Float16>(v[3])};
  uint2 u;
  static_assert(sizeof(h) == sizeof(u), "4 halves in 8 bytes");
  __builtin_memcpy(&u, h, sizeof(u));
  return u;
}

atmo_status read_file(const std::string& path, std::vector<float>* out, size_t floats) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (f == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "cannot open " + path);
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (bytes < 0 || static_cast<size_t>(bytes) != floats * sizeof(float)) {
    std::fclose(f);
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                path + ": " + std::to_string(bytes) + " bytes, the model's LUT has " +
                    std::to_string(floats * sizeof(float)));
  }
  out->resize(floats);
  const size_t got = std::fread(out->data(), sizeof(float), floats, f);
  std::fclose(f);
  if (got != floats) return fail(ATMO_ERROR_INVALID_ARGUMENT, "short read of " + path);
  return ATMO_OK;
}

}  // namespace

extern "C" {

atmo_status atmo_model_export_luts(const atmo_model* m, const char* directory) {
  if (m == nullptr || directory == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  if (!m->initialized) return fail(ATMO_ERROR_NOT_INITIALIZED, "model not initialized");
  for (const LutFile& lf : kFiles) {
    if (!exported(m, lf.id)) continue;
    std::vector<float> host(texels_of(m, lf.id) * 4);
    ATMO_TRY(atmo_model_copy_lut(m, lf.id, host.data()));
    const std::string path = path_of(directory, lf.name);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (f == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "cannot create " + path);
    const size_t put = std::fwrite(host.data(), sizeof(float), host.size(), f);
    const bool ok = std::fclose(f) == 0 && put == host.size();
    if (!ok) return fail(ATMO_ERROR_INVALID_ARGUMENT, "cannot write " + path);
  }
  return ATMO_OK;
}

atmo_status atmo_model_import_luts(atmo_model* m, const char* directory) {
  if (m == nullptr || directory == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  std::vector<float> t, s, sm, e;
  ATMO_TRY(read_file(path_of(directory, "transmittance.dat"), &t,
                     texels_of(m, ATMO_LUT_TRANSMITTANCE) * 4));
  ATMO_TRY(read_file(path_of(directory, "scattering.dat"), &s,
                     texels_of(m, ATMO_LUT_SCATTERING) * 4));
  if (!m->spec.combined)
    ATMO_TRY(read_file(path_of(directory, "single_mie_scattering.dat"), &sm,
                       texels_of(m, ATMO_LUT_SINGLE_MIE_SCATTERING) * 4));
  ATMO_TRY(read_file(path_of(directory, "irradiance.dat"), &e,
                     texels_of(m, ATMO_LUT_IRRADIANCE) * 4));
  auto upload = [](void* dst, const void* src, size_t bytes) -> atmo_status {
    const hipError_t err = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    if (err != hipSuccess)
      return fail(ATMO_ERROR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(err));
    return ATMO_OK;
  };
  ATMO_TRY(upload(m->T, t.data(), t.size() * sizeof(float)));
  ATMO_TRY(upload(m->E, e.data(), e.size() * sizeof(float)));
  if (!m->spec.half_precision) {
    ATMO_TRY(upload(m->S, s.data(), s.size() * sizeof(float)));
    if (!m->spec.combined) ATMO_TRY(upload(m->SM, sm.data(), sm.size() * sizeof(float)));
  } else {
    // Per-channel scales from the largest magnitudes, as k_lut_max and
    // k_to_half do on the FP32 accumulators: scales[0..3] for S,
    // scales[4..6] for SM.
    float mx[8] = {};
    for (size_t i = 0; i < s.size(); ++i) mx[i % 4] = std::fmax(mx[i % 4], std::fabs(s[i]));
    for (size_t i = 0; i < sm.size(); ++i)
      if (i % 4 < 3) mx[4 + i % 4] = std::fmax(mx[4 + i % 4], std::fabs(sm[i]));
    float sc[8];
    for (int c = 0; c < 8; ++c) sc[c] = pow2_scale(mx[c]);
    const size_t n = s.size() / 4;
    std::vector<uint2> packed(n);
    for (size_t i = 0; i < n; ++i) {
      const float v[4] = {s[4 * i] * sc[0], s[4 * i + 1] * sc[1], s[4 * i + 2] * sc[2],
                          s[4 * i + 3] * sc[3]};
      packed[i] = pack_half4(v);
    }
    ATMO_TRY(upload(m->S16, packed.data(), n * sizeof(uint2)));
    if (!m->spec.combined) {
      for (size_t i = 0; i < n; ++i) {
        const float v[4] = {sm[4 * i] * sc[4], sm[4 * i + 1] * sc[5], sm[4 * i + 2] * sc[6],
                            0.0f};
        packed[i] = pack_half4(v);
      }
      ATMO_TRY(upload(m->SM16, packed.data(), n * sizeof(uint2)));
    }
    ATMO_TRY(upload(m->d_scales, sc, sizeof(sc)));
  }
  ATMO_TRY(build_lookup_copy(m, nullptr));
  const hipError_t err = hipStreamSynchronize(nullptr);
  if (err != hipSuccess)
    return fail(ATMO_ERROR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(err));
  m->initialized = true;
  return ATMO_OK;
}

atmo_status atmo_model_init_cached(atmo_model* m, unsigned orders, const char* directory,
                                   void* stream) {
  if (m == nullptr || directory == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  FILE* probe = std::fopen(path_of(directory, "transmittance.dat").c_str(), "rb");
  if (probe != nullptr) {
    std::fclose(probe);
    return atmo_model_import_luts(m, directory);
  }
  ATMO_TRY(atmo_model_init(m, orders, stream));
  const hipError_t err = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (err != hipSuccess)
    return fail(ATMO_ERROR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(err));
  return atmo_model_export_luts(m, directory);
}

float atmo_fp16_channel_scale(float max_abs) { return pow2_scale(max_abs); }

}  // extern "C"
