// libatmo's host code under AddressSanitizer and UndefinedBehaviorSanitizer,
// on the paths that need no GPU: argument validation of atmo_model_create
// (every rejected parameter, and the cleanup of a model whose device
// allocation fails without a GPU), parameter lowering and illuminance
// iterations, the spectrum conversion, the FP16 channel scale, the tile
// table (set/get validation, save, and load of malformed files) and
// atmo_last_error. Exits 0 if every call returns the expected status; the
// sanitizers abort on the first error.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/atmo.h"
#include "../../precomputed_atmospheric_scattering_amd/csrc/atmo_device.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "unexpected: %s (%s)\n", what, atmo_last_error());
    ++failures;
  }
}

struct Params {
  std::vector<double> wl, solar, ray, mie_s, mie_e, abs, alb;
  atmo_density_layer r_layer{0.0, 1.0, -1.0 / 8000.0, 0.0, 0.0};
  atmo_density_layer m_layer{0.0, 1.0, -1.0 / 1200.0, 0.0, 0.0};
  atmo_density_layer oz[3] = {{25000.0, 0.0, 0.0, 1.0 / 15000.0, -2.0 / 3.0},
                              {0.0, 0.0, 0.0, -1.0 / 15000.0, 8.0 / 3.0},
                              {0.0, 0.0, 0.0, 0.0, 0.0}};
  atmo_model_params p{};
  explicit Params(int n, unsigned precomputed) {
    for (int i = 0; i < n; ++i) {
      const double l = 360.0 + 470.0 * i / (n - 1);
      wl.push_back(l);
      solar.push_back(1.5);
      ray.push_back(1.24062e-6 * std::pow(l * 1e-3, -4.0));
      mie_s.push_back(3.996e-6);
      mie_e.push_back(4.44e-6);
      abs.push_back(0.0);
      alb.push_back(0.1);
    }
    p.num_wavelengths = n;
    p.wavelengths = wl.data();
    p.solar_irradiance = solar.data();
    p.sun_angular_radius = 0.00935;
    p.bottom_radius = 6360e3;
    p.top_radius = 6420e3;
    p.num_rayleigh_layers = 1;
    p.rayleigh_density = &r_layer;
    p.rayleigh_scattering = ray.data();
    p.num_mie_layers = 1;
    p.mie_density = &m_layer;
    p.mie_scattering = mie_s.data();
    p.mie_extinction = mie_e.data();
    p.mie_phase_function_g = 0.8;
    p.num_absorption_layers = 2;
    p.absorption_density = oz;
    p.absorption_extinction = abs.data();
    p.ground_albedo = alb.data();
    p.max_sun_zenith_angle = 102.0 / 180.0 * 3.14159265358979323846;
    p.length_unit_in_meters = 1000.0;
    p.num_precomputed_wavelengths = precomputed;
  }
};

atmo_status create(const atmo_model_params& p, const atmo_lut_sizes* z) {
  atmo_model* m = nullptr;
  const atmo_status st = atmo_model_create(&p, z, &m);
  if (m != nullptr) atmo_model_destroy(m);
  expect(std::strlen(atmo_last_error()) < 4096, "bounded error message");
  return st;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string tmp = argc > 1 ? argv[1] : "/tmp";
  const double nan = std::numeric_limits<double>::quiet_NaN();

  // Rejected parameters: each returns ATMO_ERROR_INVALID_ARGUMENT before any
  // device call.
  {
    Params g(16, 3);
    expect(atmo_model_create(&g.p, nullptr, nullptr) == ATMO_ERROR_INVALID_ARGUMENT, "NULL out");
    expect(atmo_model_create(nullptr, nullptr, nullptr) == ATMO_ERROR_INVALID_ARGUMENT,
           "NULL params");
    auto bad = [&](void (*edit)(Params&), const char* what) {
      Params q(16, 3);
      edit(q);
      expect(create(q.p, nullptr) == ATMO_ERROR_INVALID_ARGUMENT, what);
    };
    bad([](Params& q) { q.p.top_radius = q.p.bottom_radius; }, "top <= bottom");
    bad([](Params& q) { q.p.sun_angular_radius = 0.0; }, "sun radius");
    bad([](Params& q) { q.p.length_unit_in_meters = -1.0; }, "length unit");
    bad([](Params& q) { q.p.num_wavelengths = 0; }, "no wavelengths");
    bad([](Params& q) { q.wl[3] = q.wl[2]; }, "wavelengths not increasing");
    bad([](Params& q) { q.p.num_rayleigh_layers = 3; }, "three layers");
    bad([](Params& q) { q.p.num_mie_layers = -1; }, "negative layers");
    bad([](Params& q) { q.p.rayleigh_density = nullptr; }, "NULL layers");
    bad([](Params& q) { q.p.mie_extinction = nullptr; }, "NULL spectrum");
    bad([](Params& q) { q.ray[5] = std::numeric_limits<double>::quiet_NaN(); }, "NaN");
    bad([](Params& q) { q.p.mie_phase_function_g = 1.0; }, "g = 1");
    const atmo_lut_sizes odd_mu{256, 64, 32, 127, 32, 8, 64, 16};
    expect(create(g.p, &odd_mu) == ATMO_ERROR_INVALID_ARGUMENT, "odd MU");
    const atmo_lut_sizes huge{4096, 4096, 4096, 4096, 4096, 32, 512, 64};
    expect(create(g.p, &huge) == ATMO_ERROR_INVALID_ARGUMENT, "2^28 texels");
    const atmo_lut_sizes tiny{1, 64, 32, 128, 32, 8, 64, 16};
    expect(create(g.p, &tiny) == ATMO_ERROR_INVALID_ARGUMENT, "size 1");
    const atmo_lut_sizes wide_e{256, 64, 32, 128, 32, 8, 1024, 16};
    expect(create(g.p, &wide_e) == ATMO_ERROR_INVALID_ARGUMENT, "irradiance width");
    // Valid parameters: without a GPU the device allocation fails and the
    // partly built model is released.
    const atmo_status st = create(g.p, nullptr);
    expect(st == ATMO_OK || st == ATMO_ERROR_DEVICE || st == ATMO_ERROR_OUT_OF_MEMORY,
           "valid parameters");
    expect(atmo_model_destroy(nullptr) == ATMO_OK, "destroy NULL");
    expect(atmo_model_init(nullptr, 4, nullptr) == ATMO_ERROR_INVALID_ARGUMENT, "init NULL");
    expect(atmo_model_init_static_scratch(nullptr, 4, nullptr) == ATMO_ERROR_INVALID_ARGUMENT,
           "static init NULL");
    (void)nan;
  }

  // Lowering, iterations and the spectrum conversion (host-only).
  for (unsigned pre : {3u, 15u, 16u, 47u}) {
    Params q(47, pre);
    const int it = atmo_num_iterations(&q.p);
    expect(it >= 1 && it <= 16, "iterations");
    for (int i = 0; i < it; ++i) {
      double l3[3];
      float m9[9];
      expect(atmo_iteration(&q.p, i, l3, m9) == ATMO_OK, "iteration");
      atmo_atmosphere a;
      expect(atmo_lower_parameters(&q.p, l3, &a) == ATMO_OK, "lower");
      expect(std::isfinite(a.rayleigh_scattering[2]) && a.mu_s_min < 0.0, "lowered values");
    }
    double l3[3];
    float m9[9];
    expect(atmo_iteration(&q.p, it, l3, m9) == ATMO_ERROR_INVALID_ARGUMENT, "iteration range");
    double r, g, b;
    expect(atmo_convert_spectrum_to_linear_srgb(47, q.wl.data(), q.solar.data(), &r, &g, &b) ==
               ATMO_OK && std::isfinite(r + g + b),
           "spectrum to sRGB");
  }
  {
    double r, g, b;
    const double w[1] = {550.0}, v[1] = {1.0};
    expect(atmo_convert_spectrum_to_linear_srgb(0, w, v, &r, &g, &b) != ATMO_OK, "empty spectrum");
    expect(atmo_convert_spectrum_to_linear_srgb(1, nullptr, v, &r, &g, &b) != ATMO_OK,
           "NULL spectrum");
  }

  // FP16 channel scales: finite, positive, powers of two, for every input.
  const float probes[] = {0.0f, -1.0f, 1e-45f, 1e-40f, 1.17549435e-38f, 1e-20f, 1.0f,
                          65504.0f, 3e38f, std::numeric_limits<float>::infinity(),
                          std::numeric_limits<float>::quiet_NaN()};
  for (float x : probes) {
    const float sc = atmo_fp16_channel_scale(x);
    int e = 0;
    expect(std::isfinite(sc) && sc > 0.0f && std::frexp(sc, &e) == 0.5f &&
               std::isfinite(1.0f / sc),
           "fp16 scale");
  }

  // 3 x 21-bit float texels (delta J of FP16 models): each channel keeps
  // its FP32 exponent and 12 mantissa bits rounded to nearest, so the
  // relative error is at most 2^-13 at any magnitude (normal FP32 values),
  // whatever the other channels; zero and signs are kept.
  {
    const float vals[] = {0.0f, 1.17549435e-38f, 3e-20f, 1e-7f, 0.37f, 1.0f, 1.9999999f,
                          6.5e4f, 1e30f, 3.0e38f, -2.5f, 0.1f, 0.3333333f};
    for (float x : vals)
      for (float y : vals)
        for (float z : {0.0f, 7.0e-12f, 123.456f}) {
          const float in[3] = {x, y, z};
          const float3 out = atmo::unpack_f21x3(atmo::pack_f21x3(x, y, z));
          const float o[3] = {out.x, out.y, out.z};
          for (int c = 0; c < 3; ++c)
            expect(std::fabs(o[c] - in[c]) <= std::fabs(in[c]) * 1.2207031e-4f &&
                       std::signbit(o[c]) == std::signbit(in[c]),
                   "3 x 21-bit float round trip");
        }
    // Rounding is to nearest: 1 + 2^-13 rounds to 1, 1 + 3 * 2^-13 up to
    // 1 + 2^-11, and a value already on the 12-bit grid is kept exactly.
    const float3 q = atmo::unpack_f21x3(atmo::pack_f21x3(1.0f + 0.00012207031f,
                                                        1.0f + 3.0f * 0.00012207031f,
                                                        1.0f + 0.00048828125f));
    expect(q.x == 1.0f && q.y == 1.0f + 0.00048828125f && q.z == 1.0f + 0.00048828125f,
           "3 x 21-bit rounding");
  }

  // Tile table: validation, save, and loads of malformed files.
  {
    atmo_tiles t{};
    const atmo_lut_sizes z{256, 64, 32, 128, 32, 8, 64, 16};
    expect(atmo_get_tiles(&z, &t) == ATMO_OK, "get tiles");
    expect(atmo_get_tiles(nullptr, &t) == ATMO_ERROR_INVALID_ARGUMENT, "get tiles NULL");
    atmo_tiles bad = t;
    bad.ms_row_order = 7;
    expect(atmo_set_tiles(&z, &bad) == ATMO_ERROR_INVALID_ARGUMENT, "bad row order");
    bad = t;
    bad.sd_block_threads = 96;
    expect(atmo_set_tiles(&z, &bad) == ATMO_ERROR_INVALID_ARGUMENT, "bad block");
    bad = t;
    bad.ms_block_threads = 1024;
    expect(atmo_set_tiles(&z, &bad) == ATMO_ERROR_INVALID_ARGUMENT, "bad ms block");
    t.ms_row_order = 2;
    expect(atmo_set_tiles(&z, &t) == ATMO_OK, "set tiles");
    const std::string good = tmp + "/tiles_good.txt";
    expect(atmo_tune_save(good.c_str()) == ATMO_OK, "save tiles");
    expect(atmo_tune_load(good.c_str()) == ATMO_OK, "load tiles");
    const char* bad_files[] = {
        "sd 64 4 3 256\n", "ms 32 128 8 4 1 256 9\n", "ms 128 8 4 1 256\n", "xx\n",
        "sd -5 4 4 256\n", "sd 99999999999999999999 4 4 256\n",
        "ms 32 128 32 8 1 256 0 trailing words and more words and more\n",
        "sd 64 4 4 256 ms 32 128 32 8 1 256 0\n", "ms 32 128 32 8 4 256 0\n"};
    int k = 0;
    for (const char* text : bad_files) {
      const std::string path = tmp + "/tiles_bad" + std::to_string(k++) + ".txt";
      FILE* f = std::fopen(path.c_str(), "w");
      std::fputs(text, f);
      std::fclose(f);
      expect(atmo_tune_load(path.c_str()) == ATMO_ERROR_INVALID_ARGUMENT, text);
    }
    // A line longer than the reader's buffer.
    const std::string path = tmp + "/tiles_long.txt";
    FILE* f = std::fopen(path.c_str(), "w");
    std::fputs(std::string(5000, '7').c_str(), f);
    std::fclose(f);
    expect(atmo_tune_load(path.c_str()) == ATMO_ERROR_INVALID_ARGUMENT, "long line");
    expect(atmo_tune_load((tmp + "/absent/tiles.txt").c_str()) == ATMO_ERROR_INVALID_ARGUMENT,
           "absent file");
    expect(atmo_tune_save(nullptr) == ATMO_ERROR_INVALID_ARGUMENT, "save NULL");
    atmo_tune_reset();
  }
  std::printf("capi driver: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
