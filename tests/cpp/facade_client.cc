// A C++ caller of the atmo::Model facade (include/atmo_model.h), written the
// way the reference's demo builds its model (atmosphere/demo/demo.cc:173-284:
// 10 nm bins from 360 to 830 nm, a constant solar spectrum, Rayleigh
// lambda^-4, Angstrom Mie, an ozone tent profile, then
// Model(...) and Init()). It runs Init, one batch of each lookup, and writes
// what it got for the tests to compare with the Python path:
//   <out>/params.txt         the constructor parameters, one per line (%.17g)
//   <out>/<lut>.f32          each LUT as float32 RGBA (atmo_model_copy_lut)
//   <out>/lookups.f32        GetSkyRadiance, GetSkyRadianceToPoint and
//                            GetSunAndSkyIrradiance outputs (6 x n x 3)
//   <out>/solar.f32          GetSolarRadiance and ConvertSpectrumToLinearSrgb
//                            of the solar spectrum
//
//     facade_client <out dir> <num_scattering_orders> <num_rays> <seed>
//
// Build: g++ -std=c++17 -I include tests/cpp/facade_client.cc -L<pkg> -latmo
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "atmo_model.h"

namespace {

void write(const std::string& path, const void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (f == nullptr || std::fwrite(data, 1, bytes, f) != bytes)
    throw std::runtime_error("cannot write " + path);
  std::fclose(f);
}

void check(atmo_status s, const char* what) {
  if (s != ATMO_OK)
    throw std::runtime_error(std::string(what) + ": " + atmo_last_error());
}

std::vector<float> copy_lut(const atmo::Model& m, atmo_lut_id id, const atmo_lut_view& v) {
  std::vector<float> host(static_cast<size_t>(v.width) * v.height * v.depth * 4);
  check(atmo_model_copy_lut(m.handle(), id, host.data()), "copy_lut");
  return host;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s <out dir> <orders> <num_rays> <seed>\n", argv[0]);
    return 2;
  }
  const std::string out = argv[1];
  const unsigned orders = static_cast<unsigned>(std::atoi(argv[2]));
  const int64_t n = std::atoll(argv[3]);
  const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
  try {
    using atmo::DensityProfileLayer;
    constexpr double kPi = 3.14159265358979323846;
    constexpr double kSunAngularRadius = 0.00935 / 2.0;
    constexpr double kLengthUnitInMeters = 1000.0;
    constexpr double kConstantSolarIrradiance = 1.5;
    constexpr double kBottomRadius = 6360000.0;
    constexpr double kTopRadius = 6420000.0;
    constexpr double kRayleigh = 1.24062e-6;
    constexpr double kRayleighScaleHeight = 8000.0;
    constexpr double kMieScaleHeight = 1200.0;
    constexpr double kMieAngstromAlpha = 0.0;
    constexpr double kMieAngstromBeta = 5.328e-3;
    constexpr double kMieSingleScatteringAlbedo = 0.9;
    constexpr double kMiePhaseFunctionG = 0.8;
    constexpr double kGroundAlbedo = 0.1;
    // 300 Dobson units over the tent profile's 15 km integral, times a
    // smooth synthetic cross-section (m^2) peaking near 600 nm.
    constexpr double kMaxOzoneNumberDensity = 300.0 * 2.687e20 / 15000.0;
    const double max_sun_zenith_angle = 102.0 / 180.0 * kPi;

    DensityProfileLayer rayleigh_layer(0.0, 1.0, -1.0 / kRayleighScaleHeight, 0.0, 0.0);
    DensityProfileLayer mie_layer(0.0, 1.0, -1.0 / kMieScaleHeight, 0.0, 0.0);
    std::vector<DensityProfileLayer> ozone_density;
    ozone_density.push_back(DensityProfileLayer(25000.0, 0.0, 0.0, 1.0 / 15000.0, -2.0 / 3.0));
    ozone_density.push_back(DensityProfileLayer(0.0, 0.0, 0.0, -1.0 / 15000.0, 8.0 / 3.0));

    std::vector<double> wavelengths, solar_irradiance, rayleigh_scattering, mie_scattering,
        mie_extinction, absorption_extinction, ground_albedo;
    for (int l = 360; l <= 830; l += 10) {
      const double lambda = static_cast<double>(l) * 1e-3;  // micrometres
      const double mie = kMieAngstromBeta / kMieScaleHeight * std::pow(lambda, -kMieAngstromAlpha);
      const double x = (l - 600.0) / 60.0;
      wavelengths.push_back(l);
      solar_irradiance.push_back(kConstantSolarIrradiance);
      rayleigh_scattering.push_back(kRayleigh * std::pow(lambda, -4));
      mie_scattering.push_back(mie * kMieSingleScatteringAlbedo);
      mie_extinction.push_back(mie);
      absorption_extinction.push_back(kMaxOzoneNumberDensity * 4.5e-25 * std::exp(-x * x));
      ground_albedo.push_back(kGroundAlbedo);
    }

    // The constructor call of atmosphere/demo/demo.cc:278-284.
    atmo::Model model(wavelengths, solar_irradiance, kSunAngularRadius,
        kBottomRadius, kTopRadius, {rayleigh_layer}, rayleigh_scattering,
        {mie_layer}, mie_scattering, mie_extinction, kMiePhaseFunctionG,
        ozone_density, absorption_extinction, ground_albedo, max_sun_zenith_angle,
        kLengthUnitInMeters, 3, false, false);
    model.Init(orders);

    FILE* f = std::fopen((out + "/params.txt").c_str(), "w");
    if (f == nullptr) throw std::runtime_error("cannot write params.txt");
    auto vec = [&](const char* name, const std::vector<double>& v) {
      std::fprintf(f, "%s", name);
      for (double x : v) std::fprintf(f, " %.17g", x);
      std::fprintf(f, "\n");
    };
    vec("wavelengths", wavelengths);
    vec("solar_irradiance", solar_irradiance);
    vec("rayleigh_scattering", rayleigh_scattering);
    vec("mie_scattering", mie_scattering);
    vec("mie_extinction", mie_extinction);
    vec("absorption_extinction", absorption_extinction);
    vec("ground_albedo", ground_albedo);
    vec("scalars", {kSunAngularRadius, kBottomRadius, kTopRadius, kMiePhaseFunctionG,
                    max_sun_zenith_angle, kLengthUnitInMeters, -1.0 / kRayleighScaleHeight,
                    -1.0 / kMieScaleHeight, 1.0 / 15000.0, -2.0 / 3.0, -1.0 / 15000.0,
                    8.0 / 3.0});
    std::fclose(f);

    const struct {
      const char* name;
      atmo_lut_id id;
      atmo_lut_view view;
    } luts[] = {
        {"transmittance", ATMO_LUT_TRANSMITTANCE, model.transmittance()},
        {"scattering", ATMO_LUT_SCATTERING, model.scattering()},
        {"single_mie_scattering", ATMO_LUT_SINGLE_MIE_SCATTERING, model.single_mie_scattering()},
        {"irradiance", ATMO_LUT_IRRADIANCE, model.irradiance()},
    };
    for (const auto& l : luts) {
      const std::vector<float> h = copy_lut(model, l.id, l.view);
      write(out + "/" + l.name + ".f32", h.data(), h.size() * sizeof(float));
    }

    // One batch of each lookup on seeded rays (device arrays).
    const size_t b3 = static_cast<size_t>(n) * 3 * sizeof(float);
    float *cam, *view, *sun, *point, *normal, *shadow;
    float* res[6];
    check(atmo_malloc(reinterpret_cast<void**>(&cam), b3), "malloc");
    check(atmo_malloc(reinterpret_cast<void**>(&view), b3), "malloc");
    check(atmo_malloc(reinterpret_cast<void**>(&sun), b3), "malloc");
    check(atmo_malloc(reinterpret_cast<void**>(&point), b3), "malloc");
    check(atmo_malloc(reinterpret_cast<void**>(&normal), b3), "malloc");
    check(atmo_malloc(reinterpret_cast<void**>(&shadow), b3 / 3), "malloc");
    for (float*& r : res) check(atmo_malloc(reinterpret_cast<void**>(&r), b3), "malloc");
    check(atmo_generate_sky_queries(seed, n, kBottomRadius / kLengthUnitInMeters,
                                    kTopRadius / kLengthUnitInMeters, 100.0, 50.0,
                                    100.0 / 180.0 * kPi, cam, view, sun, point, normal,
                                    shadow, nullptr),
          "generate queries");
    model.GetSkyRadiance(n, cam, view, shadow, sun, res[0], res[1]);
    model.GetSkyRadianceToPoint(n, cam, point, shadow, sun, res[2], res[3]);
    model.GetSunAndSkyIrradiance(n, point, normal, sun, res[4], res[5]);
    std::vector<float> host(static_cast<size_t>(n) * 3 * 6);
    for (int k = 0; k < 6; ++k)
      check(atmo_memcpy_to_host(host.data() + static_cast<size_t>(k) * n * 3, res[k], b3,
                                nullptr),
            "copy");
    check(atmo_stream_synchronize(nullptr), "synchronize");
    write(out + "/lookups.f32", host.data(), host.size() * sizeof(float));
    for (float* p : {cam, view, sun, point, normal, shadow}) atmo_free(p);
    for (float* r : res) atmo_free(r);

    float solar[6];
    model.GetSolarRadiance(solar);
    double r, g, b;
    atmo::Model::ConvertSpectrumToLinearSrgb(wavelengths, solar_irradiance, &r, &g, &b);
    solar[3] = static_cast<float>(r);
    solar[4] = static_cast<float>(g);
    solar[5] = static_cast<float>(b);
    write(out + "/solar.f32", solar, sizeof(solar));
    std::printf("facade client: Init(%u) and %lld rays of each lookup done\n", orders,
                static_cast<long long>(n));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "facade client: %s\n", e.what());
    return 1;
  }
  return 0;
}
