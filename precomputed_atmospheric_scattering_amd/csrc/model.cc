// C++ facade atmo::Model over the C ABI (include/atmo_model.h).
#include "atmo_model.h"

#include <stdexcept>
#include <string>

namespace atmo {
namespace {

void check(atmo_status s, const char* what) {
  if (s != ATMO_OK)
    throw std::runtime_error(std::string(what) + ": " + atmo_last_error());
}

std::vector<atmo_density_layer> layers(const std::vector<DensityProfileLayer>& v) {
  std::vector<atmo_density_layer> out;
  for (const DensityProfileLayer& l : v)
    out.push_back({l.width, l.exp_term, l.exp_scale, l.linear_term, l.constant_term});
  return out;
}

void same_size(const std::vector<double>& a, const std::vector<double>& b,
               const char* name) {
  if (a.size() != b.size())
    throw std::invalid_argument(std::string(name) +
                                " must have one value per wavelength");
}

}  // namespace

Model::Model(const std::vector<double>& wavelengths,
             const std::vector<double>& solar_irradiance,
             double sun_angular_radius, double bottom_radius,
             double top_radius,
             const std::vector<DensityProfileLayer>& rayleigh_density,
             const std::vector<double>& rayleigh_scattering,
             const std::vector<DensityProfileLayer>& mie_density,
             const std::vector<double>& mie_scattering,
             const std::vector<double>& mie_extinction,
             double mie_phase_function_g,
             const std::vector<DensityProfileLayer>& absorption_density,
             const std::vector<double>& absorption_extinction,
             const std::vector<double>& ground_albedo,
             double max_sun_zenith_angle, double length_unit_in_meters,
             unsigned int num_precomputed_wavelengths,
             bool combine_scattering_textures, bool half_precision,
             const LutSizes* sizes) {
  same_size(wavelengths, solar_irradiance, "solar_irradiance");
  same_size(wavelengths, rayleigh_scattering, "rayleigh_scattering");
  same_size(wavelengths, mie_scattering, "mie_scattering");
  same_size(wavelengths, mie_extinction, "mie_extinction");
  same_size(wavelengths, absorption_extinction, "absorption_extinction");
  same_size(wavelengths, ground_albedo, "ground_albedo");
  std::vector<atmo_density_layer> r = layers(rayleigh_density),
                                   m = layers(mie_density),
                                   a = layers(absorption_density);
  atmo_model_params p;
  p.num_wavelengths = static_cast<int>(wavelengths.size());
  p.wavelengths = wavelengths.data();
  p.solar_irradiance = solar_irradiance.data();
  p.sun_angular_radius = sun_angular_radius;
  p.bottom_radius = bottom_radius;
  p.top_radius = top_radius;
  p.num_rayleigh_layers = static_cast<int>(r.size());
  p.rayleigh_density = r.data();
  p.rayleigh_scattering = rayleigh_scattering.data();
  p.num_mie_layers = static_cast<int>(m.size());
  p.mie_density = m.data();
  p.mie_scattering = mie_scattering.data();
  p.mie_extinction = mie_extinction.data();
  p.mie_phase_function_g = mie_phase_function_g;
  p.num_absorption_layers = static_cast<int>(a.size());
  p.absorption_density = a.data();
  p.absorption_extinction = absorption_extinction.data();
  p.ground_albedo = ground_albedo.data();
  p.max_sun_zenith_angle = max_sun_zenith_angle;
  p.length_unit_in_meters = length_unit_in_meters;
  p.num_precomputed_wavelengths = num_precomputed_wavelengths;
  p.combine_scattering_textures = combine_scattering_textures ? 1 : 0;
  p.half_precision = half_precision ? 1 : 0;
  atmo_lut_sizes z;
  const atmo_lut_sizes* zp = nullptr;
  if (sizes != nullptr) {
    z = {sizes->transmittance_width, sizes->transmittance_height,
         sizes->scattering_r,        sizes->scattering_mu,
         sizes->scattering_mu_s,     sizes->scattering_nu,
         sizes->irradiance_width,    sizes->irradiance_height};
    zp = &z;
  }
  check(atmo_model_create(&p, zp, &model_), "Model");
}

Model::~Model() { atmo_model_destroy(model_); }

void Model::Init(unsigned int num_scattering_orders) {
  check(atmo_model_init(model_, num_scattering_orders, nullptr), "Init");
}

static atmo_lut_view view(const atmo_model* m, atmo_lut_id id) {
  atmo_lut_view v;
  check(atmo_model_get_lut(m, id, &v), "LUT view");
  return v;
}

atmo_lut_view Model::transmittance() const {
  return view(model_, ATMO_LUT_TRANSMITTANCE);
}
atmo_lut_view Model::scattering() const {
  return view(model_, ATMO_LUT_SCATTERING);
}
atmo_lut_view Model::single_mie_scattering() const {
  return view(model_, ATMO_LUT_SINGLE_MIE_SCATTERING);
}
atmo_lut_view Model::irradiance() const {
  return view(model_, ATMO_LUT_IRRADIANCE);
}

void Model::GetSkyRadiance(int64_t n, const float* camera,
                           const float* view_ray, const float* shadow_length,
                           const float* sun_direction, float* radiance,
                           float* transmittance, void* stream) const {
  check(atmo_get_sky_radiance(model_, n, camera, view_ray, shadow_length,
                              sun_direction, radiance, transmittance, stream),
        "GetSkyRadiance");
}

void Model::GetSkyRadianceToPoint(int64_t n, const float* camera,
                                  const float* point,
                                  const float* shadow_length,
                                  const float* sun_direction, float* radiance,
                                  float* transmittance, void* stream) const {
  check(atmo_get_sky_radiance_to_point(model_, n, camera, point, shadow_length,
                                       sun_direction, radiance, transmittance,
                                       stream),
        "GetSkyRadianceToPoint");
}

void Model::GetSunAndSkyIrradiance(int64_t n, const float* point,
                                   const float* normal,
                                   const float* sun_direction,
                                   float* sun_irradiance,
                                   float* sky_irradiance, void* stream) const {
  check(atmo_get_sun_and_sky_irradiance(model_, n, point, normal, sun_direction,
                                        sun_irradiance, sky_irradiance, stream),
        "GetSunAndSkyIrradiance");
}

void Model::GetSolarRadiance(float* rgb) const {
  check(atmo_get_solar_radiance(model_, rgb), "GetSolarRadiance");
}

void Model::ConvertSpectrumToLinearSrgb(const std::vector<double>& wavelengths,
                                        const std::vector<double>& spectrum,
                                        double* r, double* g, double* b) {
  same_size(wavelengths, spectrum, "spectrum");
  check(atmo_convert_spectrum_to_linear_srgb(
            static_cast<int>(wavelengths.size()), wavelengths.data(),
            spectrum.data(), r, g, b),
        "ConvertSpectrumToLinearSrgb");
}

}  // namespace atmo
