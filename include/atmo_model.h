// C++ facade of libatmo with the interface of the reference's
// atmosphere::Model (atmosphere/model.h:160-337): same constructor
// parameters, Init(num_scattering_orders), ConvertSpectrumToLinearSrgb and
// kLambdaR/G/B. The GL-specific shader() and SetProgramUniforms() are
// replaced by device LUT views and batched lookup calls. Errors throw
// std::runtime_error (the reference asserts).
#ifndef ATMO_MODEL_H_
#define ATMO_MODEL_H_

#include <cstdint>
#include <vector>

#include "atmo.h"

namespace atmo {

// atmosphere/model.h:165-178
class DensityProfileLayer {
 public:
  DensityProfileLayer() : DensityProfileLayer(0.0, 0.0, 0.0, 0.0, 0.0) {}
  DensityProfileLayer(double width, double exp_term, double exp_scale,
                      double linear_term, double constant_term)
      : width(width), exp_term(exp_term), exp_scale(exp_scale),
        linear_term(linear_term), constant_term(constant_term) {}
  double width;
  double exp_term;
  double exp_scale;
  double linear_term;
  double constant_term;
};

// Per-model LUT resolution (compile-time constants in the reference,
// atmosphere/constants.h:47-61).
struct LutSizes {
  int transmittance_width = 256;
  int transmittance_height = 64;
  int scattering_r = 32;
  int scattering_mu = 128;
  int scattering_mu_s = 32;
  int scattering_nu = 8;
  int irradiance_width = 64;
  int irradiance_height = 16;
};

class Model {
 public:
  // atmosphere/model.h:182-281; 'sizes' is an extension (nullptr selects
  // the reference's sizes).
  Model(const std::vector<double>& wavelengths,
        const std::vector<double>& solar_irradiance,
        double sun_angular_radius, double bottom_radius, double top_radius,
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
        const LutSizes* sizes = nullptr);
  ~Model();
  Model(const Model&) = delete;
  Model& operator=(const Model&) = delete;

  // atmosphere/model.h:285
  void Init(unsigned int num_scattering_orders = 4);

  // Device LUT views (in place of shader()/SetProgramUniforms,
  // atmosphere/model.h:287-294).
  atmo_lut_view transmittance() const;
  atmo_lut_view scattering() const;
  atmo_lut_view single_mie_scattering() const;
  atmo_lut_view irradiance() const;

  // The shader API of atmosphere/model.cc:221-281 over device arrays of
  // n rays / points (n x 3 floats, lengths in the model's length unit).
  void GetSkyRadiance(int64_t n, const float* camera, const float* view_ray,
                      const float* shadow_length, const float* sun_direction,
                      float* radiance, float* transmittance,
                      void* stream = nullptr) const;
  void GetSkyRadianceToPoint(int64_t n, const float* camera,
                             const float* point, const float* shadow_length,
                             const float* sun_direction, float* radiance,
                             float* transmittance,
                             void* stream = nullptr) const;
  void GetSunAndSkyIrradiance(int64_t n, const float* point,
                              const float* normal, const float* sun_direction,
                              float* sun_irradiance, float* sky_irradiance,
                              void* stream = nullptr) const;
  void GetSolarRadiance(float* rgb) const;

  // atmosphere/model.h:301-304
  static void ConvertSpectrumToLinearSrgb(
      const std::vector<double>& wavelengths,
      const std::vector<double>& spectrum, double* r, double* g, double* b);

  // atmosphere/model.h:306-308
  static constexpr double kLambdaR = 680.0;
  static constexpr double kLambdaG = 550.0;
  static constexpr double kLambdaB = 440.0;

  atmo_model* handle() const { return model_; }

 private:
  atmo_model* model_ = nullptr;
};

}  // namespace atmo

#endif  // ATMO_MODEL_H_
