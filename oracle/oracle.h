/*
 * FP64 CPU oracle for the atmosphere LUT precomputation and sky lookups.
 *
 * TEST INFRASTRUCTURE ONLY. This library gives the reference results that the
 * HIP product is compared against; it is loaded by tests/, by __graft_entry__.smoke() and
 * by the cpu_baseline leg of bench.py, and by nothing else. It is never on the
 * product path.
 *
 * It restates, in double precision and from the reference's text:
 *   - atmosphere/functions.glsl (every function on the precompute and lookup
 *     path; each definition in oracle.cc cites the lines it follows),
 *   - atmosphere/model.cc:535-595, 613-795, 866-975, 1048-1215 (parameter
 *     lowering, luminance factors, pass DAG, illuminance-mode iterations and
 *     the additive-blend output packing of the GL programs),
 *   - atmosphere/reference/model.cc:73-245 (per-texel CPU loop structure).
 * Texture reads follow the OpenGL 3.3 LINEAR + CLAMP_TO_EDGE rule with exact
 * FP64 weights (section 3.8.11 of the GL 3.3 specification).
 *
 * LUT sizes are runtime values. Every LUT is an array of 4-channel doubles,
 * texel-major with x fastest: element (x, y, z, c) is at
 * ((z * height + y) * width + x) * 4 + c. The packed scattering layout is
 * x = nu_index * MU_S + mu_s_index, y = mu_index, z = r_index.
 */
#ifndef ATMO_ORACLE_H_
#define ATMO_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_layer {
  double width, exp_term, exp_scale, linear_term, constant_term;
} or_layer;

/* AtmosphereParameters (atmosphere/definitions.glsl:205-255) at 3 channels,
 * lengths already in the model's length unit. */
typedef struct or_atmosphere {
  double solar_irradiance[3];
  double sun_angular_radius;
  double bottom_radius;
  double top_radius;
  or_layer rayleigh_density[2];
  double rayleigh_scattering[3];
  or_layer mie_density[2];
  double mie_scattering[3];
  double mie_extinction[3];
  double mie_phase_function_g;
  or_layer absorption_density[2];
  double absorption_extinction[3];
  double ground_albedo[3];
  double mu_s_min;
} or_atmosphere;

/* LUT resolution (atmosphere/constants.h:47-61 are the defaults). */
typedef struct or_sizes {
  int transmittance_width, transmittance_height;
  int scattering_r, scattering_mu, scattering_mu_s, scattering_nu;
  int irradiance_width, irradiance_height;
} or_sizes;

/* The Model constructor parameters (atmosphere/model.h:182-281). */
typedef struct or_model_params {
  int num_wavelengths;
  const double* wavelengths;
  const double* solar_irradiance;
  double sun_angular_radius;
  double bottom_radius;
  double top_radius;
  int num_rayleigh_layers;
  const or_layer* rayleigh_density;
  const double* rayleigh_scattering;
  int num_mie_layers;
  const or_layer* mie_density;
  const double* mie_scattering;
  const double* mie_extinction;
  double mie_phase_function_g;
  int num_absorption_layers;
  const or_layer* absorption_density;
  const double* absorption_extinction;
  const double* ground_albedo;
  double max_sun_zenith_angle;
  double length_unit_in_meters;
  unsigned num_precomputed_wavelengths;
  int combine_scattering_textures;
  int half_precision;
} or_model_params;

void or_set_threads(int n);
int or_get_threads(void);

/* ---- scalar functions (functions.glsl), for closed-form tests ---- */
double or_distance_to_top(const or_atmosphere* a, double r, double mu);
double or_distance_to_bottom(const or_atmosphere* a, double r, double mu);
int or_ray_intersects_ground(const or_atmosphere* a, double r, double mu);
double or_distance_to_nearest(const or_atmosphere* a, double r, double mu,
                              int ground);
double or_profile_density(const or_layer* two_layers, double altitude);
double or_optical_length(const or_atmosphere* a, const or_layer* two_layers,
                         double r, double mu);
void or_transmittance_to_top(const or_atmosphere* a, double r, double mu,
                             double* out3);
double or_tex_coord_from_unit_range(double x, int n);
double or_unit_range_from_tex_coord(double u, int n);
void or_transmittance_uv_from_rmu(const or_atmosphere* a, const or_sizes* s,
                                  double r, double mu, double* out_uv);
void or_rmu_from_transmittance_uv(const or_atmosphere* a, const or_sizes* s,
                                  double u, double v, double* r, double* mu);
void or_get_transmittance_to_top(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, double r, double mu,
                                 double* out3);
void or_get_transmittance(const or_atmosphere* a, const or_sizes* s,
                          const double* t_lut, double r, double mu, double d,
                          int ground, double* out3);
void or_get_transmittance_to_sun(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, double r, double mu_s,
                                 double* out3);
void or_single_scattering_integrand(const or_atmosphere* a, const or_sizes* s,
                                    const double* t_lut, double r, double mu,
                                    double mu_s, double nu, double d,
                                    int ground, double* rayleigh3,
                                    double* mie3);
void or_single_scattering(const or_atmosphere* a, const or_sizes* s,
                          const double* t_lut, double r, double mu,
                          double mu_s, double nu, int ground,
                          double* rayleigh3, double* mie3);
double or_rayleigh_phase(double nu);
double or_mie_phase(double g, double nu);
void or_scattering_uvwz_from_rmumusnu(const or_atmosphere* a,
                                      const or_sizes* s, double r, double mu,
                                      double mu_s, double nu, int ground,
                                      double* out_uvwz);
void or_rmumusnu_from_scattering_uvwz(const or_atmosphere* a,
                                      const or_sizes* s, const double* uvwz,
                                      double* out_r_mu_mus_nu, int* ground);
void or_rmumusnu_from_scattering_texel(const or_atmosphere* a,
                                       const or_sizes* s, int x, int y, int z,
                                       double* out_r_mu_mus_nu, int* ground);
void or_get_scattering(const or_atmosphere* a, const or_sizes* s,
                       const double* lut, double r, double mu, double mu_s,
                       double nu, int ground, double* out4);
void or_scattering_density(const or_atmosphere* a, const or_sizes* s,
                           const double* t_lut, const double* single_rayleigh,
                           const double* single_mie,
                           const double* multiple_scattering,
                           const double* irradiance, double r, double mu,
                           double mu_s, double nu, int order, double* out3);
void or_multiple_scattering(const or_atmosphere* a, const or_sizes* s,
                            const double* t_lut, const double* density_lut,
                            double r, double mu, double mu_s, double nu,
                            int ground, double* out3);
void or_direct_irradiance(const or_atmosphere* a, const or_sizes* s,
                          const double* t_lut, double r, double mu_s,
                          double* out3);
void or_indirect_irradiance(const or_atmosphere* a, const or_sizes* s,
                            const double* single_rayleigh,
                            const double* single_mie,
                            const double* multiple_scattering, double r,
                            double mu_s, int order, double* out3);
void or_irradiance_uv_from_rmus(const or_atmosphere* a, const or_sizes* s,
                                double r, double mu_s, double* out_uv);
void or_rmus_from_irradiance_uv(const or_atmosphere* a, const or_sizes* s,
                                double u, double v, double* r, double* mu_s);
void or_get_irradiance(const or_atmosphere* a, const or_sizes* s,
                       const double* e_lut, double r, double mu_s,
                       double* out3);

/* ---- whole-LUT passes (the six GL programs, atmosphere/model.cc:125-208,
 * without the luminance matrix / blending epilogue) ---- */
void or_pass_transmittance(const or_atmosphere* a, const or_sizes* s,
                           double* t_out);
void or_pass_direct_irradiance(const or_atmosphere* a, const or_sizes* s,
                               const double* t_lut, double* delta_e_out);
void or_pass_single_scattering(const or_atmosphere* a, const or_sizes* s,
                               const double* t_lut, double* delta_rayleigh_out,
                               double* delta_mie_out);
void or_pass_scattering_density(const or_atmosphere* a, const or_sizes* s,
                                const double* t_lut,
                                const double* delta_rayleigh,
                                const double* delta_mie,
                                const double* delta_multiple,
                                const double* delta_irradiance, int order,
                                double* delta_j_out);
void or_pass_indirect_irradiance(const or_atmosphere* a, const or_sizes* s,
                                 const double* delta_rayleigh,
                                 const double* delta_mie,
                                 const double* delta_multiple, int order,
                                 double* delta_e_out);
/* nu_out (may be NULL) receives the clamped nu of every texel. */
void or_pass_multiple_scattering(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, const double* delta_j,
                                 double* delta_multiple_out, double* nu_out);

/* Same passes on a list of texels only (pass-local checks). Texel i is at
 * flat index idx[i] = x + W * (y + H * z); out gets 4 doubles per texel. */
void or_texels_single_scattering(const or_atmosphere* a, const or_sizes* s,
                                 const double* t_lut, int n, const int* idx,
                                 double* rayleigh_out, double* mie_out);
void or_texels_scattering_density(const or_atmosphere* a, const or_sizes* s,
                                  const double* t_lut,
                                  const double* delta_rayleigh,
                                  const double* delta_mie,
                                  const double* delta_multiple,
                                  const double* delta_irradiance, int order,
                                  int n, const int* idx, double* out);
void or_texels_multiple_scattering(const or_atmosphere* a, const or_sizes* s,
                                   const double* t_lut, const double* delta_j,
                                   int n, const int* idx, double* out);

/* ---- model level (atmosphere/model.cc) ---- */
double or_interpolate(int n, const double* wavelengths, const double* values,
                      double lambda);
double or_cie_table_value(double lambda, int column);
void or_luminance_factors(int n, const double* wavelengths,
                          const double* solar, double lambda_power,
                          double* k3);
void or_convert_spectrum_to_linear_srgb(int n, const double* wavelengths,
                                        const double* spectrum, double* rgb3);
/* Lower the constructor parameters to AtmosphereParameters at 3 wavelengths
 * (atmosphere/model.cc:636-666, 718-734), without text quantization. */
void or_lower(const or_model_params* p, const double* lambdas3,
              or_atmosphere* out);
/* The 3 wavelengths and luminance_from_radiance matrix (row-major, doubles)
 * of illuminance-mode iteration i (atmosphere/model.cc:914-943), and the
 * iteration count. */
int or_num_iterations(const or_model_params* p);
void or_iteration(const or_model_params* p, int i, double* lambdas3,
                  double* matrix9);

/* Model::Init (atmosphere/model.cc:866-975). Outputs are 4-channel LUTs.
 * single_mie_out is left untouched when combine_scattering_textures is set.
 * Returns 0 on success. */
int or_model_init(const or_model_params* p, const or_sizes* s,
                  unsigned num_scattering_orders, double* transmittance_out,
                  double* scattering_out, double* single_mie_out,
                  double* irradiance_out);
/* Init(n) for every n = 1..orders in one run: scattering[n-1] etc. receive
 * the LUTs of Init(n) (arrays of 'orders' pointers). */
int or_model_init_orders(const or_model_params* p, const or_sizes* s,
                         unsigned orders, double* transmittance,
                         double* const* scattering, double* const* single_mie,
                         double* const* irradiance);
/* One Model::Precompute call (atmosphere/model.cc:1048-1215) on
 * caller-provided LUTs, so illuminance iterations can be checked one by one.
 * matrix9 is luminance_from_radiance (row-major). */
void or_precompute(const or_atmosphere* a, const or_sizes* s,
                   const double* matrix9, int blend,
                   unsigned num_scattering_orders, int combined,
                   double* transmittance, double* scattering,
                   double* single_mie, double* irradiance);

/* ---- sky-shading lookups (functions.glsl:1633-1896) ----
 * Arrays are n x 3 doubles, lengths in the model's length unit. */
void or_sky_radiance(const or_atmosphere* a, const or_sizes* s,
                     const double* t_lut, const double* scattering,
                     const double* single_mie, int combined, int n,
                     const double* camera, const double* view_ray,
                     const double* shadow_length, const double* sun_direction,
                     double* radiance_out, double* transmittance_out);
void or_sky_radiance_to_point(const or_atmosphere* a, const or_sizes* s,
                              const double* t_lut, const double* scattering,
                              const double* single_mie, int combined, int n,
                              const double* camera, const double* point,
                              const double* shadow_length,
                              const double* sun_direction,
                              double* radiance_out,
                              double* transmittance_out);
void or_sun_and_sky_irradiance(const or_atmosphere* a, const or_sizes* s,
                               const double* t_lut, const double* e_lut,
                               int n, const double* point,
                               const double* normal,
                               const double* sun_direction,
                               double* sun_irradiance_out,
                               double* sky_irradiance_out);

#ifdef __cplusplus
}
#endif

#endif  /* ATMO_ORACLE_H_ */
