/*
 * atmo: MI355X (gfx950) precomputation of Bruneton-style atmospheric
 * scattering LUTs and the sky-shading lookups that read them.
 *
 * C ABI of libatmo.so. It replaces the OpenGL-backed atmosphere::Model of the
 * reference (atmosphere/model.h:160-337, atmosphere/model.cc:613-1215) and the
 * GLSL rendering API it exposes (atmosphere/model.h:51-89,
 * atmosphere/model.cc:221-281). Conventions:
 *   - every entry point returns an atmo_status; atmo_last_error() gives the
 *     message of the last failure on the calling thread;
 *   - a model handle owns its device memory; one handle must not be used from
 *     two threads at once (distinct handles are independent);
 *   - 'stream' is a hipStream_t (NULL = the default stream); all device work
 *     is enqueued on it and the call returns without waiting, except where a
 *     function says otherwise;
 *   - array arguments of the lookup functions are device pointers, n x 3
 *     floats (AoS) unless stated; lengths are in the model's length unit, as
 *     in the reference's shader API.
 */
#ifndef ATMO_H_
#define ATMO_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum atmo_status {
  ATMO_OK = 0,
  ATMO_ERROR_INVALID_ARGUMENT = 1,
  ATMO_ERROR_OUT_OF_MEMORY = 2,
  ATMO_ERROR_DEVICE = 3,       /* a HIP runtime call failed */
  ATMO_ERROR_NOT_INITIALIZED = 4,
  ATMO_ERROR_UNSUPPORTED = 5,
} atmo_status;

/* DensityProfileLayer (atmosphere/model.h:165-178): width in m, exp_scale and
 * linear_term in 1/m; density = clamp(exp_term * exp(exp_scale * h) +
 * linear_term * h + constant_term, 0, 1). */
typedef struct atmo_density_layer {
  double width;
  double exp_term;
  double exp_scale;
  double linear_term;
  double constant_term;
} atmo_density_layer;

/* The 17 Model constructor parameters (atmosphere/model.h:182-281), same
 * meaning and units. Spectra have num_wavelengths samples; layer lists have
 * at most 2 entries (fewer are padded with null layers at the front, as
 * atmosphere/model.cc:653-658 does). */
typedef struct atmo_model_params {
  int num_wavelengths;
  const double* wavelengths;          /* nm, increasing */
  const double* solar_irradiance;     /* W/m^2/nm */
  double sun_angular_radius;          /* rad */
  double bottom_radius;               /* m */
  double top_radius;                  /* m */
  int num_rayleigh_layers;
  const atmo_density_layer* rayleigh_density;
  const double* rayleigh_scattering;  /* 1/m */
  int num_mie_layers;
  const atmo_density_layer* mie_density;
  const double* mie_scattering;       /* 1/m */
  const double* mie_extinction;       /* 1/m */
  double mie_phase_function_g;
  int num_absorption_layers;
  const atmo_density_layer* absorption_density;
  const double* absorption_extinction; /* 1/m */
  const double* ground_albedo;
  double max_sun_zenith_angle;        /* rad */
  double length_unit_in_meters;
  unsigned num_precomputed_wavelengths;
  int combine_scattering_textures;
  int half_precision;
} atmo_model_params;

/* LUT resolution. The reference fixes these at compile time
 * (atmosphere/constants.h:47-61); here they are per model. NULL selects the
 * reference values: 256x64, (32, 128, 32, 8), 64x16. Every size is in
 * [2, 4096], scattering_mu is even and >= 4 (two halves of at least 2
 * texels), scattering_nu <= 32, irradiance_width
 * <= 512 (the density pass keeps nu tables and an irradiance row in LDS),
 * and R x MU x MU_S x NU <= 2^28. */
typedef struct atmo_lut_sizes {
  int transmittance_width;
  int transmittance_height;
  int scattering_r;
  int scattering_mu;
  int scattering_mu_s;
  int scattering_nu;
  int irradiance_width;
  int irradiance_height;
} atmo_lut_sizes;

typedef enum atmo_lut_id {
  ATMO_LUT_TRANSMITTANCE = 0,
  ATMO_LUT_SCATTERING = 1,
  ATMO_LUT_SINGLE_MIE_SCATTERING = 2, /* absent when textures are combined */
  ATMO_LUT_IRRADIANCE = 3,
  /* Temporaries of the last Init pass, kept only when the model was created
   * with ATMO_OPTION_RETAIN_INTERMEDIATES (checks and debugging). */
  ATMO_LUT_DELTA_IRRADIANCE = 4,
  ATMO_LUT_DELTA_RAYLEIGH = 5,        /* also delta multiple scattering */
  ATMO_LUT_DELTA_MIE = 6,
  ATMO_LUT_DELTA_SCATTERING_DENSITY = 7,
} atmo_lut_id;

/* F16 texels hold value * scale[c]. RGB21 (the delta scattering density of
 * FP16 models): 8 bytes per texel, three 21-bit floats (the FP32 sign,
 * exponent and top 12 mantissa bits, rounded to nearest), channel c in bits
 * 21c .. 21c + 20 of the little-endian 64-bit word. */
typedef enum atmo_dtype {
  ATMO_DTYPE_F32 = 0,
  ATMO_DTYPE_F16 = 1,
  ATMO_DTYPE_RGB21 = 2,
} atmo_dtype;

/* A LUT in device memory: width x height x depth texels of 'channels'
 * channels (4, RGBA; 3, RGB, for the delta scattering density), x fastest,
 * in the GL texture layout (the packed scattering LUT has
 * x = nu_index * MU_S + mu_s_index, y = mu_index, z = r_index). A stored
 * value equals the physical value times scale[c] (scale is 1 for F32). */
typedef struct atmo_lut_view {
  void* data;
  int width;
  int height;
  int depth;
  int channels;
  atmo_dtype dtype;
  float scale[4];
} atmo_lut_view;

typedef enum atmo_option {
  ATMO_OPTION_RETAIN_INTERMEDIATES = 0, /* keep the delta LUTs after Init */
  ATMO_OPTION_PASS_TIMING = 1,          /* time every kernel launch of Init */
  /* Separate scattering textures only (default on; no counterpart in the
   * reference): also keep a lookup copy of S and S_Mie in which, for every
   * (r, mu) row and (nu, mu_s) anchor, the S and S_Mie texels of the 2 x 2
   * (nu, mu_s) taps a lookup reads there are contiguous (128 B FP32), and
   * with FP16 storage those of both mu taps too (128 B), so that a 16-tap
   * scattering lookup reads 4 (FP32) or 2 (FP16) 128-byte blocks. Costs
   * about 4 (FP32) or 8 (FP16) times the size of S and S_Mie
   * (x (NU+1)(MU_S+1)/(NU MU_S), and (MU+1)/MU with FP16);
   * lookups give identical results with and without it; the LUT views and
   * files are unchanged. Turning it on for an initialized model builds the
   * copy at once (synchronous). */
  ATMO_OPTION_INTERLEAVED_LOOKUPS = 2,
} atmo_option;

/* Kernel passes of the precomputation (atmosphere/model.cc:1085-1211). */
typedef enum atmo_pass {
  ATMO_PASS_TRANSMITTANCE = 0,
  ATMO_PASS_DIRECT_IRRADIANCE = 1,
  ATMO_PASS_SINGLE_SCATTERING = 2,
  ATMO_PASS_SCATTERING_DENSITY = 3,
  ATMO_PASS_INDIRECT_IRRADIANCE = 4,
  ATMO_PASS_MULTIPLE_SCATTERING = 5,
} atmo_pass;

/* Duration of one kernel launch, measured with HIP events recorded on the
 * Init stream around the launch. */
typedef struct atmo_pass_timing {
  int pass;   /* atmo_pass */
  int order;  /* scattering order computed (SD, MS) or read (II); 0 else */
  float ms;
} atmo_pass_timing;

typedef struct atmo_model atmo_model;

/* Model::Model (atmosphere/model.h:182-281, model.cc:613-795). Validates and
 * copies the parameters, allocates the result LUTs on the current device. */
atmo_status atmo_model_create(const atmo_model_params* params,
                              const atmo_lut_sizes* sizes, atmo_model** out);
/* Model::~Model (atmosphere/model.cc:801-811). NULL is a no-op. A model
 * held by a live atmo_batch is not destroyed (ATMO_ERROR_INVALID_ARGUMENT):
 * destroy the batch first. */
atmo_status atmo_model_destroy(atmo_model* model);
atmo_status atmo_model_set_option(atmo_model* model, atmo_option option,
                                  int value);
/* Launch timings of the last Init of a model created with
 * ATMO_OPTION_PASS_TIMING (waits for that Init to finish). Writes at most
 * capacity entries and the total count to *count. */
atmo_status atmo_model_pass_timings(atmo_model* model, atmo_pass_timing* out,
                                    int capacity, int* count);
/* Model::Init (atmosphere/model.h:285, model.cc:866-975): precomputes the
 * LUTs with num_scattering_orders scattering orders (>= 1). */
atmo_status atmo_model_init(atmo_model* model, unsigned num_scattering_orders,
                            void* stream);
/* Allocation-free Init. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It computes the same LUTs
 * as atmo_model_init, bit for bit, but carves the Init temporaries from a
 * static device array of ATMO_STATIC_SCRATCH_BYTES per GPU (held from the
 * library's first use on that GPU until the process ends) and calls no
 * hipMalloc or hipFree. A model whose temporaries exceed that size
 * (atmo_model_footprint's init_scratch; e.g. 46.2 MB for 3 wavelengths at
 * the reference LUT sizes in FP32, 568 MB for 15 wavelengths) is rejected
 * with ATMO_ERROR_INVALID_ARGUMENT, as are models in a live batch or owned
 * by a server. Returns without waiting; the temporaries are not retained and
 * no pass timings are recorded. Every call on one GPU shares that GPU's
 * array: calls whose work overlaps on the device (on different streams, or
 * from different threads) race, so the caller orders them (one stream, or
 * events). The model must have been created on the current device. */
#define ATMO_STATIC_SCRATCH_BYTES (64u << 20)
atmo_status atmo_model_init_static_scratch(atmo_model* model,
                                           unsigned num_scattering_orders,
                                           void* stream);
/* Init of several models in one batched submission (no counterpart in the
 * reference, which precomputes one Model at a time): all models' passes of
 * one kind run in a single launch, the launch sequence is captured once in a
 * HIP graph and replayed. orders[i] belongs to models[i]. */
atmo_status atmo_model_init_batch(atmo_model* const* models,
                                  const unsigned* orders, int count,
                                  void* stream);
/* A reusable batched precompute plan (no counterpart in the reference): the
 * pass sequence of several models, one launch per pass level over all of
 * them, recorded once. With use_graph the launches are captured in a HIP
 * graph and atmo_batch_launch replays the graph. The batch's launches use
 * the models' temporaries and lookup copies, so while it lives: a model
 * cannot join a second batch (atmo_batch_create fails), atmo_model_init
 * keeps the temporaries, turning ATMO_OPTION_INTERLEAVED_LOOKUPS on or off
 * and atmo_model_destroy fail. The temporaries are freed when the batch is
 * destroyed (unless the model retains them). Models owned by a server
 * cannot be batched. atmo_model_init_batch is create + launch + wait +
 * destroy. */
typedef struct atmo_batch atmo_batch;
atmo_status atmo_batch_create(atmo_model* const* models,
                              const unsigned* orders, int count, int use_graph,
                              atmo_batch** out);
atmo_status atmo_batch_launch(atmo_batch* batch, void* stream);
int atmo_batch_num_launches(const atmo_batch* batch);
void atmo_batch_destroy(atmo_batch* batch);

/* Tile shapes and their autotuner (no counterpart in the reference, whose
 * passes are fragment-shader draws). The scattering-density pass runs with
 * sd_nu_per_thread nu outputs per thread in blocks of sd_block_threads; the
 * multiple-scattering pass with ms_texels_per_thread texels per thread in
 * blocks of ms_block_threads, its blocks walking the (r, mu) rows of the LUT r-major (ms_row_order 0, the
 * LUT's own order), mu-major (1) or in ray order (2: rows sorted by the
 * impact parameter r sqrt(1 - mu^2) of their view ray, so that blocks that
 * run together read the same source rows).
 * Every shape computes each texel identically: the LUTs do not depend on
 * the shapes. Shapes are chosen per LUT size class, (MU, NU) for the
 * density pass and (R, MU, MU_S, NU) for multiple scattering, in a table
 * shared by the process; a model's Init, batch or serving request reads it
 * when its job table is built. Classes never tuned or set use the defaults
 * (4 nu per thread in blocks of 128 threads when MU x ceil(NU / 4) <= 256,
 * else 256; the smallest multiple-scattering tile that holds a row (1 x 256,
 * 2 x 256, then 2 texels x 512 threads); ray order). */
typedef struct atmo_tiles {
  int sd_nu_per_thread;
  int sd_block_threads;
  int ms_texels_per_thread;
  int ms_block_threads;
  int ms_row_order;
  int tuned; /* atmo_get_tiles: bit 0 density class, bit 1 multiple class set */
} atmo_tiles;
/* Candidate shapes (returns the count; fills up to capacity entries). */
int atmo_sd_tile_candidates(int* nu_per_thread, int* block_threads,
                            int capacity);
int atmo_ms_tile_candidates(int* texels_per_thread, int* block_threads,
                            int capacity);
/* Times every candidate shape of the two passes on a batch of Earth-like
 * atmospheres of each listed size (enough copies to fill the GPU) on
 * 'stream', and records the fastest shape per size class; a shape replaces
 * the default only if it is at least 10% faster. Classes already in the
 * table are skipped unless retune is nonzero. Synchronous. */
atmo_status atmo_autotune(const atmo_lut_sizes* sizes, int count, int retune,
                          void* stream);
atmo_status atmo_get_tiles(const atmo_lut_sizes* sizes, atmo_tiles* out);
/* Pins the shapes of the size classes of 'sizes' (the tuned field is
 * ignored); an error if a shape is not a candidate, ms_row_order is not 0,
 * 1 or 2, or if NU x MU_S texels per row exceed ms_texels_per_thread x
 * ms_block_threads (rows of at most 1024 texels; longer rows take a gather
 * kernel whatever the shape). */
atmo_status atmo_set_tiles(const atmo_lut_sizes* sizes, const atmo_tiles* tiles);
/* The table as text ("sd MU NU nu_per_thread block_threads",
 * "ms R MU MU_S NU texels_per_thread block_threads row_order"), merged back by
 * atmo_tune_load. */
atmo_status atmo_tune_save(const char* path);
atmo_status atmo_tune_load(const char* path);
void atmo_tune_reset(void);

/* The device LUTs (the reference binds them to texture units instead,
 * Model::SetProgramUniforms, atmosphere/model.h:289-294). */
atmo_status atmo_model_get_lut(const atmo_model* model, atmo_lut_id id,
                               atmo_lut_view* view);
/* Synchronous copy of a LUT to host memory as float RGBA, decoded to
 * physical values (width*height*depth*4 floats). */
atmo_status atmo_model_copy_lut(const atmo_model* model, atmo_lut_id id,
                                float* host_rgba);
/* Device bytes held by the model's result LUTs, and the peak extra bytes
 * Init allocates for its temporaries. */
atmo_status atmo_model_footprint(const atmo_model* model, size_t* resident,
                                 size_t* init_scratch);

/* LUT files: raw float32 RGBA texels, x fastest, no header, decoded to
 * physical values - the format the reference's WebGL exporter writes with
 * glGetTexImage (atmosphere/demo/webgl/precompute.cc:63-106) - named
 * transmittance.dat, scattering.dat, single_mie_scattering.dat (separate
 * Mie only) and irradiance.dat in 'directory', as the reference CPU model's
 * cache (atmosphere/reference/model.cc:74-84, :240-244). Export waits for
 * the device. */
atmo_status atmo_model_export_luts(const atmo_model* model,
                                   const char* directory);
/* Loads the files into the model's LUTs (their sizes must match the
 * model's; FP16 models re-encode with the same per-channel scale rule as
 * Init). The model is then initialised, as after Init. */
atmo_status atmo_model_import_luts(atmo_model* model, const char* directory);
/* The per-channel scale of FP16 LUT storage (half_precision; no
 * counterpart in the reference, whose half-float textures are unscaled):
 * for a channel whose largest magnitude is max_abs, the power of two
 * 2^min(14 - e, 126), where max_abs = f 2^e with f in [0.5, 1). It maps
 * max_abs into [2^13, 2^14), well inside the FP16 range, and stays finite
 * (with a normal inverse) for tiny and denormal maxima. Zero, negative, NaN
 * and infinite max_abs give 1. Init and atmo_model_import_luts use it;
 * atmo_lut_view.scale reports it. Host-only. */
float atmo_fp16_channel_scale(float max_abs);
/* reference::Model::Init with a cache directory
 * (atmosphere/reference/model.cc:74-84): loads the LUTs if
 * directory/transmittance.dat exists, else runs Init and saves them. */
atmo_status atmo_model_init_cached(atmo_model* model,
                                   unsigned num_scattering_orders,
                                   const char* directory, void* stream);

/* GetSolarRadiance (atmosphere/model.cc:227-230): host-side, 3 floats. */
atmo_status atmo_get_solar_radiance(const atmo_model* model, float* rgb);
/* GetSkyRadiance (atmosphere/functions.glsl:1705-1769, wrapped at
 * model.cc:231-237). camera, view_ray, sun_direction: n x 3; shadow_length:
 * n floats. Outputs radiance and transmittance: n x 3. */
atmo_status atmo_get_sky_radiance(const atmo_model* model, int64_t n,
                                  const float* camera, const float* view_ray,
                                  const float* shadow_length,
                                  const float* sun_direction, float* radiance,
                                  float* transmittance, void* stream);
/* GetSkyRadianceToPoint (functions.glsl:1787-1863, model.cc:238-244). */
atmo_status atmo_get_sky_radiance_to_point(
    const atmo_model* model, int64_t n, const float* camera,
    const float* point, const float* shadow_length, const float* sun_direction,
    float* radiance, float* transmittance, void* stream);
/* GetSunAndSkyIrradiance (functions.glsl:1878-1896, model.cc:245-250).
 * Returns the sun irradiance in sun_irradiance and the sky irradiance in
 * sky_irradiance, n x 3 each. */
atmo_status atmo_get_sun_and_sky_irradiance(
    const atmo_model* model, int64_t n, const float* point,
    const float* normal, const float* sun_direction, float* sun_irradiance,
    float* sky_irradiance, void* stream);

/* AtmosphereParameters (atmosphere/definitions.glsl:205-255) at three
 * wavelengths, as the constructor lowers them (atmosphere/model.cc:636-666,
 * 718-734) but in exact binary values: lengths divided by the length unit,
 * coefficients multiplied by it, profiles padded to two layers. Host-only. */
typedef struct atmo_atmosphere {
  double solar_irradiance[3];
  double sun_angular_radius;
  double bottom_radius;
  double top_radius;
  atmo_density_layer rayleigh_density[2];
  double rayleigh_scattering[3];
  atmo_density_layer mie_density[2];
  double mie_scattering[3];
  double mie_extinction[3];
  double mie_phase_function_g;
  atmo_density_layer absorption_density[2];
  double absorption_extinction[3];
  double ground_albedo[3];
  double mu_s_min;
} atmo_atmosphere;
atmo_status atmo_lower_parameters(const atmo_model_params* params,
                                  const double* lambdas3,
                                  atmo_atmosphere* out);
/* Number of Precompute iterations (1, or ceil(num_precomputed_wavelengths/3)
 * in illuminance mode) and, for iteration i, its 3 wavelengths and
 * luminance_from_radiance matrix (row-major), atmosphere/model.cc:907-943.
 * Host-only. */
int atmo_num_iterations(const atmo_model_params* params);
atmo_status atmo_iteration(const atmo_model_params* params, int i,
                           double* lambdas3, float* matrix9);

/* The luminance API of atmosphere/model.cc:252-281: GetSolarLuminance,
 * GetSkyLuminance, GetSkyLuminanceToPoint and GetSunAndSkyIlluminance, i.e.
 * the radiance lookups times SKY_SPECTRAL_RADIANCE_TO_LUMINANCE (sky terms)
 * or SUN_SPECTRAL_RADIANCE_TO_LUMINANCE (sun terms). They are available in
 * every mode; the radiance functions above only when at most 3 wavelengths
 * are precomputed (as in the reference, model.cc:226 and 769-772), and
 * return ATMO_ERROR_UNSUPPORTED otherwise. */
atmo_status atmo_get_solar_luminance(const atmo_model* model, float* rgb);
atmo_status atmo_get_sky_luminance(const atmo_model* model, int64_t n,
                                   const float* camera, const float* view_ray,
                                   const float* shadow_length,
                                   const float* sun_direction,
                                   float* luminance, float* transmittance,
                                   void* stream);
atmo_status atmo_get_sky_luminance_to_point(
    const atmo_model* model, int64_t n, const float* camera,
    const float* point, const float* shadow_length, const float* sun_direction,
    float* luminance, float* transmittance, void* stream);
atmo_status atmo_get_sun_and_sky_illuminance(
    const atmo_model* model, int64_t n, const float* point,
    const float* normal, const float* sun_direction, float* sun_illuminance,
    float* sky_illuminance, void* stream);
/* The two factor vectors (SKY_..., SUN_SPECTRAL_RADIANCE_TO_LUMINANCE). */
atmo_status atmo_get_luminance_factors(const atmo_model* model, double* sky_k,
                                       double* sun_k);

/* Model::ConvertSpectrumToLinearSrgb (atmosphere/model.h:301-304,
 * model.cc:1020-1040). Host-only. */
atmo_status atmo_convert_spectrum_to_linear_srgb(int n,
                                                 const double* wavelengths,
                                                 const double* spectrum,
                                                 double* r, double* g,
                                                 double* b);

/* ------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference). A worker thread takes
 * precompute requests and lookup batches as they arrive. Precompute requests
 * are admitted in arrival order by device-memory footprint (resident LUTs of
 * all models held + Init temporaries of running batches <= budget) into up
 * to max_inflight_batches concurrent batches, each on its own stream:
 * a request that arrives while other batches run joins the next batch as soon
 * as a slot and memory are free, and its work overlaps the running ones
 * (continuous batching). Finished LUT sets stay resident until released: a
 * request that does not fit waits for batches to finish and models to be
 * released; one that cannot fit in an empty server completes with
 * ATMO_ERROR_OUT_OF_MEMORY.
 * Lookups run on resident models (or wait on the device for a model whose
 * precompute is in flight). Every request completes with its submit, start
 * and end times, for latency statistics.
 * ---------------------------------------------------------------------- */
typedef struct atmo_server atmo_server;
typedef struct atmo_server_config {
  size_t memory_budget;      /* bytes of device memory; 0 = 80% of free */
  int max_inflight_batches;  /* concurrent precompute batches; 0 = 3 */
  int max_batch_models;      /* models per batch; 0 = 16 */
  int lookup_streams;        /* streams for lookups; 0 = 2 */
} atmo_server_config;

typedef enum atmo_request_kind {
  ATMO_REQUEST_PRECOMPUTE = 0,
  ATMO_REQUEST_LOOKUP = 1,
} atmo_request_kind;

typedef enum atmo_lookup {
  ATMO_LOOKUP_SKY = 0,          /* a=camera b=view_ray c=shadow_length d=sun */
  ATMO_LOOKUP_SKY_TO_POINT = 1, /* a=camera b=point c=shadow_length d=sun */
  ATMO_LOOKUP_SUN_AND_SKY = 2,  /* a=point b=normal c=sun d unused */
} atmo_lookup;

typedef struct atmo_completion {
  uint64_t request_id;
  int kind;            /* atmo_request_kind */
  int status;          /* atmo_status */
  uint64_t model_id;   /* precompute: id of the new model (= request_id) */
  int batch_models;    /* precompute: models in the batch that ran it */
  double submit_ms;    /* host times since the server started */
  double start_ms;     /* admitted and launched */
  double end_ms;       /* device work observed complete */
} atmo_completion;

typedef struct atmo_server_stats {
  size_t budget_bytes;
  size_t resident_bytes;
  size_t scratch_bytes;
  int resident_models;
  int pending_precomputes;
  int pending_lookups;
  int inflight_batches;
  int inflight_lookups;
  uint64_t completed_precomputes;
  uint64_t completed_lookups;
  uint64_t failed;
} atmo_server_stats;

atmo_status atmo_server_create(const atmo_server_config* config,
                               atmo_server** out);
/* Waits for in-flight work, then frees every model and buffer. */
void atmo_server_destroy(atmo_server* server);
/* Copies the parameters; the model id of the result is *request_id. */
atmo_status atmo_server_submit_precompute(atmo_server* server,
                                          const atmo_model_params* params,
                                          const atmo_lut_sizes* sizes,
                                          unsigned num_scattering_orders,
                                          uint64_t* request_id);
/* Device arrays as for the lookup functions; they must stay valid until the
 * request completes. luminance selects the luminance API. */
atmo_status atmo_server_submit_lookup(atmo_server* server, uint64_t model_id,
                                      atmo_lookup lookup, int luminance,
                                      int64_t n, const float* a,
                                      const float* b, const float* c,
                                      const float* d, float* out0,
                                      float* out1, uint64_t* request_id);
/* Moves up to capacity completions to out; waits up to timeout_ms for at
 * least one when none is ready (0 = do not wait). */
atmo_status atmo_server_poll(atmo_server* server, atmo_completion* out,
                             int capacity, int* count, int timeout_ms);
/* The LUTs of the model are freed once its launched lookups are done;
 * lookups on it still queued complete with ATMO_ERROR_NOT_INITIALIZED. */
atmo_status atmo_server_release(atmo_server* server, uint64_t model_id);
/* The resident model (for LUT views); valid until released. */
atmo_status atmo_server_model(atmo_server* server, uint64_t model_id,
                              atmo_model** model);
atmo_status atmo_server_get_stats(atmo_server* server,
                                  atmo_server_stats* stats);

const char* atmo_last_error(void);

/* Seeded synthetic query batch on the device (no counterpart in the
 * reference; the input generator of the lookup benchmarks and serving
 * workload). Ray i uses outputs 6i+1 .. 6i+6 of the SplitMix64 sequence of
 * 'seed': camera at altitude U[0, max_altitude] above (0, 0, bottom_radius),
 * view ray uniform on the sphere, sun zenith U[0, max_sun_zenith_angle] and
 * azimuth U[0, 2 pi), point = camera + U[0, max_distance] * view ray moved
 * radially to within 1e-3 of the shell [bottom_radius, top_radius],
 * normal = point / |point|, shadow length 0. Lengths in the model's unit;
 * arrays n x 3 (shadow_length n), device memory. */
atmo_status atmo_generate_sky_queries(uint64_t seed, int64_t n,
                                      double bottom_radius, double top_radius,
                                      double max_altitude, double max_distance,
                                      double max_sun_zenith_angle,
                                      float* camera, float* view_ray,
                                      float* sun_direction, float* point,
                                      float* normal, float* shadow_length,
                                      void* stream);

/* Minimal device-memory and stream helpers, so that callers without their
 * own HIP runtime (ctypes, cgo) can drive the lookups. */
atmo_status atmo_malloc(void** ptr, size_t bytes);
atmo_status atmo_free(void* ptr);
/* dst is device memory; src is host or device memory (asynchronous). */
atmo_status atmo_memcpy_to_device(void* dst, const void* src, size_t bytes,
                                  void* stream);
atmo_status atmo_memcpy_to_host(void* dst, const void* src, size_t bytes,
                                void* stream);
atmo_status atmo_stream_create(void** stream);
atmo_status atmo_stream_destroy(void* stream);
atmo_status atmo_stream_synchronize(void* stream);
atmo_status atmo_device_synchronize(void);
/* HIP events, for timing work on a stream (elapsed time in ms). */
atmo_status atmo_event_create(void** event);
atmo_status atmo_event_destroy(void* event);
atmo_status atmo_event_record(void* event, void* stream);
atmo_status atmo_event_elapsed_ms(void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* ATMO_H_ */
