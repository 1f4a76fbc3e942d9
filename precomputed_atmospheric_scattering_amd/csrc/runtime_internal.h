// Internal interface of libatmo's runtime, shared by runtime.cpp (the C ABI
// of models and batches) and server.cpp (the serving loop). Not installed.
#ifndef ATMO_RUNTIME_INTERNAL_H_
#define ATMO_RUNTIME_INTERNAL_H_

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/atmo.h"
#include "atmo_device.h"
#include "kernels.h"

namespace atmo_rt {

// The constructor parameters, copied.
struct Spec {
  std::vector<double> wavelengths, solar, rayleigh, mie_sca, mie_ext, absorption,
      albedo;
  std::vector<atmo_density_layer> prof_r, prof_m, prof_a;
  double sun_angular_radius, bottom_radius, top_radius, g, max_sun_zenith_angle,
      length_unit;
  unsigned num_precomputed_wavelengths;
  bool combined, half_precision;
};


struct Sizes {
  int tw, th, nr, nmu, nmus, nnu, ew, eh;
  size_t t_texels() const { return static_cast<size_t>(tw) * th; }
  size_t s_texels() const {
    return static_cast<size_t>(nr) * nmu * nmus * nnu;
  }
  size_t e_texels() const { return static_cast<size_t>(ew) * eh; }
};


enum Pass { kT = 0, kE0, kSS, kSD, kII, kMS, kAcc, kHalf, kClear, kPack, kOD, kPasses };


// One launch of a batched pass sequence: pass kind, order argument, and the
// segment of the plan's job and block-prefix arrays it uses.
struct Level {
  Pass pass;
  int order;  // SD: order computed; II: order read; others unused
  int job_offset, njobs, fb_offset, total_blocks, max_nnu, max_ew, max_w, min_w, max_lds_w,
      kinds, dj_kinds;
  int tile;   // SD and MS: the tile shape shared by the level's jobs
};

// The launch sequence of several models' Init (see plan_levels).
struct LevelPlan {
  std::vector<atmo::Job> jobs;
  std::vector<int> first_block;
  std::vector<Level> levels;
};

}  // namespace atmo_rt

// A model: parameters, LUTs, temporaries, job tables.

struct atmo_model {
  using Spec = atmo_rt::Spec;
  using Sizes = atmo_rt::Sizes;
  static constexpr int kPasses = atmo_rt::kPasses;
  Spec spec;
  Sizes sizes;
  int device = 0;
  bool retain = false;
  bool initialized = false;
  bool external_temporaries = false;  // temporaries carved from caller memory
  // Live atmo_batch objects holding this model (0 or 1): their job tables
  // point at the model's temporaries and lookup copy, which must then stay.
  int live_batches = 0;
  float4 *T = nullptr, *S = nullptr, *SM = nullptr, *E = nullptr;
  // Init temporaries, carved from one block (owned, or caller memory with
  // external_temporaries). In illuminance mode each Precompute iteration has
  // its own copy, so that the iterations run as parallel jobs: iteration k's
  // delta LUTs are at dR_it + k * (scattering texels) and so on, its
  // transmittance and irradiance accumulators at T_it and E_it. dE, dR, dM,
  // dJ point at the last iteration's copies (the retained intermediates).
  char* temp_block = nullptr;
  float4 *T_it = nullptr, *E_it = nullptr;
  float4 *dE_it = nullptr, *dR_it = nullptr, *dM_it = nullptr;
  float4 *dE = nullptr, *dR = nullptr, *dM = nullptr;
  // Delta J: RGB float texels, or 3 x 21-bit floats in FP16 mode.
  char *dJ_it = nullptr, *dJ = nullptr;
  // FP16 storage: resident half LUTs; S and SM above are then FP32
  // accumulators that live only during Init, like the deltas.
  uint2 *S16 = nullptr, *SM16 = nullptr;
  // ATMO_OPTION_INTERLEAVED_LOOKUPS: the lookup copy of S and SM, in blocks
  // of the texels a (nu, mu_s) tap set reads (interleaved_bytes).
  bool interleave = true;
  void* Q = nullptr;
  unsigned last_orders = 4;
  // FP64 accumulators of S and SM in illuminance mode (Init temporaries),
  // one per iteration: iteration k's at S64 + 4 * k * (scattering texels).
  double *S64 = nullptr, *SM64 = nullptr;
  // Optical depths of the transmittance texels, shared by the iterations'
  // transmittance LUTs and the final one (Init temporary, illuminance mode).
  double* OD = nullptr;
  float* d_scales = nullptr;
  unsigned* d_maxbits = nullptr;
  // The (r, mu) rows in ray order (Job::ms_rows), R x MU entries.
  int* d_ms_rows = nullptr;
  // Device job table: [iterations..., rgb transmittance job, query job].
  std::vector<atmo::Job> jobs;
  atmo::Job* d_jobs = nullptr;
  // Per-pass block prefixes of the Precompute iterations' jobs (all of one
  // size): pass p's are entries p * (iters + 1) .. p * (iters + 1) + iters,
  // {0, b, 2b, ...}; a launch of the first n jobs uses the first n + 1.
  int iters = 1;
  int* d_first_block = nullptr;
  std::vector<int> first_block_host;
  float solar_radiance[3] = {};
  double sky_k[3] = {}, sun_k[3] = {};
  // Launch timing (ATMO_OPTION_PASS_TIMING).
  bool timing = false;
  struct Timed {
    int pass, order;
    hipEvent_t start, stop;
  };
  std::vector<Timed> timed;   // events of the last Init, in launch order
  std::vector<hipEvent_t> event_pool;
};


#define ATMO_TRY(x)                 \
  do {                              \
    atmo_status st_ = (x);          \
    if (st_ != ATMO_OK) return st_; \
  } while (0)

namespace atmo_rt {

atmo_status fail(atmo_status s, const std::string& msg);
atmo_status check_launch(hipError_t e, const char* what);
atmo_status validate_params(const atmo_model_params* p);
atmo_status validate_sizes(const atmo_lut_sizes& z);
int num_iterations(const Spec& s);
// Whether Init accumulates S and SM in FP64 (illuminance mode).
inline bool accumulates_fp64(const Spec& s) { return s.num_precomputed_wavelengths > 3; }
// Bytes of the lookup copy of a model with these sizes and storage: one
// 128-byte block per (r, mu) row (FP32) or (r, mu anchor) (FP16) and
// (nu, mu_s) anchor (atmo::lookup_copy_blocks); 0 with combined textures,
// which have no S_Mie.
inline size_t interleaved_bytes(const Sizes& z, bool half, bool combined) {
  return combined ? 0
                  : static_cast<size_t>(
                        atmo::lookup_copy_blocks(z.nr, z.nmu, z.nmus, z.nnu, half)) *
                        8 * sizeof(uint4);
}
int pass_blocks(Pass p, const atmo::Job& j);
// Tile shapes for a model's LUT sizes: the autotuner's choice, or the
// default shapes for sizes it has not tuned (autotune.cpp).
struct Tiles {
  int sd, ms;
  int ms_row_order;  // row order of the multiple-scattering blocks
};
Tiles tiles_for(const Sizes& z);
// Rebuild a model's job table for 'orders' and copy it to the device on
// 'stream' (asynchronous with respect to other streams).
atmo_status upload_jobs(atmo_model* m, unsigned orders, hipStream_t stream);
// Fills the interleaved lookup copy from S and SM (if the model has one).
atmo_status build_lookup_copy(atmo_model* m, hipStream_t stream);
// Bytes of Init temporaries of a model (or of a model of these sizes and
// storage with 'iters' Precompute iterations), and carving them out of
// caller-owned memory (the model then does not free them).
size_t temporaries_bytes(const atmo_model* m);
size_t temporaries_bytes_for(const Sizes& z, bool half, bool combined, int iters);
void use_external_temporaries(atmo_model* m, char* base);
void drop_external_temporaries(atmo_model* m);
// Plan the levels of the models' Init (their job tables must be current).
void plan_levels(atmo_model* const* models, const unsigned* orders, int count,
                 LevelPlan* plan);
// Streams and events on which the per-tile launches of one pass level run
// side by side: a level split into k launches (one per tile shape) forks
// k - 1 of them onto side[0..k-2] and joins them back. Default priority,
// non-blocking streams; ev holds kMaxFork events.
constexpr int kMaxFork = atmo::kSdTiles;
struct Fork {
  hipStream_t side[kMaxFork - 1] = {};
  hipEvent_t ev[kMaxFork] = {};
};
atmo_status create_fork(Fork* f);
void destroy_fork(Fork* f);
// Launch a plan whose jobs and block prefixes are at d_jobs / d_first_block.
// With a fork, the launches of one level that differ only in tile shape run
// concurrently; without, one after the other.
atmo_status enqueue_levels(const LevelPlan& plan, const atmo::Job* d_jobs,
                           const int* d_first_block, hipStream_t stream,
                           const Fork* fork = nullptr);

}  // namespace atmo_rt

#endif  // ATMO_RUNTIME_INTERNAL_H_
