// Host-side launchers of the precompute and lookup kernels.
#ifndef ATMO_KERNELS_H_
#define ATMO_KERNELS_H_

#include <hip/hip_runtime.h>

#include "atmo_device.h"

namespace atmo {

// One launch of a pass over a set of jobs. Block b of the launch belongs to
// job i with first_block[i] <= b < first_block[i + 1] (device arrays).
struct PassLaunch {
  const Job* jobs;
  const int* first_block;
  int njobs;
  int total_blocks;
  int max_nnu;  // largest nu size among the jobs (LDS sizing)
  int max_ew;   // largest irradiance width among the jobs (LDS sizing)
  int max_w;    // largest packed row (nu x mu_s texels) among the jobs
  int min_w;    // smallest packed row among the jobs
  int max_lds_w;  // longest packed row of at most kMsMaxRow texels (0: none)
  int kinds;    // kJobsF32 | kJobsAcc64: which kinds of jobs the launch holds
  int tile;     // tile shape of the jobs (scattering density, multiple
                // scattering): every job of a launch has the same one
  int dj_kinds;  // kDjF32 | kDjRgb21: the delta J formats of the jobs
};

constexpr int kDjF32 = 1;    // delta J as RGB float (FP32-storage jobs)
constexpr int kDjRgb21 = 2;  // delta J as 3 x 21-bit floats (FP16-storage jobs)

constexpr int kJobsF32 = 1;    // jobs accumulating S and SM in FP32
constexpr int kJobsAcc64 = 2;  // illuminance-mode jobs with FP64 accumulators

constexpr int kBlock = 256;

// Tile shapes of the two passes that dominate Init, picked per LUT size
// class by the autotuner (autotune.cpp). Every shape computes each texel
// with the same operations in the same order, so the LUTs do not depend on
// the choice.
//
// Scattering density: nu outputs per thread (they share the phase function
// of the outgoing direction) and threads per block. A (r, mu_s) plane of
// MU x ceil(NU / chunk) threads is split into ceil(plane / block) blocks,
// each of which builds the plane's nu tables in LDS.
struct SdTile {
  int chunk, block;
};
constexpr int kSdTiles = 6;
constexpr SdTile kSdTile[kSdTiles] = {{4, 256}, {2, 256}, {1, 256},
                                      {4, 128}, {2, 128}, {4, 64}};
// Default shape: 4 nu per thread, in blocks of 128 threads when a plane has
// at most 256 threads (MU x ceil(NU / 4)), else 256. Measured on MI355X
// with the other shapes: at MU x NU = 128 x 8 (a 256-thread plane) the
// 128-thread blocks are 4% faster at every order; at 256 x 16 (1024) the
// 256-thread blocks are, the blocks then rebuilding a plane's tables from
// a LUT beyond the caches 4 times instead of 8.
ATMO_HD int sd_default_tile(int nmu, int nnu) {
  return nmu * ((nnu + 3) / 4) <= 256 ? 3 : 0;
}

// Multiple scattering: packed rows of at most kMsMaxRow texels are processed
// through LDS by blocks of kMsTile[tile].block threads, each thread holding
// kMsTile[tile].tpt texels: ms_rows_per_block(W, capacity) whole rows per
// block, capacity = tpt x block texels (0: the row is too long for LDS and
// takes the gather kernel; -1: it does not fit this tile). Rows of up to
// 1024 texels take 2 texels per thread in 512-thread blocks rather than 4 in
// 256: the blocks hold the same rows and the kernel runs at more waves per
// SIMD (the high-resolution pass, rows of 1024 texels: 83.0 against 88.8 ms
// per Init with FP32 storage, 81.5 against 83.0 with FP16; 1 texel in
// 1024-thread blocks took 103-105 ms).
struct MsTile {
  int tpt, block;
};
constexpr int kMsTiles = 3;
constexpr MsTile kMsTile[kMsTiles] = {{1, 256}, {2, 256}, {2, 512}};
constexpr int kMsMaxRow = 4 * kBlock;
constexpr int ms_tile_capacity(int t) { return kMsTile[t].tpt * kMsTile[t].block; }
// Rows per block are capped at 4 (blocks of 256 texels) or 8, which keeps
// the block's LDS (2 x capacity RGB floats + 3.2 KB of samples per row)
// under 64 KB.
ATMO_HD int ms_rows_per_block(int w, int capacity) {
  if (w > kMsMaxRow) return 0;
  if (w > capacity) return -1;
  const int r = capacity / w;
  const int cap = capacity <= kBlock ? 4 : 8;
  return r < cap ? r : cap;
}
// The smallest tile that holds a row of w texels.
ATMO_HD int ms_default_tile(int w) {
  if (w <= kBlock) return 0;
  if (w <= 2 * kBlock) return 1;
  return 2;
}
// Row orders of the multiple-scattering blocks (Job::ms_row_order):
// 0 r-major (the LUT's own order), 1 mu-major, 2 ray order (rows sorted by
// the impact parameter of their view ray, Job::ms_rows). The samples of a
// row's ray lie on the ray's line, whose points have the same impact
// parameter r sqrt(1 - mu^2), so rows of nearly equal impact parameter read
// nearly the same source rows of the scattering density: in ray order the
// blocks resident on one XCD share their source rows in its L2.
constexpr int kMsRowOrders = 3;
// Default: ray order at every size. At the reference sizes (a 12.6 MB
// delta J per iteration) it fetches 0.25 GB per launch of 5 iterations
// against 1.05 GB in r-major order, and runs 1.4% faster.
ATMO_HD int ms_default_row_order(int nr, int nmu, int w) {
  (void)nr;
  (void)nmu;
  (void)w;
  return 2;
}

// Blocks one job needs for each pass.
int blocks_optical_depth(const Job& j);
int blocks_transmittance(const Job& j);
int blocks_direct_irradiance(const Job& j);
int blocks_single_scattering(const Job& j);
int blocks_scattering_density(const Job& j);
int blocks_indirect_irradiance(const Job& j);
int blocks_multiple_scattering(const Job& j);
int blocks_half_conversion(const Job& j);
int blocks_lookup_copy(const Job& j);
int blocks_combine(const Job& j);

// The optical depths of the jobs' transmittance texels into their OD (one
// job per model), read by launch_transmittance for jobs whose OD is set.
hipError_t launch_optical_depth(const PassLaunch& p, hipStream_t s);
hipError_t launch_transmittance(const PassLaunch& p, hipStream_t s);
hipError_t launch_direct_irradiance(const PassLaunch& p, hipStream_t s);
hipError_t launch_single_scattering(const PassLaunch& p, hipStream_t s);
// 'order' is the scattering order being computed (>= 2).
hipError_t launch_scattering_density(const PassLaunch& p, int order,
                                     hipStream_t s);
// 'order' is the order of the scattering read by the pass (>= 1).
hipError_t launch_indirect_irradiance(const PassLaunch& p, int order,
                                      hipStream_t s);
hipError_t launch_multiple_scattering(const PassLaunch& p, hipStream_t s);
// FP16 storage of the scattering LUTs of half-precision jobs:
// launch_clear_max (one block per job) resets the per-channel maxima, then
// launch_to_half (blocks_half_conversion blocks per job) computes the maxima
// and writes the scaled FP16 texels.
hipError_t launch_clear_max(const PassLaunch& p, hipStream_t s);
// Illuminance mode: sums the Precompute iterations' FP64 accumulators into S
// and SM, and their irradiance accumulators into E (blocks_combine blocks
// per job; the job's niter, S64, SM64 and E_it name the accumulators).
hipError_t launch_acc64_to_f32(const PassLaunch& p, hipStream_t s);
hipError_t launch_to_half(const PassLaunch& p, hipStream_t s);
// The lookup copy of S and SM, in quads (blocks_lookup_copy blocks per
// job).
hipError_t launch_interleave(const PassLaunch& p, hipStream_t s);

// Lookups (one ray / point per thread). 'job' is a device pointer. With
// 'luminance' the results are scaled by the job's radiance-to-luminance
// factors (the Luminance API of atmosphere/model.cc:252-281).
hipError_t launch_sky_radiance(const Job* job, int64_t n, const float* camera,
                               const float* view_ray,
                               const float* shadow_length,
                               const float* sun_direction, float* radiance,
                               float* transmittance, int luminance, hipStream_t s);
hipError_t launch_sky_radiance_to_point(const Job* job, int64_t n,
                                        const float* camera,
                                        const float* point,
                                        const float* shadow_length,
                                        const float* sun_direction,
                                        float* radiance, float* transmittance,
                                        int luminance, hipStream_t s);
// Seeded synthetic query batch (camera, view ray, sun direction, point,
// normal: n x 3; shadow length: n).
hipError_t launch_generate_queries(uint64_t seed, int64_t n, double bottom_radius,
                                  double top_radius, double max_altitude,
                                  double max_distance, double max_sun_zenith,
                                  float* camera, float* view_ray, float* sun_direction,
                                  float* point, float* normal, float* shadow_length,
                                  hipStream_t s);
hipError_t launch_sun_and_sky_irradiance(const Job* job, int64_t n,
                                         const float* point,
                                         const float* normal,
                                         const float* sun_direction,
                                         float* sun_irradiance,
                                         float* sky_irradiance, int luminance,
                                         hipStream_t s);

}  // namespace atmo

#endif  // ATMO_KERNELS_H_
