// Tile-shape autotuner of libatmo (include/atmo.h, atmo_autotune and
// friends). The scattering-density and multiple-scattering passes each come
// in several tile shapes (kernels.h: kSdTile, kMsTile) that compute every
// texel identically. The best shape depends on the LUT sizes: a (r, mu_s)
// plane of the density pass holds MU x NU outputs, which a 256-thread block
// of 4-nu threads only half fills at MU = 64, NU = 8, and a packed row of
// the multiple-scattering pass holds NU x MU_S texels. The tuner times every
// shape on a batch of Earth-like atmospheres of the size in question, enough
// copies to fill the GPU, and records the fastest per size class:
//   scattering density:   (MU, NU)
//   multiple scattering:  (R, MU, MU_S, NU), texels per thread, block
//                         threads and the order of the blocks over the
//                         (r, mu) rows
// The choice is read when a model's job table is built (Init, batches, the
// serving loop); sizes never tuned get the default shapes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/atmo.h"
#include "atmo_device.h"
#include "kernels.h"
#include "runtime_internal.h"

namespace atmo_rt {
namespace {

using SdKey = std::array<int, 2>;  // (MU, NU)
using MsKey = std::array<int, 4>;  // (R, MU, MU_S, NU)

// A multiple-scattering shape: tile (index into kMsTile) and row order.
struct MsShape {
  int tile, row_order;
};

std::mutex g_tiles_mu;
std::map<SdKey, int> g_sd;
std::map<MsKey, MsShape> g_ms;

SdKey sd_key(const Sizes& z) { return {z.nmu, z.nnu}; }
MsKey ms_key(const Sizes& z) { return {z.nr, z.nmu, z.nmus, z.nnu}; }

MsShape ms_default(const Sizes& z) {
  const int w = z.nnu * z.nmus;
  return {atmo::ms_default_tile(w), atmo::ms_default_row_order(z.nr, z.nmu, w)};
}

Sizes sizes_of(const atmo_lut_sizes& z) {
  return Sizes{z.transmittance_width, z.transmittance_height, z.scattering_r,
               z.scattering_mu,       z.scattering_mu_s,      z.scattering_nu,
               z.irradiance_width,    z.irradiance_height};
}

// Whether multiple-scattering tile t can hold the packed rows of z (rows
// too long for LDS take the gather kernel whatever the tile).
bool ms_tile_fits(const Sizes& z, int t) {
  return atmo::ms_rows_per_block(z.nnu * z.nmus, atmo::ms_tile_capacity(t)) >= 0;
}

int sd_tile_of_shape(int chunk, int block) {
  for (int t = 0; t < atmo::kSdTiles; ++t)
    if (atmo::kSdTile[t].chunk == chunk && atmo::kSdTile[t].block == block) return t;
  return -1;
}

int ms_tile_of_shape(int tpt, int block) {
  for (int t = 0; t < atmo::kMsTiles; ++t)
    if (atmo::kMsTile[t].tpt == tpt && atmo::kMsTile[t].block == block) return t;
  return -1;
}

#define TUNE_HIP(call)                                                      \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(e_ == hipErrorOutOfMemory ? ATMO_ERROR_OUT_OF_MEMORY      \
                                            : ATMO_ERROR_DEVICE,            \
                  std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

// The tuning atmosphere: Earth at 680, 550 and 440 nm (Rayleigh 8 km,
// Mie 1.2 km scale heights, no ozone), in km.
struct EarthParams {
  double wl[3] = {440.0, 550.0, 680.0};
  double solar[3] = {1.5, 1.5, 1.5};
  double rayleigh[3] = {33.1e-6, 13.558e-6, 5.802e-6};
  double mie_sca[3] = {3.996e-6, 3.996e-6, 3.996e-6};
  double mie_ext[3] = {4.44e-6, 4.44e-6, 4.44e-6};
  double absorption[3] = {0.0, 0.0, 0.0};
  double albedo[3] = {0.1, 0.1, 0.1};
  atmo_density_layer ray{0.0, 1.0, -1.0 / 8000.0, 0.0, 0.0};
  atmo_density_layer mie{0.0, 1.0, -1.0 / 1200.0, 0.0, 0.0};
  atmo_model_params params() const {
    atmo_model_params p{};
    p.num_wavelengths = 3;
    p.wavelengths = wl;
    p.solar_irradiance = solar;
    p.sun_angular_radius = 0.00935;
    p.bottom_radius = 6360000.0;
    p.top_radius = 6420000.0;
    p.num_rayleigh_layers = 1;
    p.rayleigh_density = &ray;
    p.rayleigh_scattering = rayleigh;
    p.num_mie_layers = 1;
    p.mie_density = &mie;
    p.mie_scattering = mie_sca;
    p.mie_extinction = mie_ext;
    p.mie_phase_function_g = 0.8;
    p.num_absorption_layers = 0;
    p.absorption_density = nullptr;
    p.absorption_extinction = absorption;
    p.ground_albedo = albedo;
    p.max_sun_zenith_angle = 102.0 / 180.0 * 3.141592653589793;
    p.length_unit_in_meters = 1000.0;
    p.num_precomputed_wavelengths = 3;
    p.combine_scattering_textures = 0;
    p.half_precision = 0;
    return p;
  }
};

// Scattering texels of one tuning batch: enough blocks for every CU of the
// GPU several times over with any tile (2x the reference LUT).
constexpr double kTuneTexels = 2.0 * 1048576.0;
constexpr int kMaxCopies = 64;
constexpr int kReps = 3;
// A shape replaces the default only if it is this much faster: a batch of
// mixed sizes takes one launch per distinct shape of a pass, and the split
// launches fill the GPU less well than the timing of one size class alone
// shows (with 3% the tuned 64-atmosphere batch ran 1-3% slower than the
// default shapes, with 64-65 launches instead of 44).
constexpr double kMinGain = 1.10;

struct DeviceArrays {
  atmo::Job* jobs = nullptr;
  int* first_block = nullptr;
  hipEvent_t start = nullptr, stop = nullptr;
  ~DeviceArrays() {
    if (jobs != nullptr) (void)hipFree(jobs);
    if (first_block != nullptr) (void)hipFree(first_block);
    if (start != nullptr) (void)hipEventDestroy(start);
    if (stop != nullptr) (void)hipEventDestroy(stop);
  }
};

// Times every shape of the passes asked for on k copies of the tuning
// atmosphere at sizes z; writes the fastest shapes.
atmo_status tune_size(const atmo_lut_sizes& lz, bool do_sd, bool do_ms, hipStream_t s,
                      int* sd_out, MsShape* ms_out) {
  const Sizes z = sizes_of(lz);
  const double texels = static_cast<double>(z.s_texels());
  const int k = std::max(1, std::min(kMaxCopies, static_cast<int>(std::ceil(kTuneTexels /
                                                                               texels))));
  const EarthParams ep;
  const atmo_model_params p = ep.params();
  std::vector<std::unique_ptr<atmo_model, void (*)(atmo_model*)>> owned;
  std::vector<atmo_model*> models;
  for (int i = 0; i < k; ++i) {
    atmo_model* m = nullptr;
    ATMO_TRY(atmo_model_create(&p, &lz, &m));
    owned.emplace_back(m, [](atmo_model* x) { (void)atmo_model_destroy(x); });
    models.push_back(m);
    m->retain = true;  // keep the deltas: the timed launches read them
  }
  // Orders 1 and 2 give every input of the timed launches: the delta
  // scattering, irradiance and density LUTs.
  std::vector<unsigned> orders(k, 2);
  ATMO_TRY(atmo_model_init_batch(models.data(), orders.data(), k, s));

  DeviceArrays d;
  TUNE_HIP(hipMalloc(&d.jobs, k * sizeof(atmo::Job)));
  TUNE_HIP(hipMalloc(&d.first_block, (k + 1) * sizeof(int)));
  TUNE_HIP(hipEventCreate(&d.start));
  TUNE_HIP(hipEventCreate(&d.stop));
  const int w = z.nnu * z.nmus;
  auto time_pass = [&](Pass pass, int order, int tile, int row_order,
                       float* best) -> atmo_status {
    std::vector<atmo::Job> js(k);
    std::vector<int> fb(k + 1);
    int total = 0;
    for (int i = 0; i < k; ++i) {
      js[i] = models[i]->jobs[0];
      (pass == kSD ? js[i].sd_tile : js[i].ms_tile) = tile;
      if (pass == kMS) js[i].ms_row_order = row_order;
      fb[i] = total;
      total += pass_blocks(pass, js[i]);
    }
    fb[k] = total;
    TUNE_HIP(hipMemcpy(d.jobs, js.data(), k * sizeof(atmo::Job), hipMemcpyHostToDevice));
    TUNE_HIP(hipMemcpy(d.first_block, fb.data(), (k + 1) * sizeof(int),
                       hipMemcpyHostToDevice));
    atmo::PassLaunch pl;
    pl.jobs = d.jobs;
    pl.first_block = d.first_block;
    pl.njobs = k;
    pl.total_blocks = total;
    pl.max_nnu = z.nnu;
    pl.max_ew = z.ew;
    pl.max_w = w;
    pl.min_w = w;
    pl.max_lds_w = w <= atmo::kMsMaxRow ? w : 0;
    pl.kinds = atmo::kJobsF32;
    pl.dj_kinds = atmo::kDjF32;
    pl.tile = tile;
    auto launch = [&]() {
      return pass == kSD ? atmo::launch_scattering_density(pl, order, s)
                         : atmo::launch_multiple_scattering(pl, s);
    };
    ATMO_TRY(check_launch(launch(), "tuning launch"));  // warm-up
    *best = std::numeric_limits<float>::max();
    for (int r = 0; r < kReps; ++r) {
      TUNE_HIP(hipEventRecord(d.start, s));
      ATMO_TRY(check_launch(launch(), "tuning launch"));
      TUNE_HIP(hipEventRecord(d.stop, s));
      TUNE_HIP(hipEventSynchronize(d.stop));
      float ms = 0.0f;
      TUNE_HIP(hipEventElapsedTime(&ms, d.start, d.stop));
      *best = std::min(*best, ms);
    }
    return ATMO_OK;
  };
  auto pick = [](const std::vector<double>& t, int dflt) {
    int best = dflt;
    for (int i = 0; i < static_cast<int>(t.size()); ++i)
      if (t[i] < t[best]) best = i;
    return t[dflt] >= kMinGain * t[best] ? best : dflt;
  };
  if (do_sd) {
    // One order-2 launch and (for the default 4 orders) two higher-order
    // launches per Init: the order-2 kernel also reads the single-Mie LUT.
    std::vector<double> t(atmo::kSdTiles);
    for (int tile = 0; tile < atmo::kSdTiles; ++tile) {
      float t2 = 0.0f, t3 = 0.0f;
      ATMO_TRY(time_pass(kSD, 2, tile, 0, &t2));
      ATMO_TRY(time_pass(kSD, 3, tile, 0, &t3));
      t[tile] = t2 + 2.0 * t3;
    }
    *sd_out = pick(t, atmo::sd_default_tile(z.nmu, z.nnu));
  }
  // Candidates kMsRowOrders * tile + row order. Rows too long for LDS take
  // the gather kernel whatever the tile: only the row order is timed for
  // them.
  if (do_ms) {
    constexpr int kO = atmo::kMsRowOrders;
    const MsShape dflt = ms_default(z);
    std::vector<double> t(kO * atmo::kMsTiles, std::numeric_limits<double>::max());
    for (int tile = 0; tile < atmo::kMsTiles; ++tile) {
      if (!ms_tile_fits(z, tile)) continue;
      if (w > atmo::kMsMaxRow && tile != dflt.tile) continue;
      for (int o = 0; o < kO; ++o) {
        float tm = 0.0f;
        ATMO_TRY(time_pass(kMS, 0, tile, o, &tm));
        t[kO * tile + o] = tm;
      }
    }
    const int best = pick(t, kO * dflt.tile + dflt.row_order);
    *ms_out = MsShape{best / kO, best % kO};
  }
  TUNE_HIP(hipStreamSynchronize(s));
  return ATMO_OK;
}

}  // namespace

Tiles tiles_for(const Sizes& z) {
  std::lock_guard<std::mutex> lock(g_tiles_mu);
  MsShape ms = ms_default(z);
  const auto b = g_ms.find(ms_key(z));
  if (b != g_ms.end() && ms_tile_fits(z, b->second.tile)) ms = b->second;
  Tiles t{atmo::sd_default_tile(z.nmu, z.nnu), ms.tile, ms.row_order};
  const auto a = g_sd.find(sd_key(z));
  if (a != g_sd.end()) t.sd = a->second;
  return t;
}

}  // namespace atmo_rt

namespace {

// One line of a tile table: a kind word, then at most 'cap' integers in
// [-2^20, 2^20] separated by blanks, and nothing else. Returns the number of
// integers, or -1 for a malformed line (no conversion overflows).
int parse_tile_line(const char* line, std::string* kind, long* v, int cap) {
  auto blank = [](char c) { return c == ' ' || c == '\t'; };
  auto end_of_line = [](char c) { return c == '\n' || c == '\r' || c == '\0'; };
  const char* p = line;
  while (blank(*p)) ++p;
  const char* k0 = p;
  while (!blank(*p) && !end_of_line(*p)) ++p;
  kind->assign(k0, p);
  int n = 0;
  for (;;) {
    while (blank(*p)) ++p;
    if (end_of_line(*p)) break;
    if (n == cap) return -1;
    char* end = nullptr;
    errno = 0;
    const long x = std::strtol(p, &end, 10);
    if (end == p || errno == ERANGE || x < -(1L << 20) || x > (1L << 20)) return -1;
    if (!blank(*end) && !end_of_line(*end)) return -1;
    v[n++] = x;
    p = end;
  }
  return n;
}

}  // namespace

using namespace atmo_rt;

extern "C" {

int atmo_sd_tile_candidates(int* nu_per_thread, int* block_threads, int capacity) {
  for (int t = 0; t < atmo::kSdTiles && t < capacity; ++t) {
    if (nu_per_thread != nullptr) nu_per_thread[t] = atmo::kSdTile[t].chunk;
    if (block_threads != nullptr) block_threads[t] = atmo::kSdTile[t].block;
  }
  return atmo::kSdTiles;
}

int atmo_ms_tile_candidates(int* texels_per_thread, int* block_threads, int capacity) {
  for (int t = 0; t < atmo::kMsTiles && t < capacity; ++t) {
    if (texels_per_thread != nullptr) texels_per_thread[t] = atmo::kMsTile[t].tpt;
    if (block_threads != nullptr) block_threads[t] = atmo::kMsTile[t].block;
  }
  return atmo::kMsTiles;
}

atmo_status atmo_get_tiles(const atmo_lut_sizes* sizes, atmo_tiles* out) {
  if (sizes == nullptr || out == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  ATMO_TRY(validate_sizes(*sizes));
  const Sizes z = sizes_of(*sizes);
  const Tiles t = tiles_for(z);
  out->sd_nu_per_thread = atmo::kSdTile[t.sd].chunk;
  out->sd_block_threads = atmo::kSdTile[t.sd].block;
  out->ms_texels_per_thread = atmo::kMsTile[t.ms].tpt;
  out->ms_block_threads = atmo::kMsTile[t.ms].block;
  out->ms_row_order = t.ms_row_order;
  std::lock_guard<std::mutex> lock(g_tiles_mu);
  out->tuned = (g_sd.count(sd_key(z)) ? 1 : 0) | (g_ms.count(ms_key(z)) ? 2 : 0);
  return ATMO_OK;
}

atmo_status atmo_set_tiles(const atmo_lut_sizes* sizes, const atmo_tiles* tiles) {
  if (sizes == nullptr || tiles == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  ATMO_TRY(validate_sizes(*sizes));
  const Sizes z = sizes_of(*sizes);
  const int sd = sd_tile_of_shape(tiles->sd_nu_per_thread, tiles->sd_block_threads);
  if (sd < 0)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "not a scattering-density tile shape");
  const int ms = ms_tile_of_shape(tiles->ms_texels_per_thread, tiles->ms_block_threads);
  if (ms < 0)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "not a multiple-scattering tile shape");
  if (!ms_tile_fits(z, ms))
    return fail(ATMO_ERROR_INVALID_ARGUMENT,
                "multiple-scattering rows of NU x MU_S texels do not fit this tile");
  if (tiles->ms_row_order < 0 || tiles->ms_row_order >= atmo::kMsRowOrders)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "ms_row_order must be 0, 1 or 2");
  std::lock_guard<std::mutex> lock(g_tiles_mu);
  g_sd[sd_key(z)] = sd;
  g_ms[ms_key(z)] = MsShape{ms, tiles->ms_row_order};
  return ATMO_OK;
}

void atmo_tune_reset(void) {
  std::lock_guard<std::mutex> lock(g_tiles_mu);
  g_sd.clear();
  g_ms.clear();
}

atmo_status atmo_autotune(const atmo_lut_sizes* sizes, int count, int retune,
                          void* stream) {
  if (count < 0 || (count > 0 && sizes == nullptr))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "bad sizes array");
  for (int i = 0; i < count; ++i) ATMO_TRY(validate_sizes(sizes[i]));
  hipStream_t s = static_cast<hipStream_t>(stream);
  // Size classes already settled by this call (or before it, unless retune).
  std::map<SdKey, bool> sd_done;
  std::map<MsKey, bool> ms_done;
  for (int i = 0; i < count; ++i) {
    const Sizes z = sizes_of(sizes[i]);
    bool do_sd = !sd_done.count(sd_key(z));
    bool do_ms = !ms_done.count(ms_key(z));
    if (!retune) {
      std::lock_guard<std::mutex> lock(g_tiles_mu);
      do_sd = do_sd && !g_sd.count(sd_key(z));
      do_ms = do_ms && !g_ms.count(ms_key(z));
    }
    sd_done[sd_key(z)] = true;
    ms_done[ms_key(z)] = true;
    if (!do_sd && !do_ms) continue;
    int sd = atmo::sd_default_tile(z.nmu, z.nnu);
    MsShape ms = ms_default(z);
    ATMO_TRY(tune_size(sizes[i], do_sd, do_ms, s, &sd, &ms));
    std::lock_guard<std::mutex> lock(g_tiles_mu);
    if (do_sd) g_sd[sd_key(z)] = sd;
    if (do_ms) g_ms[ms_key(z)] = ms;
  }
  return ATMO_OK;
}

// Text file, one size class per line:
//   sd <MU> <NU> <nu per thread> <block threads>
//   ms <R> <MU> <MU_S> <NU> <texels per thread> <block threads> <row_order>
atmo_status atmo_tune_save(const char* path) {
  if (path == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "path is NULL");
  FILE* f = std::fopen(path, "w");
  if (f == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, std::string("cannot create ") + path);
  bool ok = std::fprintf(f, "# libatmo tile shapes\n") > 0;
  {
    std::lock_guard<std::mutex> lock(g_tiles_mu);
    for (const auto& e : g_sd)
      ok = ok && std::fprintf(f, "sd %d %d %d %d\n", e.first[0], e.first[1],
                              atmo::kSdTile[e.second].chunk,
                              atmo::kSdTile[e.second].block) > 0;
    for (const auto& e : g_ms)
      ok = ok && std::fprintf(f, "ms %d %d %d %d %d %d %d\n", e.first[0], e.first[1],
                              e.first[2], e.first[3], atmo::kMsTile[e.second.tile].tpt,
                              atmo::kMsTile[e.second.tile].block, e.second.row_order) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(ATMO_ERROR_INVALID_ARGUMENT, std::string("cannot write ") + path);
  return ATMO_OK;
}

atmo_status atmo_tune_load(const char* path) {
  if (path == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "path is NULL");
  FILE* f = std::fopen(path, "r");
  if (f == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  std::map<SdKey, int> sd;
  std::map<MsKey, MsShape> ms;
  char line[256];
  int lineno = 0;
  atmo_status st = ATMO_OK;
  while (st == ATMO_OK && std::fgets(line, sizeof(line), f) != nullptr) {
    ++lineno;
    const std::string where = std::string(path) + ":" + std::to_string(lineno);
    const size_t len = std::strlen(line);
    if (len + 1 == sizeof(line) && line[len - 1] != '\n') {
      st = fail(ATMO_ERROR_INVALID_ARGUMENT, where + ": line too long");
      break;
    }
    if (line[0] == '#' || line[0] == '\n') continue;
    std::string kind;
    long v[7] = {};
    const int n = parse_tile_line(line, &kind, v, 7);
    if (n == 4 && kind == "sd") {
      const int t = sd_tile_of_shape(static_cast<int>(v[2]), static_cast<int>(v[3]));
      if (t < 0 || v[0] < 2 || v[1] < 2)
        st = fail(ATMO_ERROR_INVALID_ARGUMENT, where + ": not a scattering-density entry");
      else
        sd[{static_cast<int>(v[0]), static_cast<int>(v[1])}] = t;
    } else if (n == 7 && kind == "ms") {
      const int t = ms_tile_of_shape(static_cast<int>(v[4]), static_cast<int>(v[5]));
      Sizes z{};
      z.nr = static_cast<int>(v[0]);
      z.nmu = static_cast<int>(v[1]);
      z.nmus = static_cast<int>(v[2]);
      z.nnu = static_cast<int>(v[3]);
      const long h = v[6];
      if (t < 0 || z.nr < 2 || z.nmu < 2 || z.nmus < 2 || z.nnu < 2 || h < 0 ||
          h >= atmo::kMsRowOrders || !ms_tile_fits(z, t))
        st = fail(ATMO_ERROR_INVALID_ARGUMENT, where + ": not a multiple-scattering entry");
      else
        ms[{z.nr, z.nmu, z.nmus, z.nnu}] = MsShape{t, static_cast<int>(h)};
    } else {
      st = fail(ATMO_ERROR_INVALID_ARGUMENT, where + ": unrecognized line");
    }
  }
  std::fclose(f);
  if (st != ATMO_OK) return st;
  std::lock_guard<std::mutex> lock(g_tiles_mu);
  for (const auto& e : sd) g_sd[e.first] = e.second;
  for (const auto& e : ms) g_ms[e.first] = e.second;
  return ATMO_OK;
}

}  // extern "C"
