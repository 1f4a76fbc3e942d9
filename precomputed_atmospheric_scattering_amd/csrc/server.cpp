// Serving loop of libatmo (include/atmo.h, atmo_server_*): continuous
// batching of precompute requests with device-memory admission, resident LUT
// sets, and lookups chained on the device after the precompute they need.
//
// One worker thread owns every HIP call of the server. Precompute batches run
// on up to max_inflight_batches "slots", each with its own stream, reusable
// scratch arena (Init temporaries of the batch's models) and job tables, so
// admitting a batch allocates only the new models' resident LUTs. Freed
// memory (released models, outgrown arenas) is returned only when the device
// is idle, because hipFree waits for the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime_internal.h"

using namespace atmo_rt;

namespace {

// Deep copy of the constructor parameters of a request.
struct OwnedParams {
  std::vector<double> wl, solar, ray, mie_sca, mie_ext, absorption, albedo;
  std::vector<atmo_density_layer> lr, lm, la;
  atmo_model_params p{};

  void copy(const atmo_model_params& src) {
    const int n = src.num_wavelengths;
    auto v = [n](const double* x) { return std::vector<double>(x, x + n); };
    wl = v(src.wavelengths);
    solar = v(src.solar_irradiance);
    ray = v(src.rayleigh_scattering);
    mie_sca = v(src.mie_scattering);
    mie_ext = v(src.mie_extinction);
    absorption = v(src.absorption_extinction);
    albedo = v(src.ground_albedo);
    lr.assign(src.rayleigh_density, src.rayleigh_density + src.num_rayleigh_layers);
    lm.assign(src.mie_density, src.mie_density + src.num_mie_layers);
    la.assign(src.absorption_density, src.absorption_density + src.num_absorption_layers);
    p = src;
    rebind();
  }
  void rebind() {
    p.wavelengths = wl.data();
    p.solar_irradiance = solar.data();
    p.rayleigh_scattering = ray.data();
    p.mie_scattering = mie_sca.data();
    p.mie_extinction = mie_ext.data();
    p.absorption_extinction = absorption.data();
    p.ground_albedo = albedo.data();
    p.rayleigh_density = lr.empty() ? nullptr : lr.data();
    p.mie_density = lm.empty() ? nullptr : lm.data();
    p.absorption_density = la.empty() ? nullptr : la.data();
  }
};

struct PendingPrecompute {
  uint64_t id;
  std::unique_ptr<OwnedParams> params;
  atmo_lut_sizes sizes;
  unsigned orders;
  double submit_ms;
  size_t resident;  // LUT bytes kept after Init
  size_t scratch;   // Init temporaries
};

struct PendingLookup {
  uint64_t id, model_id;
  int lookup, luminance;
  int64_t n;
  const float* in[4];
  float* out[2];
  double submit_ms;
};

struct Resident {
  atmo_model* model = nullptr;
  size_t bytes = 0;
  bool ready = false;  // its precompute has completed
  int slot = -1;       // slot whose batch computes it, while not ready
  int lookups = 0;     // lookups launched and not completed
  bool released = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  atmo::Job* d_jobs = nullptr;
  size_t jobs_cap = 0;
  int* d_fb = nullptr;
  size_t fb_cap = 0;
  atmo::Job* h_jobs = nullptr;  // pinned staging of the job table
  int* h_fb = nullptr;
  LevelPlan plan;
  std::vector<uint64_t> ids;
  std::vector<double> submit;
  double start_ms = 0.0;
};

struct InflightLookup {
  PendingLookup req;
  hipEvent_t done;
  double start_ms;
};

size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Resident and scratch bytes of a model with these sizes and storage (the
// same layout as atmo_model_footprint); iters: its Precompute iterations.
void footprint(const atmo_lut_sizes& z, bool half, bool combined, int iters,
               size_t* resident, size_t* scratch) {
  const size_t nt = static_cast<size_t>(z.transmittance_width) * z.transmittance_height;
  const size_t ne = static_cast<size_t>(z.irradiance_width) * z.irradiance_height;
  const size_t ns = static_cast<size_t>(z.scattering_r) * z.scattering_mu *
                    z.scattering_mu_s * z.scattering_nu;
  const size_t texel = half ? 8 : 16;
  const Sizes sz{z.transmittance_width, z.transmittance_height, z.scattering_r,
                 z.scattering_mu,       z.scattering_mu_s,      z.scattering_nu,
                 z.irradiance_width,    z.irradiance_height};
  // The models keep the default interleaved lookup copy.
  *resident = round256(nt * 16) + round256(ne * 16) +
              (combined ? 1 : 2) * round256(ns * texel) +
              round256(interleaved_bytes(sz, half, combined));
  *scratch = temporaries_bytes_for(sz, half, combined, iters);
}

}  // namespace

struct atmo_server {
  atmo_server_config cfg{};
  int device = 0;
  size_t budget = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stopping = false;
  uint64_t next_id = 1;
  // Guarded by mu.
  std::deque<PendingPrecompute> pending;
  std::deque<PendingLookup> lookups;
  std::vector<uint64_t> releases;
  std::deque<atmo_completion> done;
  std::unordered_map<uint64_t, Resident> models;
  size_t resident_bytes = 0;
  atmo_server_stats counters{};
  // Worker-owned.
  std::vector<Slot> slots;
  std::vector<hipStream_t> lstreams;
  size_t next_lstream = 0;
  std::vector<InflightLookup> inflight;
  std::vector<hipEvent_t> free_events;
  // Released models and their admitted resident bytes, freed when idle.
  std::vector<std::pair<atmo_model*, size_t>> graveyard;
  std::vector<void*> graveyard_ptrs;   // outgrown buffers, freed when idle
  std::thread worker;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
  }

  void complete(uint64_t id, int kind, atmo_status st, uint64_t model, int batch,
                double submit, double start, double end) {
    atmo_completion c;
    c.request_id = id;
    c.kind = kind;
    c.status = st;
    c.model_id = model;
    c.batch_models = batch;
    c.submit_ms = submit;
    c.start_ms = start;
    c.end_ms = end;
    done.push_back(c);
    if (st != ATMO_OK) counters.failed++;
    else if (kind == ATMO_REQUEST_PRECOMPUTE) counters.completed_precomputes++;
    else counters.completed_lookups++;
    cv_done.notify_all();
  }

  size_t arenas_bytes() const {
    size_t b = 0;
    for (const Slot& s : slots) b += s.arena_bytes;
    return b;
  }

  bool any_inflight() const {
    if (!inflight.empty()) return true;
    for (const Slot& s : slots)
      if (s.busy) return true;
    return false;
  }

  hipEvent_t take_event() {
    if (!free_events.empty()) {
      hipEvent_t e = free_events.back();
      free_events.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
  }

  // Completed batches: their models become ready, their arenas reusable.
  bool poll_slots() {
    bool progress = false;
    for (size_t k = 0; k < slots.size(); ++k) {
      Slot& s = slots[k];
      if (!s.busy) continue;
      hipError_t e = hipEventQuery(s.done);
      if (e == hipErrorNotReady) continue;
      const double end = now_ms();
      std::lock_guard<std::mutex> lk(mu);
      for (size_t i = 0; i < s.ids.size(); ++i) {
        auto it = models.find(s.ids[i]);
        atmo_status st = e == hipSuccess ? ATMO_OK : ATMO_ERROR_DEVICE;
        if (it != models.end()) {
          drop_external_temporaries(it->second.model);
          it->second.ready = true;
          it->second.slot = -1;
        }
        complete(s.ids[i], ATMO_REQUEST_PRECOMPUTE, st, s.ids[i],
                 static_cast<int>(s.ids.size()), s.submit[i], s.start_ms, end);
      }
      s.busy = false;
      s.ids.clear();
      s.submit.clear();
      progress = true;
    }
    return progress;
  }

  bool poll_lookups() {
    bool progress = false;
    for (size_t i = 0; i < inflight.size();) {
      hipError_t e = hipEventQuery(inflight[i].done);
      if (e == hipErrorNotReady) {
        ++i;
        continue;
      }
      const double end = now_ms();
      {
        std::lock_guard<std::mutex> lk(mu);
        auto it = models.find(inflight[i].req.model_id);
        if (it != models.end()) it->second.lookups--;
        complete(inflight[i].req.id, ATMO_REQUEST_LOOKUP,
                 e == hipSuccess ? ATMO_OK : ATMO_ERROR_DEVICE, inflight[i].req.model_id, 0,
                 inflight[i].req.submit_ms, inflight[i].start_ms, end);
      }
      free_events.push_back(inflight[i].done);
      inflight[i] = inflight.back();
      inflight.pop_back();
      progress = true;
    }
    return progress;
  }

  // Released models without pending lookups go to the graveyard; the
  // graveyard is freed when the device is idle.
  bool handle_releases() {
    bool progress = false;
    std::lock_guard<std::mutex> lk(mu);
    for (auto it = models.begin(); it != models.end();) {
      Resident& r = it->second;
      if (r.released && r.ready && r.lookups == 0) {
        graveyard.emplace_back(r.model, r.bytes);
        it = models.erase(it);
        progress = true;
      } else {
        ++it;
      }
    }
    if (!graveyard.empty() || !graveyard_ptrs.empty()) {
      if (!any_inflight()) {
        for (auto& g : graveyard) {
          resident_bytes -= std::min(resident_bytes, g.second);
          atmo_model_destroy(g.first);
        }
        graveyard.clear();
        for (void* p : graveyard_ptrs) (void)hipFree(p);
        graveyard_ptrs.clear();
        progress = true;
      }
    }
    return progress;
  }

  // Grow a device buffer (the old one is retired to the graveyard).
  template <class T>
  bool ensure(T** ptr, size_t* cap, size_t need) {
    if (*cap >= need) return true;
    size_t n = std::max(need, *cap * 2);
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return false;
    if (*ptr != nullptr) graveyard_ptrs.push_back(*ptr);
    *ptr = static_cast<T*>(p);
    *cap = n;
    return true;
  }

  // Admit pending precomputes into free slots (FIFO, by memory footprint).
  bool admit() {
    bool progress = false;
    for (size_t k = 0; k < slots.size(); ++k) {
      Slot& s = slots[k];
      if (s.busy) continue;
      std::vector<PendingPrecompute> batch;
      size_t scratch = 0;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (pending.empty()) return progress;
        const size_t others = arenas_bytes() - s.arena_bytes;
        size_t add = 0;
        while (!pending.empty() &&
               batch.size() < static_cast<size_t>(cfg.max_batch_models)) {
          const PendingPrecompute& r = pending.front();
          const size_t arena = std::max(s.arena_bytes, scratch + r.scratch);
          if (resident_bytes + add + r.resident + others + arena > budget) break;
          add += r.resident;
          scratch += r.scratch;
          batch.push_back(std::move(pending.front()));
          pending.pop_front();
        }
        if (batch.empty()) {
          // The head request does not fit now. One that cannot fit even in
          // an empty server fails; otherwise it waits for batches to finish
          // and models to be released. Arenas are returned to the device
          // while no batch runs, so that they do not hold memory it needs.
          const PendingPrecompute& r = pending.front();
          if (r.resident + r.scratch > budget) {
            PendingPrecompute f = std::move(pending.front());
            pending.pop_front();
            complete(f.id, ATMO_REQUEST_PRECOMPUTE, ATMO_ERROR_OUT_OF_MEMORY, f.id, 0,
                     f.submit_ms, now_ms(), now_ms());
            return true;
          }
          bool batches = false;
          for (const Slot& o : slots) batches |= o.busy;
          if (!batches && arenas_bytes() > 0) {
            for (Slot& o : slots) {
              if (o.arena != nullptr) (void)hipFree(o.arena);
              o.arena = nullptr;
              o.arena_bytes = 0;
            }
            return true;
          }
          return progress;
        }
        resident_bytes += add;
      }
      launch_batch(s, batch, scratch);
      progress = true;
    }
    return progress;
  }

  void fail_batch(std::vector<PendingPrecompute>& batch, std::vector<atmo_model*>& made,
                  atmo_status st) {
    std::lock_guard<std::mutex> lk(mu);
    for (size_t i = 0; i < made.size(); ++i) {
      if (made[i] == nullptr) continue;
      drop_external_temporaries(made[i]);
      graveyard.emplace_back(made[i], 0);
    }
    for (PendingPrecompute& r : batch) {
      resident_bytes -= std::min(resident_bytes, r.resident);
      models.erase(r.id);
      complete(r.id, ATMO_REQUEST_PRECOMPUTE, st, r.id, 0, r.submit_ms, now_ms(), now_ms());
    }
  }

  void launch_batch(Slot& s, std::vector<PendingPrecompute>& batch, size_t scratch) {
    std::vector<atmo_model*> made(batch.size(), nullptr);
    if (s.arena_bytes < scratch) {
      if (s.arena != nullptr) graveyard_ptrs.push_back(s.arena);
      s.arena = nullptr;
      s.arena_bytes = 0;
      void* p = nullptr;
      if (hipMalloc(&p, scratch) != hipSuccess) {
        fail_batch(batch, made, ATMO_ERROR_OUT_OF_MEMORY);
        return;
      }
      s.arena = static_cast<char*>(p);
      s.arena_bytes = scratch;
    }
    std::vector<unsigned> orders;
    size_t offset = 0;
    for (size_t i = 0; i < batch.size(); ++i) {
      atmo_status st = atmo_model_create(&batch[i].params->p, &batch[i].sizes, &made[i]);
      if (st != ATMO_OK) {
        fail_batch(batch, made, st);
        return;
      }
      use_external_temporaries(made[i], s.arena + offset);
      offset += temporaries_bytes(made[i]);
      made[i]->initialized = true;  // lookups are ordered after the batch
      if (upload_jobs(made[i], batch[i].orders, s.stream) != ATMO_OK) {
        fail_batch(batch, made, ATMO_ERROR_DEVICE);
        return;
      }
      orders.push_back(batch[i].orders);
    }
    plan_levels(made.data(), orders.data(), static_cast<int>(made.size()), &s.plan);
    const size_t nj = s.plan.jobs.size(), nf = s.plan.first_block.size();
    size_t jcap = s.jobs_cap, fcap = s.fb_cap;
    if (!ensure(&s.d_jobs, &s.jobs_cap, nj) || !ensure(&s.d_fb, &s.fb_cap, nf)) {
      fail_batch(batch, made, ATMO_ERROR_OUT_OF_MEMORY);
      return;
    }
    if (s.h_jobs == nullptr || jcap < nj || s.jobs_cap != jcap) {
      if (s.h_jobs != nullptr) (void)hipHostFree(s.h_jobs);
      if (hipHostMalloc(&s.h_jobs, s.jobs_cap * sizeof(atmo::Job)) != hipSuccess) {
        s.h_jobs = nullptr;
        fail_batch(batch, made, ATMO_ERROR_OUT_OF_MEMORY);
        return;
      }
    }
    if (s.h_fb == nullptr || fcap < nf || s.fb_cap != fcap) {
      if (s.h_fb != nullptr) (void)hipHostFree(s.h_fb);
      if (hipHostMalloc(&s.h_fb, s.fb_cap * sizeof(int)) != hipSuccess) {
        s.h_fb = nullptr;
        fail_batch(batch, made, ATMO_ERROR_OUT_OF_MEMORY);
        return;
      }
    }
    std::memcpy(s.h_jobs, s.plan.jobs.data(), nj * sizeof(atmo::Job));
    std::memcpy(s.h_fb, s.plan.first_block.data(), nf * sizeof(int));
    if (hipMemcpyAsync(s.d_jobs, s.h_jobs, nj * sizeof(atmo::Job), hipMemcpyHostToDevice,
                       s.stream) != hipSuccess ||
        hipMemcpyAsync(s.d_fb, s.h_fb, nf * sizeof(int), hipMemcpyHostToDevice,
                       s.stream) != hipSuccess ||
        enqueue_levels(s.plan, s.d_jobs, s.d_fb, s.stream) != ATMO_OK ||
        hipEventRecord(s.done, s.stream) != hipSuccess) {
      fail_batch(batch, made, ATMO_ERROR_DEVICE);
      return;
    }
    s.busy = true;
    s.start_ms = now_ms();
    s.ids.clear();
    s.submit.clear();
    std::lock_guard<std::mutex> lk(mu);
    for (size_t i = 0; i < batch.size(); ++i) {
      Resident r;
      r.model = made[i];
      r.bytes = batch[i].resident;
      r.slot = static_cast<int>(&s - slots.data());
      models[batch[i].id] = r;
      s.ids.push_back(batch[i].id);
      s.submit.push_back(batch[i].submit_ms);
    }
  }

  bool launch_lookups() {
    bool progress = false;
    std::unique_lock<std::mutex> lk(mu);
    for (auto it = lookups.begin(); it != lookups.end();) {
      PendingLookup q = *it;
      auto mit = models.find(q.model_id);
      if (mit == models.end() || mit->second.released) {
        bool waiting = false;
        for (const PendingPrecompute& p : pending)
          if (p.id == q.model_id) waiting = true;
        if (!waiting) {
          complete(q.id, ATMO_REQUEST_LOOKUP, ATMO_ERROR_NOT_INITIALIZED, q.model_id, 0,
                   q.submit_ms, now_ms(), now_ms());
          it = lookups.erase(it);
          progress = true;
        } else {
          ++it;
        }
        continue;
      }
      Resident& r = mit->second;
      hipStream_t st = lstreams[next_lstream++ % lstreams.size()];
      hipEvent_t ev = take_event();
      atmo_status status = ev == nullptr ? ATMO_ERROR_DEVICE : ATMO_OK;
      if (status == ATMO_OK && !r.ready && r.slot >= 0 &&
          hipStreamWaitEvent(st, slots[r.slot].done, 0) != hipSuccess)
        status = ATMO_ERROR_DEVICE;
      if (status == ATMO_OK) {
        switch (q.lookup) {
          case ATMO_LOOKUP_SKY:
            status = q.luminance
                         ? atmo_get_sky_luminance(r.model, q.n, q.in[0], q.in[1], q.in[2],
                                                  q.in[3], q.out[0], q.out[1], st)
                         : atmo_get_sky_radiance(r.model, q.n, q.in[0], q.in[1], q.in[2],
                                                 q.in[3], q.out[0], q.out[1], st);
            break;
          case ATMO_LOOKUP_SKY_TO_POINT:
            status = q.luminance ? atmo_get_sky_luminance_to_point(
                                       r.model, q.n, q.in[0], q.in[1], q.in[2], q.in[3],
                                       q.out[0], q.out[1], st)
                                 : atmo_get_sky_radiance_to_point(
                                       r.model, q.n, q.in[0], q.in[1], q.in[2], q.in[3],
                                       q.out[0], q.out[1], st);
            break;
          default:
            status = q.luminance
                         ? atmo_get_sun_and_sky_illuminance(r.model, q.n, q.in[0], q.in[1],
                                                            q.in[2], q.out[0], q.out[1], st)
                         : atmo_get_sun_and_sky_irradiance(r.model, q.n, q.in[0], q.in[1],
                                                           q.in[2], q.out[0], q.out[1], st);
        }
      }
      if (status == ATMO_OK && hipEventRecord(ev, st) != hipSuccess) status = ATMO_ERROR_DEVICE;
      if (status != ATMO_OK) {
        if (ev != nullptr) free_events.push_back(ev);
        complete(q.id, ATMO_REQUEST_LOOKUP, status, q.model_id, 0, q.submit_ms, now_ms(),
                 now_ms());
      } else {
        r.lookups++;
        inflight.push_back({q, ev, now_ms()});
      }
      it = lookups.erase(it);
      progress = true;
    }
    return progress;
  }

  void run() {
    (void)hipSetDevice(device);
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        const bool idle = pending.empty() && lookups.empty() && !any_inflight() &&
                          graveyard.empty() && graveyard_ptrs.empty();
        bool has_released = false;
        for (auto& kv : models)
          if (kv.second.released) has_released = true;
        if (idle && !has_released) {
          if (stopping) break;
          cv_work.wait_for(lk, std::chrono::milliseconds(5));
          continue;
        }
      }
      bool progress = poll_slots();
      progress |= poll_lookups();
      progress |= handle_releases();
      progress |= admit();
      progress |= launch_lookups();
      if (!progress) {
        std::unique_lock<std::mutex> lk(mu);
        const bool device_work = any_inflight();
        if (stopping && pending.empty() && lookups.empty() && !device_work) break;
        // Device work is polled; otherwise only a submission or a release
        // can make progress possible.
        if (device_work)
          cv_work.wait_for(lk, std::chrono::microseconds(50));
        else
          cv_work.wait_for(lk, std::chrono::milliseconds(2));
      }
    }
    // Drain and free everything.
    (void)hipDeviceSynchronize();
    for (auto& kv : models) atmo_model_destroy(kv.second.model);
    models.clear();
    for (auto& g : graveyard) atmo_model_destroy(g.first);
    for (void* p : graveyard_ptrs) (void)hipFree(p);
    for (Slot& s : slots) {
      if (s.arena) (void)hipFree(s.arena);
      if (s.d_jobs) (void)hipFree(s.d_jobs);
      if (s.d_fb) (void)hipFree(s.d_fb);
      if (s.h_jobs) (void)hipHostFree(s.h_jobs);
      if (s.h_fb) (void)hipHostFree(s.h_fb);
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (hipStream_t st : lstreams) (void)hipStreamDestroy(st);
    for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
    for (InflightLookup& l : inflight) (void)hipEventDestroy(l.done);
  }
};

extern "C" {

atmo_status atmo_server_create(const atmo_server_config* config, atmo_server** out) {
  if (out == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  std::unique_ptr<atmo_server> s(new atmo_server());
  if (config != nullptr) s->cfg = *config;
  if (s->cfg.max_inflight_batches <= 0) s->cfg.max_inflight_batches = 3;
  if (s->cfg.max_batch_models <= 0) s->cfg.max_batch_models = 16;
  if (s->cfg.lookup_streams <= 0) s->cfg.lookup_streams = 2;
  if (hipGetDevice(&s->device) != hipSuccess)
    return fail(ATMO_ERROR_DEVICE, "hipGetDevice failed");
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
    return fail(ATMO_ERROR_DEVICE, "hipMemGetInfo failed");
  s->budget = s->cfg.memory_budget > 0 ? s->cfg.memory_budget
                                       : static_cast<size_t>(0.8 * static_cast<double>(free_b));
  s->slots.resize(s->cfg.max_inflight_batches);
  auto release_streams = [&s] {
    for (Slot& sl : s->slots) {
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (hipStream_t st : s->lstreams) (void)hipStreamDestroy(st);
  };
  for (Slot& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      release_streams();
      return fail(ATMO_ERROR_DEVICE, "stream/event creation failed");
    }
  }
  for (int i = 0; i < s->cfg.lookup_streams; ++i) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      release_streams();
      return fail(ATMO_ERROR_DEVICE, "stream creation failed");
    }
    s->lstreams.push_back(st);
  }
  s->t0 = std::chrono::steady_clock::now();
  atmo_server* raw = s.get();
  s->worker = std::thread([raw] { raw->run(); });
  *out = s.release();
  return ATMO_OK;
}

void atmo_server_destroy(atmo_server* s) {
  if (s == nullptr) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    s->pending.clear();
    s->lookups.clear();
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  delete s;
}

atmo_status atmo_server_submit_precompute(atmo_server* s, const atmo_model_params* params,
                                          const atmo_lut_sizes* sizes, unsigned orders,
                                          uint64_t* request_id) {
  if (s == nullptr || params == nullptr || request_id == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  if (orders < 1) return fail(ATMO_ERROR_INVALID_ARGUMENT, "num_scattering_orders must be >= 1");
  ATMO_TRY(validate_params(params));
  atmo_lut_sizes z = sizes != nullptr ? *sizes
                                      : atmo_lut_sizes{256, 64, 32, 128, 32, 8, 64, 16};
  ATMO_TRY(validate_sizes(z));
  PendingPrecompute r;
  r.params.reset(new OwnedParams());
  r.params->copy(*params);
  r.sizes = z;
  r.orders = orders;
  footprint(z, params->half_precision != 0, params->combine_scattering_textures != 0,
            atmo_num_iterations(params), &r.resident, &r.scratch);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r.id = s->next_id++;
    r.submit_ms = s->now_ms();
    *request_id = r.id;
    s->pending.push_back(std::move(r));
  }
  s->cv_work.notify_all();
  return ATMO_OK;
}

atmo_status atmo_server_submit_lookup(atmo_server* s, uint64_t model_id, atmo_lookup lookup,
                                      int luminance, int64_t n, const float* a, const float* b,
                                      const float* c, const float* d, float* out0, float* out1,
                                      uint64_t* request_id) {
  if (s == nullptr || request_id == nullptr)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  if (lookup < ATMO_LOOKUP_SKY || lookup > ATMO_LOOKUP_SUN_AND_SKY || n < 0)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "bad lookup request");
  PendingLookup q;
  q.model_id = model_id;
  q.lookup = lookup;
  q.luminance = luminance;
  q.n = n;
  q.in[0] = a;
  q.in[1] = b;
  q.in[2] = c;
  q.in[3] = d;
  q.out[0] = out0;
  q.out[1] = out1;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    q.id = s->next_id++;
    q.submit_ms = s->now_ms();
    *request_id = q.id;
    s->lookups.push_back(q);
  }
  s->cv_work.notify_all();
  return ATMO_OK;
}

atmo_status atmo_server_poll(atmo_server* s, atmo_completion* out, int capacity, int* count,
                             int timeout_ms) {
  if (s == nullptr || count == nullptr || (capacity > 0 && out == nullptr))
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && timeout_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                        [s] { return !s->done.empty(); });
  int k = 0;
  while (k < capacity && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *count = k;
  return ATMO_OK;
}

atmo_status atmo_server_release(atmo_server* s, uint64_t model_id) {
  if (s == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "server is NULL");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->models.find(model_id);
    if (it == s->models.end()) return fail(ATMO_ERROR_INVALID_ARGUMENT, "unknown model id");
    it->second.released = true;
  }
  s->cv_work.notify_all();
  return ATMO_OK;
}

atmo_status atmo_server_model(atmo_server* s, uint64_t model_id, atmo_model** model) {
  if (s == nullptr || model == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->models.find(model_id);
  if (it == s->models.end() || it->second.released)
    return fail(ATMO_ERROR_INVALID_ARGUMENT, "unknown model id");
  if (!it->second.ready) return fail(ATMO_ERROR_NOT_INITIALIZED, "precompute not finished");
  *model = it->second.model;
  return ATMO_OK;
}

atmo_status atmo_server_get_stats(atmo_server* s, atmo_server_stats* st) {
  if (s == nullptr || st == nullptr) return fail(ATMO_ERROR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->counters;
  st->budget_bytes = s->budget;
  st->resident_bytes = s->resident_bytes;
  st->scratch_bytes = 0;
  for (const Slot& sl : s->slots) st->scratch_bytes += sl.arena_bytes;
  st->resident_models = static_cast<int>(s->models.size());
  st->pending_precomputes = static_cast<int>(s->pending.size());
  st->pending_lookups = static_cast<int>(s->lookups.size());
  st->inflight_batches = 0;
  for (const Slot& sl : s->slots)
    if (sl.busy) st->inflight_batches++;
  st->inflight_lookups = static_cast<int>(s->inflight.size());
  return ATMO_OK;
}

}  // extern "C"
