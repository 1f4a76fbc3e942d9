// On-device rate ceilings for the roofline figures of bench.py
// (libatmo_peaks.so, a measurement tool; the product library does not use
// it). Each figure is the best of a few timed launches (HIP events):
//
//   out[0] FP32 FMA issue rate: 8 independent v_fma_f32 chains per lane,
//          lane-instructions/s (plain, non-packed instructions);
//   out[1] VALU issue rate of a mix that the SIMDs issue two at a time: per
//          chain step one v_fma_f32, one v_add_f32 and one v_fract_f32,
//          lane-instructions/s;
//   out[2] FP64 FMA issue rate (v_fma_f64 chains), lane-instructions/s;
//   out[3] HBM streaming read, bytes/s (float4 loads over 2 GiB);
//   out[4] HBM copy, bytes read + written per second (1 GiB to 1 GiB);
//   out[5] random 128-byte line gathers over 1 GiB, lines/s;
//   out[6] random 128-byte line gathers over 2.25 GiB, lines/s;
//   out[7] VALU issue rate of one v_fma_f32 and one v_add_f32 per chain
//          step, lane-instructions/s (FP32 arithmetic with FP32 arithmetic);
//   out[8] VALU issue rate of one v_fma_f32 and one v_lshl_add_u32 per
//          chain step, lane-instructions/s (FP32 arithmetic with integer).
//
// The gathers read one 128-byte block per lane at a hashed block index (the
// access pattern of a sky lookup on random rays: every block a different
// line, none reused).
//
// atmo_peaks_flush(stream) writes a 512 MiB buffer on a stream: launches
// enqueued after it find none of the data of earlier launches in the L2s or
// the 256 MB Infinity Cache, and their start is not delayed by the host's
// enqueue time (the GPU is busy with the fill meanwhile), so HIP events
// around them time the kernels alone.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_fma32(float* out, int iters, float b, float c) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = threadIdx.x * 1e-3f + k;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = fmaf(a[k], b, c);
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  if (s == 12345.678f) out[blockIdx.x] = s;  // keeps the chains live
}

__global__ __launch_bounds__(kThreads) void k_mix32(float* out, int iters, float b, float c) {
  float a[8], t[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = threadIdx.x * 1e-3f + k;
    t[k] = 0.5f + k * 1e-2f;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = fmaf(a[k], b, c);
      t[k] = __builtin_amdgcn_fractf(t[k] + 0.37f);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k] + t[k];
  if (s == 12345.678f) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_fma_add32(float* out, int iters, float b, float c) {
  float a[8], t[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = threadIdx.x * 1e-3f + k;
    t[k] = 0.5f + k * 1e-2f;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = fmaf(a[k], b, c);
      t[k] = t[k] + c;
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k] + t[k];
  if (s == 12345.678f) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_fma_int(float* out, int iters, float b, float c) {
  float a[8];
  unsigned q[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = threadIdx.x * 1e-3f + k;
    q[k] = threadIdx.x + k;
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = fmaf(a[k], b, c);
      q[k] = (q[k] << 5) + 77u;
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k] + static_cast<float>(q[k]);
  if (s == 12345.678f) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_fma64(double* out, int iters, double b, double c) {
  double a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = threadIdx.x * 1e-3 + k;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = fma(a[k], b, c);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  if (s == 12345.678) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_stream(const float4* __restrict__ buf, size_t n4,
                                                     float* __restrict__ out) {
  float s = 0.0f;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n4;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const float4 v = buf[i];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 1234.5f) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_copy(const float4* __restrict__ a,
                                                   float4* __restrict__ b, size_t n4) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n4;
       i += static_cast<size_t>(gridDim.x) * kThreads)
    b[i] = a[i];
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// Lane i reads block hash(i) % nblocks: 8 float4, one 128-byte line.
__global__ __launch_bounds__(kThreads) void k_gather(const float4* __restrict__ buf,
                                                     uint32_t nblocks, uint32_t n,
                                                     uint32_t seed, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint32_t blk = hash32(i * 2654435761u + seed) % nblocks;
  const float4* p = buf + static_cast<size_t>(blk) * 8;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float4 v = p[k];
    s += v.x + v.y + v.z + v.w;
  }
  out[i] = s;
}

__global__ __launch_bounds__(kThreads) void k_fill(float4* __restrict__ p, size_t n4) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n4;
       i += static_cast<size_t>(gridDim.x) * kThreads)
    p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

struct Timer {
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t init() {
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    return e;
  }
  ~Timer() {
    if (e0 != nullptr) (void)hipEventDestroy(e0);
    if (e1 != nullptr) (void)hipEventDestroy(e1);
    if (s != nullptr) (void)hipStreamDestroy(s);
  }
  // Best (shortest) of 'reps' timed runs of f(stream), after one untimed run.
  template <class F>
  hipError_t best_ms(int reps, F&& f, double* ms) {
    f(s);
    hipError_t e = hipGetLastError();
    double best = 1e30;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      e = hipEventRecord(e0, s);
      if (e != hipSuccess) break;
      f(s);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e == hipSuccess && t < best) best = t;
    }
    *ms = best;
    return e;
  }
};

#define PEAKS_TRY(call)                  \
  do {                                   \
    const hipError_t e_ = (call);        \
    if (e_ != hipSuccess) {              \
      err = e_;                          \
      goto done;                         \
    }                                    \
  } while (0)

}  // namespace

extern "C" {

// Fills out[0..n) (n <= 9) with the figures listed above. Returns 0, or the
// HIP error code of the first failing call. Allocates 2.25 GiB + 32 MiB of
// device memory for the duration of the call.
int atmo_peaks_measure(double* out, int n) {
  if (out == nullptr || n < 0) return static_cast<int>(hipErrorInvalidValue);
  double r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  hipError_t err = hipSuccess;
  Timer tm;
  float4* buf = nullptr;
  float* sink = nullptr;
  const size_t kBuf = (size_t(9) << 30) / 4;  // 2.25 GiB
  const uint32_t kLanes = 1u << 23;
  const int kBlocks = 256 * 8 * 4;  // 8 blocks of 256 threads per CU, 4 rounds
  const int kIters = 4096;
  double ms = 0.0;
  PEAKS_TRY(tm.init());
  PEAKS_TRY(hipMalloc(&buf, kBuf));
  PEAKS_TRY(hipMalloc(&sink, size_t(kLanes) * sizeof(float)));
  PEAKS_TRY(hipMemsetAsync(buf, 0, kBuf, tm.s));
  {
    const double lane_steps = double(kBlocks) * kThreads * kIters * 8;
    PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_fma32, dim3(kBlocks), dim3(kThreads), 0, s, sink, kIters, 0.999f,
                         1e-3f);
    }, &ms));
    r[0] = lane_steps / (ms * 1e-3);
    PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_mix32, dim3(kBlocks), dim3(kThreads), 0, s, sink, kIters, 0.999f,
                         1e-3f);
    }, &ms));
    r[1] = 3.0 * lane_steps / (ms * 1e-3);  // fma, add, fract per chain step
    PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_fma64, dim3(kBlocks), dim3(kThreads), 0, s,
                         reinterpret_cast<double*>(sink), kIters, 0.999, 1e-3);
    }, &ms));
    r[2] = lane_steps / (ms * 1e-3);
    PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_fma_add32, dim3(kBlocks), dim3(kThreads), 0, s, sink, kIters,
                         0.999f, 1e-3f);
    }, &ms));
    r[7] = 2.0 * lane_steps / (ms * 1e-3);
    PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_fma_int, dim3(kBlocks), dim3(kThreads), 0, s, sink, kIters, 0.999f,
                         1e-3f);
    }, &ms));
    r[8] = 2.0 * lane_steps / (ms * 1e-3);
  }
  {
    const size_t read_bytes = size_t(2) << 30;
    PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_stream, dim3(256 * 16), dim3(kThreads), 0, s, buf, read_bytes / 16,
                         sink);
    }, &ms));
    r[3] = read_bytes / (ms * 1e-3);
    const size_t half = size_t(1) << 30;
    PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
      hipLaunchKernelGGL(k_copy, dim3(256 * 16), dim3(kThreads), 0, s, buf, buf + half / 16,
                         half / 16);
    }, &ms));
    r[4] = 2.0 * half / (ms * 1e-3);
  }
  {
    const size_t spans[2] = {size_t(1) << 30, kBuf};
    for (int k = 0; k < 2; ++k) {
      const uint32_t nblocks = static_cast<uint32_t>(spans[k] / 128);
      uint32_t seed = 1;
      PEAKS_TRY(tm.best_ms(3, [&](hipStream_t s) {
        hipLaunchKernelGGL(k_gather, dim3((kLanes + kThreads - 1) / kThreads), dim3(kThreads),
                           0, s, buf, nblocks, kLanes, seed++, sink);
      }, &ms));
      r[5 + k] = kLanes / (ms * 1e-3);
    }
  }
done:
  if (buf != nullptr) (void)hipFree(buf);
  if (sink != nullptr) (void)hipFree(sink);
  for (int k = 0; k < n && k < 9; ++k) out[k] = r[k];
  return static_cast<int>(err);
}

// Fills the flush buffer on 'stream' (a hipStream_t; NULL = the default
// stream). The buffer is allocated on the current device by the first call
// and kept until the process ends; later calls must use that device, from
// one thread at a time (a measurement tool). Returns 0 or the HIP error
// code.
int atmo_peaks_flush(void* stream) {
  static float4* buf = nullptr;
  const size_t kBytes = size_t(512) << 20;
  if (buf == nullptr) {
    const hipError_t e = hipMalloc(&buf, kBytes);
    if (e != hipSuccess) {
      buf = nullptr;
      return static_cast<int>(e);
    }
  }
  hipLaunchKernelGGL(k_fill, dim3(256 * 16), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), buf, kBytes / sizeof(float4));
  return static_cast<int>(hipGetLastError());
}

}  // extern "C"
