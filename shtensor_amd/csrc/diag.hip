// Measurement helpers (no counterpart in the reference): the FP64 MFMA rate
// a plain register-resident loop sustains on this device, and the rate of
// the grouped A-stationary kernel's step shape with its operands in cache,
// used to put the contraction kernels' matrix-core utilisation next to
// achievable rates.

#include <type_traits>

#include "common.hpp"

namespace sht {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d8 __attribute__((ext_vector_type(2), aligned(8)));

// 16 independent 16x16 accumulators per wave, iters x 16 MFMAs, operands in
// registers; the result is stored so the loop is not removed.
__global__ __launch_bounds__(256) void k_mfma_loop(double* sink, int iters, double x)
{
  constexpr int NA = 16;
  v4d acc[NA];
#pragma unroll
  for (int i = 0; i < NA; ++i) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};
  const double a = x + threadIdx.x, b = 2.0 * x + threadIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < NA; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NA; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  sink[static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x] = s;
}

// The step of the grouped A-stationary kernel without its memory misses,
// sweep boundaries, ghost waves or tail: per 4-k step every wave loads one
// 16-byte A pair (from a cache-resident buffer, four steps in flight), reads
// NS panel fragments from a four-stage LDS ring and issues 2 x NS MFMAs
// (F = 0 FP64, 1 FP32); every two steps it writes two stage rows and meets
// the workgroup barrier.  Four waves per workgroup, two workgroups per CU.
template <int F, int NS>
__global__ __launch_bounds__(256) void k_step_shape(const double* abuf, uint32_t amask, double* sink, int steps,
                                                    double x)
{
  typedef typename std::conditional<F == 0, double, float>::type TB;
  typedef typename std::conditional<F == 0, v4d, v4f>::type Acc;
  __shared__ TB stage[4][8][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int s = 0; s < 4; ++s)
    for (int q = wave; q < 8; q += 4) stage[s][q][lane] = static_cast<TB>(x + s + q + lane);
  __syncthreads();
  Acc acc[2][NS];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[t][q] = Acc{0, 0, 0, 0};
  double ar[4][2];
  uint32_t off = (blockIdx.x * 256u + threadIdx.x) * 2u;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const v2d8 v = *reinterpret_cast<const v2d8*>(abuf + (off & amask));
    ar[i][0] = v.x;
    ar[i][1] = v.y;
    off += 512u;
  }
  TB g = static_cast<TB>(x);
  auto phase = [&](auto jc) {
    constexpr int j = decltype(jc)::value;
    const v2d8 v = *reinterpret_cast<const v2d8*>(abuf + (off & amask));
    ar[(j + 3) % 4][0] = v.x;
    ar[(j + 3) % 4][1] = v.y;
    off += 512u;
    __builtin_amdgcn_sched_barrier(0);
    TB b[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) b[q] = stage[j][q][lane];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      if constexpr (F == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
          acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[q], ar[j][t], acc[t][q], 0, 0, 0);
      } else {
        const float a = static_cast<float>(ar[j][t]);
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[q], a, acc[t][q], 0, 0, 0);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (j % 2 == 1) {
      stage[(j + 1) % 4][wave][lane] = g;
      stage[(j + 2) % 4][wave + 4][lane] = g;
      g += TB(1);
      __syncthreads();
    }
  };
  for (int st = 0; st < steps; st += 4) {
    phase(std::integral_constant<int, 0>());
    phase(std::integral_constant<int, 1>());
    phase(std::integral_constant<int, 2>());
    phase(std::integral_constant<int, 3>());
  }
  double s = static_cast<double>(g);
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int q = 0; q < NS; ++q) s += acc[t][q][0] + acc[t][q][1] + acc[t][q][2] + acc[t][q][3];
  sink[static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x] = s;
}

template <int F, int NS>
void launch_step_shape(int blocks, size_t dyn, hipStream_t s, const double* abuf, uint32_t amask, double* sink,
                       int steps)
{
  hipLaunchKernelGGL((k_step_shape<F, NS>), dim3(blocks), dim3(256), dyn, s, abuf, amask, sink, steps, 1.0);
}

template <int F>
void launch_step_shape_ns(int ns, int blocks, size_t dyn, hipStream_t s, const double* abuf, uint32_t amask,
                          double* sink, int steps)
{
  switch (ns) {
    case 1: launch_step_shape<F, 1>(blocks, dyn, s, abuf, amask, sink, steps); break;
    case 2: launch_step_shape<F, 2>(blocks, dyn, s, abuf, amask, sink, steps); break;
    case 3: launch_step_shape<F, 3>(blocks, dyn, s, abuf, amask, sink, steps); break;
    case 4: launch_step_shape<F, 4>(blocks, dyn, s, abuf, amask, sink, steps); break;
    case 5: launch_step_shape<F, 5>(blocks, dyn, s, abuf, amask, sink, steps); break;
    case 6: launch_step_shape<F, 6>(blocks, dyn, s, abuf, amask, sink, steps); break;
    case 7: launch_step_shape<F, 7>(blocks, dyn, s, abuf, amask, sink, steps); break;
    default: launch_step_shape<F, 8>(blocks, dyn, s, abuf, amask, sink, steps); break;
  }
}

}  // namespace
}  // namespace sht

sht_status sht_diag_mfma_f64_rate(void* stream, double* tflops)
{
  return sht::guard([&] {
    SHT_REQUIRE(tflops, SHT_ERR_INVALID, "null argument");
    hipStream_t s = sht::as_stream(stream);
    int dev = 0, cus = 0;
    SHT_HIP(hipGetDevice(&dev));
    SHT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = 2 * cus;  // two waves per SIMD
    const int iters = 2000;
    double* sink = nullptr;
    SHT_HIP(hipMalloc(&sink, static_cast<size_t>(blocks) * 256 * sizeof(double)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) {
      hipLaunchKernelGGL(sht::k_mfma_loop, dim3(blocks), dim3(256), 0, s, sink, iters / 8, 1.0);  // warm-up
      err = hipEventRecord(e0, s);
    }
    if (err == hipSuccess) {
      hipLaunchKernelGGL(sht::k_mfma_loop, dim3(blocks), dim3(256), 0, s, sink, iters, 1.0);
      err = hipEventRecord(e1, s);
    }
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    SHT_HIP(err);
    const double flop = 2.0 * 16 * 16 * 4 * 16.0 * iters * blocks * 4;
    *tflops = flop / (static_cast<double>(ms) * 1e-3) / 1e12;
  });
}

sht_status sht_diag_step_rate(sht_precision prec, int slots, void* stream, double* tflops)
{
  return sht::guard([&] {
    SHT_REQUIRE(tflops, SHT_ERR_INVALID, "null argument");
    SHT_REQUIRE(prec == SHT_PREC_FP64 || prec == SHT_PREC_FP32, SHT_ERR_INVALID, "precision must be FP64 or FP32");
    SHT_REQUIRE(slots >= 1 && slots <= 8, SHT_ERR_INVALID, "slots must be 1..8");
    hipStream_t s = sht::as_stream(stream);
    int dev = 0, cus = 0;
    SHT_HIP(hipGetDevice(&dev));
    SHT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = 2 * cus;
    // 48 KB of dynamic LDS beside the 8-16 KB stage ring: two workgroups per CU
    const size_t dyn = 48 * 1024;
    const int steps = 8000;
    const size_t abytes = size_t{8} << 20;  // cache-resident A buffer
    const uint32_t amask = static_cast<uint32_t>(abytes / sizeof(double) - 1) & ~1u;
    double* sink = nullptr;
    double* abuf = nullptr;
    SHT_HIP(hipMalloc(&sink, static_cast<size_t>(blocks) * 256 * sizeof(double)));
    hipError_t err = hipMalloc(&abuf, abytes);
    if (err == hipSuccess) err = hipMemsetAsync(abuf, 0, abytes, s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    auto launch = [&](int n) {
      if (prec == SHT_PREC_FP64) sht::launch_step_shape_ns<0>(slots, blocks, dyn, s, abuf, amask, sink, n);
      else sht::launch_step_shape_ns<1>(slots, blocks, dyn, s, abuf, amask, sink, n);
    };
    if (err == hipSuccess) {
      launch(64);  // warm-up
      err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipEventRecord(e0, s);
    if (err == hipSuccess) {
      launch(steps);
      err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(abuf);
    (void)hipFree(sink);
    SHT_HIP(err);
    const double mfmas = 2.0 * slots * steps * static_cast<double>(blocks) * 4;
    *tflops = mfmas * 2048.0 / (static_cast<double>(ms) * 1e-3) / 1e12;
  });
}
