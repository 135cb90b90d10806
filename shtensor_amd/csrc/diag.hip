// Measurement helper: the FP64 MFMA rate a plain register-resident loop
// sustains on this device (no counterpart in the reference; used to put the
// contraction kernels' matrix-core utilisation next to an achievable rate).

#include "common.hpp"

namespace sht {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

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
