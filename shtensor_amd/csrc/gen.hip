// Data-format kernels: synthetic fill, block norms, zero fill.
// All are HBM-bound streaming kernels: one workgroup per block, lanes stride
// over the block's contiguous elements (coalesced 8-byte accesses).

#include "internal.hpp"

namespace sht {

namespace {

__device__ inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ inline void block_coords(const TensorDesc& d, int64_t id, int* c)
{
  for (int m = d.N - 1; m >= 0; --m) {
    c[m] = static_cast<int>(id / d.bstride[m]);
    id -= static_cast<int64_t>(c[m]) * d.bstride[m];
  }
}

__device__ inline int64_t block_volume(const TensorDesc& d, const int* c)
{
  int64_t v = 1;
  for (int m = 0; m < d.N; ++m) v *= d.bsizes[d.boff[m] + c[m]];
  return v;
}

template <class T>
__global__ __launch_bounds__(256) void k_gen_fill(TensorDesc d, T* data, uint64_t key, const double* factors)
{
  for (int64_t p = blockIdx.x; p < d.nnz; p += gridDim.x) {
    const int64_t id = d.indices[p];
    int c[kMaxModes];
    block_coords(d, id, c);
    const int64_t vol = block_volume(d, c);
    const int dist = d.N >= 2 ? abs(c[0] - c[1]) : 0;
    const double f = factors[dist];
    const uint64_t bkey = mix64(key ^ mix64(static_cast<uint64_t>(id) ^ 0x5851F42D4C957F2DULL));
    T* out = data + d.offsets[p];
    for (int64_t x = threadIdx.x; x < vol; x += blockDim.x) {
      const uint64_t h = mix64(bkey + static_cast<uint64_t>(x));
      const double u = static_cast<double>(h >> 11) * 0x1.0p-53;
      out[x] = static_cast<T>((2.0 * u - 1.0) * f);
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_block_norms(TensorDesc d, const T* data, int norm, double* out)
{
  __shared__ double red[256];
  for (int64_t p = blockIdx.x; p < d.nnz; p += gridDim.x) {
    int64_t id = d.indices[p];
    if (id < 0) id = -id - 1;
    int c[kMaxModes];
    block_coords(d, id, c);
    const int64_t vol = block_volume(d, c);
    const T* b = data + d.offsets[p];
    double acc = 0.0;
    for (int64_t x = threadIdx.x; x < vol; x += blockDim.x) {
      const double v = static_cast<double>(b[x]);
      acc = norm == 0 ? acc + v * v : fmax(acc, fabs(v));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        red[threadIdx.x] = norm == 0 ? red[threadIdx.x] + red[threadIdx.x + s]
                                     : fmax(red[threadIdx.x], red[threadIdx.x + s]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) out[p] = norm == 0 ? sqrt(red[0]) : red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_zero(uint4* p, size_t n16, unsigned char* tail, size_t ntail)
{
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
    p[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0)
    for (size_t i = threadIdx.x; i < ntail; i += blockDim.x) tail[i] = 0;
}

// Block gather (BaseTensor::compress): record r copies n[r] 4-byte words
// from word offset src[r] to word offset dst[r]; one workgroup per record.
__global__ __launch_bounds__(256) void k_gather_blocks(const uint32_t* in, uint32_t* out, const int64_t* recs,
                                                       int64_t nrec)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const int64_t src = recs[3 * r], dst = recs[3 * r + 1], n = recs[3 * r + 2];
    for (int64_t x = threadIdx.x; x < n; x += blockDim.x) out[dst + x] = in[src + x];
  }
}

// Element-wise copy with conversion between tensors of one block pattern
// (FP64 -> FP32 rounds to nearest even, FP32 -> FP64 is exact): the packed
// layouts are identical, so the copy is one flat stream.
template <class TI, class TO>
__global__ __launch_bounds__(256) void k_convert(const TI* in, TO* out, int64_t n)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = static_cast<TO>(in[i]);
}

int grid_for(int64_t items, int64_t cap = 65536)
{
  return static_cast<int>(items < 1 ? 1 : (items < cap ? items : cap));
}

}  // namespace

void launch_gen_fill(const TensorDesc& d, void* data, bool f64, uint64_t key, const double* d_factors, int,
                     hipStream_t s)
{
  const int g = grid_for(d.nnz);
  if (f64)
    hipLaunchKernelGGL(k_gen_fill<double>, dim3(g), dim3(256), 0, s, d, static_cast<double*>(data), key, d_factors);
  else
    hipLaunchKernelGGL(k_gen_fill<float>, dim3(g), dim3(256), 0, s, d, static_cast<float*>(data), key, d_factors);
  SHT_HIP(hipGetLastError());
}

void launch_block_norms(const TensorDesc& d, const void* data, bool f64, int norm, double* out, hipStream_t s)
{
  const int g = grid_for(d.nnz);
  if (f64)
    hipLaunchKernelGGL(k_block_norms<double>, dim3(g), dim3(256), 0, s, d, static_cast<const double*>(data), norm,
                       out);
  else
    hipLaunchKernelGGL(k_block_norms<float>, dim3(g), dim3(256), 0, s, d, static_cast<const float*>(data), norm,
                       out);
  SHT_HIP(hipGetLastError());
}

void launch_gather_blocks(const void* in, void* out, const int64_t* d_recs, int64_t nrec, hipStream_t s)
{
  if (nrec <= 0) return;
  hipLaunchKernelGGL(k_gather_blocks, dim3(grid_for(nrec)), dim3(256), 0, s, static_cast<const uint32_t*>(in),
                     static_cast<uint32_t*>(out), d_recs, nrec);
  SHT_HIP(hipGetLastError());
}

void launch_convert(const void* in, bool in_f64, void* out, bool out_f64, int64_t n, hipStream_t s)
{
  if (n <= 0) return;
  const int g = grid_for((n + 255) / 256, 16384);
  if (in_f64 && !out_f64)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(256), 0, s, static_cast<const double*>(in),
                       static_cast<float*>(out), n);
  else if (!in_f64 && out_f64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(256), 0, s, static_cast<const float*>(in),
                       static_cast<double*>(out), n);
  else if (in_f64)
    hipLaunchKernelGGL((k_convert<double, double>), dim3(g), dim3(256), 0, s, static_cast<const double*>(in),
                       static_cast<double*>(out), n);
  else
    hipLaunchKernelGGL((k_convert<float, float>), dim3(g), dim3(256), 0, s, static_cast<const float*>(in),
                       static_cast<float*>(out), n);
  SHT_HIP(hipGetLastError());
}

void launch_zero(void* p, size_t bytes, hipStream_t s)
{
  // 256-byte aligned arena pointers: 16-byte stores cover everything but a tail
  const size_t n16 = bytes / 16;
  const size_t ntail = bytes - n16 * 16;
  const int g = grid_for(static_cast<int64_t>((n16 + 255) / 256), 8192);
  hipLaunchKernelGGL(k_zero, dim3(g), dim3(256), 0, s, static_cast<uint4*>(p), n16,
                     static_cast<unsigned char*>(p) + n16 * 16, ntail);
  SHT_HIP(hipGetLastError());
}

}  // namespace sht
