// Kernel::call — packed same-shape block products, generic (any index map).
//
// One lane per C element of one op; the element is the fma chain of the
// reference JIT (src/Kernel/KernelImpl.cpp:594-710): c = beta ? beta*C : 0,
// then c = fma(alpha*A(i,k), B(k,j), c) for k ascending, with the element
// offsets taken from the scatter tables (ContractInfo.h:38-66).  Bitwise equal
// to the XMM kernel for any (expr, sizes).  Shape-specialised variants are
// generated at run time (jit.cpp); this one serves every shape.

#include "internal.hpp"

namespace sht {

namespace {

template <class T>
__device__ inline T fused(T a, T b, T c);
template <>
__device__ inline double fused<double>(double a, double b, double c) { return __fma_rn(a, b, c); }
template <>
__device__ inline float fused<float>(float a, float b, float c) { return __fmaf_rn(a, b, c); }

template <class T>
__global__ __launch_bounds__(256) void k_blockprod(const T* __restrict__ A, const T* __restrict__ B, T* C,
                                                   int64_t nb_ops, int M, int N, int K,
                                                   const int* __restrict__ offs, T alpha, T beta, int az, int bz)
{
  const int* ar = offs;
  const int* ac = ar + M;
  const int* br = ac + K;
  const int* bc = br + K;
  const int* cr = bc + N;
  const int* cc = cr + M;
  const int64_t per = static_cast<int64_t>(M) * N;
  const int64_t total = nb_ops * per;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t op = g / per;
    const int64_t r = g - op * per;
    const int j = static_cast<int>(r / M);
    const int i = static_cast<int>(r - static_cast<int64_t>(j) * M);
    const T* a = A + op * static_cast<int64_t>(M) * K;
    const T* b = B + op * static_cast<int64_t>(K) * N;
    T* c = C + op * per;
    const int oc = cr[i] + cc[j];
    T acc = bz ? T(0) : c[oc] * beta;
    if (!az) {
      const int ai = ar[i];
      const int bj = bc[j];
      for (int k = 0; k < K; ++k) acc = fused<T>(a[ai + ac[k]] * alpha, b[br[k] + bj], acc);
    }
    c[oc] = acc;
  }
}

}  // namespace

void launch_blockprod(const BlockProdArgs& p, bool f64, hipStream_t s)
{
  const int64_t total = p.nb_ops * static_cast<int64_t>(p.M) * p.N;
  if (total <= 0) return;
  const int64_t blocks = (total + 255) / 256;
  const int g = static_cast<int>(blocks < 262144 ? blocks : 262144);
  if (f64) {
    hipLaunchKernelGGL(k_blockprod<double>, dim3(g), dim3(256), 0, s, static_cast<const double*>(p.a),
                       static_cast<const double*>(p.b), static_cast<double*>(p.c), p.nb_ops, p.M, p.N, p.K, p.offs,
                       p.alpha, p.beta, p.alpha_zero, p.beta_zero);
  } else {
    hipLaunchKernelGGL(k_blockprod<float>, dim3(g), dim3(256), 0, s, static_cast<const float*>(p.a),
                       static_cast<const float*>(p.b), static_cast<float*>(p.c), p.nb_ops, p.M, p.N, p.K, p.offs,
                       static_cast<float>(p.alpha), static_cast<float>(p.beta), p.alpha_zero, p.beta_zero);
  }
  SHT_HIP(hipGetLastError());
}

}  // namespace sht
