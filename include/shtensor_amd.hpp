/*
 * shtensor_amd.hpp — the reference's C++ interface (namespace Shtensor:
 * Kernel<T>, Context, Tensor<T,N>) as a header-only layer over the C ABI in
 * sht.h, so a single-rank shtensor host program compiles against it
 * unchanged apart from the Context constructor (no MPI communicator).
 *
 * Reference interfaces mirrored:
 *   Kernel<T>       src/Kernel/Kernel.h:62-93 (ctor from any size containers,
 *                   call(T*, T*, T*, int64_t) on host data, get_info)
 *   KernelType      src/Kernel/KernelDefinitions.h:11-15
 *   BlockNorm       src/Basics/BlockSpan.h:87-91
 *   VArray<N>       src/Basics/Definitions.h:52
 *   Context         src/Shmem/Context.cpp:9-48 (one rank, one GPU)
 *   Tensor<T,N>     src/Tensor/Tensor.h:19-296, BaseTensor.h:27-53
 *                   (reserve_all, reserve(const std::vector<int>*),
 *                   get_local_block, get_sparse_info, filter, compress,
 *                   get_nb_nzblocks_local/global, get_occupation)
 * Errors are thrown as std::runtime_error carrying the library's message
 * (the reference's text where it has one).  Block data lives in device
 * memory (the context's HBM arena): get_local_block returns a device
 * pointer; upload/download move whole tensors.
 */
#ifndef SHTENSOR_AMD_HPP
#define SHTENSOR_AMD_HPP

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "sht.h"

namespace Shtensor
{

enum class KernelType { LAPACK = SHT_KERNEL_LAPACK, XMM = SHT_KERNEL_XMM };
enum class BlockNorm { FROBENIUS = SHT_NORM_FROBENIUS, MAXABS = SHT_NORM_MAXABS };
/* Block arithmetic of a contraction (no reference counterpart). */
enum class Precision {
  FP64 = SHT_PREC_FP64,
  FP64_EXACT = SHT_PREC_FP64_EXACT,
  FP32 = SHT_PREC_FP32,
  BF16X3 = SHT_PREC_BF16X3
};

template <int N>
using VArray = std::array<std::vector<int>, N>;

inline void check(sht_status s)
{
  if (s != SHT_OK) throw std::runtime_error(sht_last_error());
}

template <class T>
constexpr sht_dtype dtype_of()
{
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "float or double");
  return std::is_same<T, double>::value ? SHT_FLOAT64 : SHT_FLOAT32;
}

/* Kernel<T>: nb_ops packed same-shape block products,
 * C_op = beta*C_op + contract(alpha*A_op, B_op). */
template <class T>
class Kernel
{
 public:
  template <class ArrayIn1, class ArrayIn2, class ArrayOut>
  Kernel(const std::string& expr, const ArrayIn1& sizes_in1, const ArrayIn2& sizes_in2, const ArrayOut& sizes_out,
         T alpha, T beta, KernelType kernel_type)
  {
    const std::vector<int> a(sizes_in1.begin(), sizes_in1.end());
    const std::vector<int> b(sizes_in2.begin(), sizes_in2.end());
    const std::vector<int> c(sizes_out.begin(), sizes_out.end());
    check(sht_kernel_create(expr.c_str(), a.data(), static_cast<int>(a.size()), b.data(), static_cast<int>(b.size()),
                            c.data(), static_cast<int>(c.size()), static_cast<double>(alpha),
                            static_cast<double>(beta), dtype_of<T>(), static_cast<sht_kernel_type>(kernel_type),
                            &m_kernel));
  }
  Kernel(const Kernel&) = delete;
  Kernel& operator=(const Kernel&) = delete;
  Kernel(Kernel&& o) noexcept : m_kernel(std::exchange(o.m_kernel, nullptr)) {}
  ~Kernel() { sht_kernel_destroy(m_kernel); }

  /* The reference's call: host or device pointers; the products are done
   * when it returns. */
  void call(T* a, T* b, T* c, int64_t nb_ops)
  {
    check(sht_kernel_call(m_kernel, a, b, c, nb_ops, nullptr));
    check(sht_stream_synchronize(nullptr));
  }
  /* Asynchronous form for device pointers (a hipStream_t as void*). */
  void call_async(const T* a, const T* b, T* c, int64_t nb_ops, void* stream)
  {
    check(sht_kernel_call(m_kernel, a, b, c, nb_ops, stream));
  }
  std::string get_info() const
  {
    std::string s(1 << 16, '\0');
    check(sht_kernel_info(m_kernel, &s[0], s.size()));
    return std::string(s.c_str());
  }
  sht_kernel* handle() const { return m_kernel; }

 private:
  sht_kernel* m_kernel = nullptr;
};

/* One GPU and its HBM arena (the reference's MemoryPool window). */
class Context
{
 public:
  explicit Context(size_t pool_bytes, int device = 0) { check(sht_ctx_create(pool_bytes, device, &m_ctx)); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  ~Context() { sht_ctx_destroy(m_ctx); }
  size_t get_free_mem() const
  {
    size_t total = 0, free_bytes = 0, largest = 0;
    check(sht_ctx_arena_info(m_ctx, &total, &free_bytes, &largest));
    return free_bytes;
  }
  sht_ctx* handle() const { return m_ctx; }

 private:
  sht_ctx* m_ctx = nullptr;
};

/* A block of a tensor: device data and its dimensions (BlockSpan<T,N>). */
template <class T, int N>
struct BlockView {
  T* data;                   /* device pointer */
  std::array<int, N> sizes;  /* block dimensions */
  int64_t block_index;       /* rolled block id */
  int dim(int i) const { return sizes[i]; }
  int64_t size() const
  {
    int64_t v = 1;
    for (int x : sizes) v *= x;
    return v;
  }
};

template <class T, int N>
class Tensor
{
  static_assert(N >= 2, "Tensor has to have dimension >= 2");

 public:
  struct SparseInfo {
    const int64_t* rows;
    const int64_t* indices;
    const int64_t* offsets;
    int64_t nnz;
  };

  Tensor(const Context& ctx, const VArray<N>& block_sizes) : m_block_sizes(block_sizes)
  {
    const int* ptrs[N];
    int nblk[N];
    for (int i = 0; i < N; ++i) {
      ptrs[i] = m_block_sizes[i].data();
      nblk[i] = static_cast<int>(m_block_sizes[i].size());
    }
    check(sht_tensor_create(ctx.handle(), dtype_of<T>(), N, ptrs, nblk, &m_tensor));
  }
  Tensor(const Tensor&) = delete;
  Tensor& operator=(const Tensor&) = delete;
  Tensor(Tensor&& o) noexcept : m_block_sizes(std::move(o.m_block_sizes)), m_tensor(std::exchange(o.m_tensor, nullptr))
  {
  }
  ~Tensor() { sht_tensor_destroy(m_tensor); }

  void reserve_all() { check(sht_tensor_reserve_all(m_tensor)); }

  /* BaseTensor::reserve: N vectors of block coordinates (one per mode). */
  void reserve(const std::vector<int>* idx_per_mode)
  {
    const size_t nb = idx_per_mode[0].size();
    const int* ptrs[N];
    for (int i = 0; i < N; ++i) {
      if (idx_per_mode[i].size() != nb)
        throw std::runtime_error("reserve: Indices do not have the same size in each dimension");
      ptrs[i] = idx_per_mode[i].data();
    }
    check(sht_tensor_reserve(m_tensor, ptrs, static_cast<int64_t>(nb)));
  }
  void reserve(const VArray<N>& idx) { reserve(idx.data()); }

  SparseInfo get_sparse_info() const
  {
    SparseInfo s{};
    check(sht_tensor_sparse_info(m_tensor, &s.indices, &s.offsets, &s.rows, &s.nnz));
    return s;
  }

  std::array<int, N> get_indices(int64_t block_id) const
  {
    std::array<int, N> c{};
    for (int i = 0; i < N; ++i) {
      const int64_t n = static_cast<int64_t>(m_block_sizes[i].size());
      c[i] = static_cast<int>(block_id % n);
      block_id /= n;
    }
    return c;
  }

  /* Tensor::get_local_block: the block at position pos of the sparse list. */
  BlockView<T, N> get_local_block(int64_t pos) const
  {
    const SparseInfo s = get_sparse_info();
    if (pos < 0 || pos >= s.nnz) throw std::runtime_error("get_local_block: position out of range");
    BlockView<T, N> b{};
    b.block_index = s.indices[pos];
    const std::array<int, N> c = get_indices(b.block_index);
    for (int i = 0; i < N; ++i) b.sizes[i] = m_block_sizes[i][c[i]];
    b.data = data() + s.offsets[pos];
    return b;
  }

  int64_t get_nb_nzblocks_local() const
  {
    int64_t l = 0, g = 0;
    check(sht_tensor_nb_nzblocks(m_tensor, &l, &g));
    return l;
  }
  int64_t get_nb_nzblocks_global() const
  {
    int64_t l = 0, g = 0;
    check(sht_tensor_nb_nzblocks(m_tensor, &l, &g));
    return g;
  }
  double get_occupation() const
  {
    double o = 0.0;
    check(sht_tensor_occupation(m_tensor, &o));
    return o;
  }

  /* Tensor::filter: blocks with norm < eps are dropped (compress) or their
   * ids negated (-id-1) until compress(). */
  void filter(T eps, BlockNorm norm, bool compress = true)
  {
    check(sht_tensor_filter(m_tensor, static_cast<double>(eps), static_cast<sht_block_norm>(norm), compress ? 1 : 0));
  }
  void compress() { check(sht_tensor_compress(m_tensor)); }

  T* data() const
  {
    void* p = nullptr;
    int64_t n = 0;
    check(sht_tensor_data(m_tensor, &p, &n));
    return static_cast<T*>(p);
  }
  int64_t nelem() const
  {
    void* p = nullptr;
    int64_t n = 0;
    check(sht_tensor_data(m_tensor, &p, &n));
    return n;
  }
  void upload(const std::vector<T>& host) { check(sht_tensor_upload(m_tensor, host.data(), static_cast<int64_t>(host.size()))); }
  std::vector<T> download() const
  {
    std::vector<T> out(static_cast<size_t>(nelem()));
    check(sht_tensor_download(m_tensor, out.data(), static_cast<int64_t>(out.size())));
    return out;
  }
  const VArray<N>& get_block_sizes() const { return m_block_sizes; }
  sht_tensor* handle() const { return m_tensor; }

 private:
  VArray<N> m_block_sizes;
  sht_tensor* m_tensor = nullptr;
};

/* C = beta*C + alpha*contract(expr, A, B) on the GPU (the contraction driver
 * the reference declares but leaves empty, src/Tensor/Tensor.cpp).  An empty
 * C gets the product's block pattern; a reserved C keeps its own.  Done when
 * it returns. */
template <class TO, int NA, int NB, int NC>
void contract(Context& ctx, const std::string& expr, double alpha, const Tensor<TO, NA>& A,
              const Tensor<TO, NB>& B, double beta, Tensor<double, NC>& C, Precision prec = Precision::FP64,
              void* stream = nullptr)
{
  // operands are double for every form, or float for Precision::FP32 and
  // Precision::BF16X3 (the library rejects other combinations); C is double
  check(sht_contract(ctx.handle(), expr.c_str(), alpha, A.handle(), B.handle(), beta, C.handle(),
                     static_cast<sht_precision>(prec), stream));
}

/* No counterpart in the reference: dst takes src's block pattern and src's
 * values converted to U (sht_tensor_convert; double -> float rounds to
 * nearest even).  Done when it returns. */
template <class U, class T, int N>
void convert(const Tensor<T, N>& src, Tensor<U, N>& dst)
{
  check(sht_tensor_convert(src.handle(), dst.handle(), nullptr));
  check(sht_device_synchronize());
}

}  // namespace Shtensor

#endif
