// Internal object layouts shared by the host translation units and the
// device launch wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "arena.hpp"
#include "common.hpp"

namespace sht {

constexpr int kMaxModes = 8;
// Tensor data buffers are allocated with this many bytes of slack after the
// last element: the row-chunk kernels may read (never write) up to 64 rows
// past the end of a block column.
constexpr size_t kReadSlack = 1024;

// Block-level view of a tensor handed to kernels by value.
struct TensorDesc {
  int N;
  int nblk[kMaxModes];
  long long bstride[kMaxModes];  // rolled-index strides, mode 0 fastest
  int boff[kMaxModes];           // start of mode d inside bsizes
  const int* bsizes;             // device: concatenated block sizes
  const int64_t* indices;        // device: sorted block ids
  const int64_t* offsets;        // device: element offsets
  int64_t nnz;
};

}  // namespace sht

struct sht_ctx {
  int device = 0;
  bool host = false;  // host-backed arena, no device use
  void* d_base = nullptr;
  size_t bytes = 0;
  sht::Arena arena;
};

struct sht_tensor {
  sht_ctx* ctx = nullptr;
  sht_dtype dtype = SHT_FLOAT64;
  int N = 0;
  std::vector<int> nblk;
  std::vector<std::vector<int>> bs;
  std::vector<int64_t> bstride;
  std::vector<int64_t> indices, offsets, rows;
  int64_t nelem = 0;
  int64_t nblocks_total = 0;
  bool reserved = false;
  void* d_data = nullptr;        // arena
  int64_t* d_indices = nullptr;  // arena
  int64_t* d_offsets = nullptr;  // arena
  int* d_bsizes = nullptr;       // arena

  size_t elem_size() const { return dtype == SHT_FLOAT64 ? 8 : 4; }
  std::vector<int> coords(int64_t id) const;
  std::vector<int> dims(int64_t id) const;
  int64_t volume(int64_t id) const;
  sht::TensorDesc desc() const;
  void release_storage();
};

namespace sht {

// reserve from sorted unique ids (validated by the caller)
void tensor_reserve_sorted(sht_tensor* t, std::vector<int64_t>&& ids);

// ---- device launch wrappers (defined in .hip files) -------------------------

// gen.hip
void launch_gen_fill(const TensorDesc& d, void* data, bool f64, uint64_t key, const double* d_factors,
                     int nfactors, hipStream_t s);
void launch_block_norms(const TensorDesc& d, const void* data, bool f64, int norm, double* out, hipStream_t s);
void launch_zero(void* p, size_t bytes, hipStream_t s);
// element-wise copy of n elements between FP64/FP32 buffers
void launch_convert(const void* in, bool in_f64, void* out, bool out_f64, int64_t n, hipStream_t s);
// d_recs: nrec triples (source word offset, destination word offset, words)
void launch_gather_blocks(const void* in, void* out, const int64_t* d_recs, int64_t nrec, hipStream_t s);

// blockprod.hip — Kernel::call generic path
struct BlockProdArgs {
  const void* a;
  const void* b;
  void* c;
  int64_t nb_ops;
  int M, N, K;
  const int* offs;  // device: offA_r[M] offA_c[K] offB_r[K] offB_c[N] offC_r[M] offC_c[N]
  double alpha, beta;
  int alpha_zero, beta_zero;
};
void launch_blockprod(const BlockProdArgs& a, bool f64, hipStream_t s);

}  // namespace sht
