// Context, arena, device helpers and the block-sparse tensor.
//
// Storage format of src/Tensor/BaseTensor.cpp:221-355: block ids are rolled
// with mode 0 fastest (Utils.h:80-94), kept sorted ascending; offsets are the
// exclusive prefix sum of block volumes in that order (int64: the reference's
// int accumulation at BaseTensor.cpp:160 overflows at large sizes); data is one
// packed, zero-filled region — here in the context's HBM arena.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "internal.hpp"

namespace sht {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

}  // namespace sht

using sht::Error;
using sht::guard;

// ---------------------------------------------------------------------------
// sht_tensor methods
// ---------------------------------------------------------------------------
std::vector<int> sht_tensor::coords(int64_t id) const
{
  std::vector<int> c(N);
  for (int d = N - 1; d >= 0; --d) {
    c[d] = static_cast<int>(id / bstride[d]);
    id %= bstride[d];
  }
  return c;
}

std::vector<int> sht_tensor::dims(int64_t id) const
{
  std::vector<int> c = coords(id);
  for (int d = 0; d < N; ++d) c[d] = bs[d][c[d]];
  return c;
}

int64_t sht_tensor::volume(int64_t id) const
{
  int64_t v = 1;
  for (int d = N - 1; d >= 0; --d) {
    v *= bs[d][id / bstride[d]];
    id %= bstride[d];
  }
  return v;
}

sht::TensorDesc sht_tensor::desc() const
{
  sht::TensorDesc d{};
  d.N = N;
  int off = 0;
  for (int i = 0; i < N; ++i) {
    d.nblk[i] = nblk[i];
    d.bstride[i] = bstride[i];
    d.boff[i] = off;
    off += nblk[i];
  }
  d.bsizes = d_bsizes;
  d.indices = d_indices;
  d.offsets = d_offsets;
  d.nnz = static_cast<int64_t>(indices.size());
  return d;
}

void sht_tensor::release_storage()
{
  if (d_data) ctx->arena.release(d_data);
  if (d_indices) ctx->arena.release(d_indices);
  if (d_offsets) ctx->arena.release(d_offsets);
  d_data = nullptr;
  d_indices = nullptr;
  d_offsets = nullptr;
  indices.clear();
  offsets.clear();
  nelem = 0;
  reserved = false;
}

namespace sht {

void tensor_reserve_sorted(sht_tensor* t, std::vector<int64_t>&& ids)
{
  t->release_storage();
  t->indices = std::move(ids);
  const int64_t nb = static_cast<int64_t>(t->indices.size());
  t->offsets.resize(nb);
  std::vector<int64_t> per_row(t->nblk[0], 0);
  int64_t off = 0;
  for (int64_t p = 0; p < nb; ++p) {
    t->offsets[p] = off;
    off += t->volume(t->indices[p]);
    per_row[t->indices[p] % t->nblk[0]]++;
  }
  // rows: per-mode-0 count prefix, as the reference computes it
  // (BaseTensor.cpp:243-312); it is not a CSR pointer for the mode-0-fastest sort
  t->rows.assign(t->nblk[0], 0);
  int64_t s = 0;
  for (int r = 0; r < t->nblk[0]; ++r) {
    t->rows[r] = s;
    s += per_row[r];
  }
  t->nelem = off;
  sht_ctx* c = t->ctx;
  t->d_data = c->arena.allocate(static_cast<size_t>(std::max<int64_t>(off, 1)) * t->elem_size() + kReadSlack);
  t->d_indices = static_cast<int64_t*>(c->arena.allocate(std::max<int64_t>(nb, 1) * sizeof(int64_t)));
  t->d_offsets = static_cast<int64_t*>(c->arena.allocate(std::max<int64_t>(nb, 1) * sizeof(int64_t)));
  if (nb) {
    SHT_HIP(hipMemcpy(t->d_indices, t->indices.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice));
    SHT_HIP(hipMemcpy(t->d_offsets, t->offsets.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  launch_zero(t->d_data, static_cast<size_t>(std::max<int64_t>(off, 1)) * t->elem_size(), nullptr);
  SHT_HIP(hipStreamSynchronize(nullptr));
  t->reserved = true;
}

}  // namespace sht

static void reserve_ids_checked(sht_tensor* t, std::vector<int64_t> ids)
{
  for (int64_t id : ids)
    SHT_REQUIRE(id >= 0 && id < t->nblocks_total, SHT_ERR_INVALID, "reserve: block index out of range");
  std::sort(ids.begin(), ids.end());
  SHT_REQUIRE(std::adjacent_find(ids.begin(), ids.end()) == ids.end(), SHT_ERR_INVALID,
              "reserve: duplicate block index");
  sht::tensor_reserve_sorted(t, std::move(ids));
}

extern "C" {

const char* sht_last_error(void) { return sht::g_last_error.c_str(); }

const char* sht_version(void) { return "shtensor_amd 0.1 (gfx950)"; }

// ---- context ----------------------------------------------------------------
sht_status sht_ctx_create(size_t arena_bytes, int device, sht_ctx** out)
{
  return guard([&] {
    SHT_REQUIRE(out, SHT_ERR_INVALID, "null output");
    auto c = std::make_unique<sht_ctx>();
    c->device = device;
    SHT_HIP(hipSetDevice(device));
    const size_t bytes = std::max<size_t>(arena_bytes, sht::Arena::kAlign);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(SHT_ERR_OOM, "Out of memory: arena of " + std::to_string(bytes) + " bytes");
    }
    c->d_base = p;
    c->bytes = bytes;
    c->arena.init(p, bytes);
    *out = c.release();
  });
}

sht_status sht_ctx_create_host(size_t arena_bytes, sht_ctx** out)
{
  return guard([&] {
    SHT_REQUIRE(out, SHT_ERR_INVALID, "null output");
    auto c = std::make_unique<sht_ctx>();
    c->host = true;
    const size_t bytes = std::max<size_t>(arena_bytes, sht::Arena::kAlign);
    void* p = std::aligned_alloc(sht::Arena::kAlign, (bytes + sht::Arena::kAlign - 1) / sht::Arena::kAlign *
                                                         sht::Arena::kAlign);
    SHT_REQUIRE(p, SHT_ERR_OOM, "Out of memory");
    c->d_base = p;
    c->bytes = bytes;
    c->arena.init(p, bytes, true);
    *out = c.release();
  });
}

sht_status sht_ctx_destroy(sht_ctx* ctx)
{
  return guard([&] {
    if (!ctx) return;
    if (ctx->host) std::free(ctx->d_base);
    else if (ctx->d_base) SHT_HIP(hipFree(ctx->d_base));
    delete ctx;
  });
}

sht_status sht_ctx_arena_info(const sht_ctx* ctx, size_t* total, size_t* free_bytes, size_t* largest)
{
  return guard([&] {
    SHT_REQUIRE(ctx, SHT_ERR_INVALID, "null context");
    if (total) *total = ctx->arena.total();
    if (free_bytes) *free_bytes = ctx->arena.free_bytes();
    if (largest) *largest = ctx->arena.largest_free();
  });
}

sht_status sht_arena_alloc(sht_ctx* ctx, size_t bytes, void** dptr)
{
  return guard([&] {
    SHT_REQUIRE(ctx && dptr, SHT_ERR_INVALID, "null argument");
    *dptr = ctx->arena.allocate(bytes);
  });
}

sht_status sht_arena_realloc(sht_ctx* ctx, void* dptr, size_t bytes, void** out)
{
  return guard([&] {
    SHT_REQUIRE(ctx && out, SHT_ERR_INVALID, "null argument");
    *out = ctx->arena.reallocate(dptr, bytes);
  });
}

sht_status sht_arena_free(sht_ctx* ctx, void* dptr)
{
  return guard([&] {
    SHT_REQUIRE(ctx, SHT_ERR_INVALID, "null context");
    ctx->arena.release(dptr);
  });
}

sht_status sht_arena_size(const sht_ctx* ctx, const void* dptr, size_t* bytes)
{
  return guard([&] {
    SHT_REQUIRE(ctx && bytes, SHT_ERR_INVALID, "null argument");
    *bytes = ctx->arena.size_of(dptr);
  });
}

// ---- device helpers -------------------------------------------------------------
sht_status sht_dev_alloc(size_t bytes, void** dptr)
{
  return guard([&] {
    SHT_REQUIRE(dptr, SHT_ERR_INVALID, "null output");
    SHT_HIP(hipMalloc(dptr, std::max<size_t>(bytes, 1)));
  });
}

sht_status sht_dev_free(void* dptr)
{
  return guard([&] {
    if (dptr) SHT_HIP(hipFree(dptr));
  });
}

sht_status sht_copy_h2d(void* dst, const void* src, size_t bytes)
{
  return guard([&] {
    if (bytes) SHT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}

sht_status sht_copy_d2h(void* dst, const void* src, size_t bytes)
{
  return guard([&] {
    if (bytes) SHT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}

sht_status sht_device_synchronize(void)
{
  return guard([&] { SHT_HIP(hipDeviceSynchronize()); });
}

sht_status sht_stream_create(void** stream)
{
  return guard([&] {
    hipStream_t s;
    SHT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
  });
}

sht_status sht_stream_destroy(void* stream)
{
  return guard([&] {
    if (stream) SHT_HIP(hipStreamDestroy(sht::as_stream(stream)));
  });
}

sht_status sht_stream_synchronize(void* stream)
{
  return guard([&] { SHT_HIP(hipStreamSynchronize(sht::as_stream(stream))); });
}

// ---- tensor -------------------------------------------------------------------------
sht_status sht_tensor_create(sht_ctx* ctx, sht_dtype dtype, int N, const int* const* block_sizes, const int* nblk,
                             sht_tensor** out)
{
  return guard([&] {
    SHT_REQUIRE(ctx && out && block_sizes && nblk, SHT_ERR_INVALID, "null argument");
    SHT_REQUIRE(!ctx->host, SHT_ERR_STATE, "tensors need a device context");
    // Tensor.h:23 requires N >= 2
    SHT_REQUIRE(N >= 2 && N <= sht::kMaxModes, SHT_ERR_INVALID, "Tensor has to have dimension >= 2 (and <= 8)");
    SHT_REQUIRE(dtype == SHT_FLOAT32 || dtype == SHT_FLOAT64, SHT_ERR_INVALID, "Invalid float type");
    auto t = std::make_unique<sht_tensor>();
    t->ctx = ctx;
    t->dtype = dtype;
    t->N = N;
    t->nblk.assign(nblk, nblk + N);
    t->bs.resize(N);
    t->bstride.assign(N, 1);
    std::vector<int> flat;
    for (int d = 0; d < N; ++d) {
      SHT_REQUIRE(nblk[d] >= 1, SHT_ERR_INVALID, "every mode needs at least one block");
      t->bs[d].assign(block_sizes[d], block_sizes[d] + nblk[d]);
      for (int b : t->bs[d]) SHT_REQUIRE(b >= 1, SHT_ERR_INVALID, "block sizes must be positive");
      flat.insert(flat.end(), t->bs[d].begin(), t->bs[d].end());
      if (d) t->bstride[d] = t->bstride[d - 1] * t->nblk[d - 1];
    }
    t->nblocks_total = t->bstride[N - 1] * t->nblk[N - 1];
    t->d_bsizes = static_cast<int*>(ctx->arena.allocate(flat.size() * sizeof(int)));
    SHT_HIP(hipMemcpy(t->d_bsizes, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice));
    *out = t.release();
  });
}

sht_status sht_tensor_destroy(sht_tensor* t)
{
  return guard([&] {
    if (!t) return;
    t->release_storage();
    if (t->d_bsizes) t->ctx->arena.release(t->d_bsizes);
    delete t;
  });
}

sht_status sht_tensor_reserve(sht_tensor* t, const int* const* idx, int64_t nb)
{
  return guard([&] {
    SHT_REQUIRE(t && (idx || nb == 0), SHT_ERR_INVALID, "null argument");
    std::vector<int64_t> ids(static_cast<size_t>(nb));
    for (int64_t b = 0; b < nb; ++b) {
      int64_t r = 0;
      for (int d = 0; d < t->N; ++d) {
        const int x = idx[d][b];
        SHT_REQUIRE(x >= 0 && x < t->nblk[d], SHT_ERR_INVALID, "reserve: block index out of range");
        r += x * t->bstride[d];
      }
      ids[b] = r;
    }
    reserve_ids_checked(t, std::move(ids));
  });
}

sht_status sht_tensor_reserve_ids(sht_tensor* t, const int64_t* ids, int64_t nb)
{
  return guard([&] {
    SHT_REQUIRE(t && (ids || nb == 0), SHT_ERR_INVALID, "null argument");
    reserve_ids_checked(t, std::vector<int64_t>(ids, ids + nb));
  });
}

sht_status sht_tensor_reserve_all(sht_tensor* t)
{
  return guard([&] {
    SHT_REQUIRE(t, SHT_ERR_INVALID, "null tensor");
    std::vector<int64_t> ids(static_cast<size_t>(t->nblocks_total));
    std::iota(ids.begin(), ids.end(), 0);
    sht::tensor_reserve_sorted(t, std::move(ids));
  });
}

sht_status sht_tensor_sparse_info(const sht_tensor* t, const int64_t** indices, const int64_t** offsets,
                                  const int64_t** rows, int64_t* nnz)
{
  return guard([&] {
    SHT_REQUIRE(t, SHT_ERR_INVALID, "null tensor");
    if (indices) *indices = t->indices.data();
    if (offsets) *offsets = t->offsets.data();
    if (rows) *rows = t->rows.data();
    if (nnz) *nnz = static_cast<int64_t>(t->indices.size());
  });
}

sht_status sht_tensor_data(const sht_tensor* t, void** dptr, int64_t* nelem)
{
  return guard([&] {
    SHT_REQUIRE(t, SHT_ERR_INVALID, "null tensor");
    if (dptr) *dptr = t->d_data;
    if (nelem) *nelem = t->nelem;
  });
}

sht_status sht_tensor_nb_nzblocks(const sht_tensor* t, int64_t* local, int64_t* global)
{
  return guard([&] {
    SHT_REQUIRE(t, SHT_ERR_INVALID, "null tensor");
    const int64_t n = static_cast<int64_t>(t->indices.size());
    if (local) *local = n;
    if (global) *global = n;  // one rank: local == global
  });
}

sht_status sht_tensor_occupation(const sht_tensor* t, double* occ)
{
  return guard([&] {
    SHT_REQUIRE(t && occ, SHT_ERR_INVALID, "null argument");
    *occ = static_cast<double>(t->indices.size()) / static_cast<double>(t->nblocks_total);
  });
}

sht_status sht_tensor_upload(sht_tensor* t, const void* host, int64_t nelem)
{
  return guard([&] {
    SHT_REQUIRE(t && t->reserved, SHT_ERR_STATE, "tensor not reserved");
    SHT_REQUIRE(nelem == t->nelem, SHT_ERR_INVALID, "element count mismatch");
    if (nelem) SHT_HIP(hipMemcpy(t->d_data, host, nelem * t->elem_size(), hipMemcpyHostToDevice));
  });
}

sht_status sht_tensor_download(const sht_tensor* t, void* host, int64_t nelem)
{
  return guard([&] {
    SHT_REQUIRE(t && t->reserved, SHT_ERR_STATE, "tensor not reserved");
    SHT_REQUIRE(nelem == t->nelem, SHT_ERR_INVALID, "element count mismatch");
    if (nelem) SHT_HIP(hipMemcpy(host, t->d_data, nelem * t->elem_size(), hipMemcpyDeviceToHost));
  });
}

// No counterpart in the reference (its Tensor<T,N> fixes T per tensor):
// dst takes src's block pattern and src's values converted to dst's float
// type, so FP64 data can be given the FP32 copy the FP32 and split-BF16
// contraction forms read.
sht_status sht_tensor_convert(const sht_tensor* src, sht_tensor* dst, void* stream)
{
  return guard([&] {
    SHT_REQUIRE(src && dst && src != dst, SHT_ERR_INVALID, "null or identical tensors");
    SHT_REQUIRE(src->reserved, SHT_ERR_STATE, "source tensor not reserved");
    SHT_REQUIRE(src->ctx == dst->ctx, SHT_ERR_INVALID, "tensors of another context");
    SHT_REQUIRE(src->bs == dst->bs, SHT_ERR_INVALID, "block sizes differ");
    SHT_REQUIRE(std::all_of(src->indices.begin(), src->indices.end(), [](int64_t id) { return id >= 0; }),
                SHT_ERR_STATE, "tensor holds filtered blocks: compress the tensor first");
    if (!dst->reserved || dst->indices != src->indices) {
      SHT_REQUIRE(!dst->reserved, SHT_ERR_STATE, "destination reserved with another block pattern");
      sht::tensor_reserve_sorted(dst, std::vector<int64_t>(src->indices));
    }
    sht::launch_convert(src->d_data, src->dtype == SHT_FLOAT64, dst->d_data, dst->dtype == SHT_FLOAT64, src->nelem,
                        sht::as_stream(stream));
  });
}

sht_status sht_tensor_block_norms(const sht_tensor* t, sht_block_norm norm, double* d_out, void* stream)
{
  return guard([&] {
    SHT_REQUIRE(t && t->reserved && d_out, SHT_ERR_STATE, "tensor not reserved");
    SHT_REQUIRE(norm == SHT_NORM_FROBENIUS || norm == SHT_NORM_MAXABS, SHT_ERR_INVALID, "unknown norm");
    if (!t->indices.empty())
      sht::launch_block_norms(t->desc(), t->d_data, t->dtype == SHT_FLOAT64, static_cast<int>(norm), d_out,
                              sht::as_stream(stream));
  });
}

// Tensor::filter (Tensor.h:336-363) + BaseTensor::compress (BaseTensor.cpp:253-354)
// with the fixes: FP64 norm accumulation, correct block sizes on compaction.
sht_status sht_tensor_filter(sht_tensor* t, double eps, sht_block_norm norm, int compress)
{
  return guard([&] {
    SHT_REQUIRE(t && t->reserved, SHT_ERR_STATE, "tensor not reserved");
    const int64_t nb = static_cast<int64_t>(t->indices.size());
    if (nb == 0) return;
    double* d_norms = static_cast<double*>(t->ctx->arena.allocate(nb * sizeof(double)));
    std::vector<double> norms(nb);
    try {
      sht::launch_block_norms(t->desc(), t->d_data, t->dtype == SHT_FLOAT64, static_cast<int>(norm), d_norms,
                              nullptr);
      SHT_HIP(hipMemcpy(norms.data(), d_norms, nb * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
      t->ctx->arena.release(d_norms);
      throw;
    }
    t->ctx->arena.release(d_norms);
    for (int64_t p = 0; p < nb; ++p)
      if (norms[p] < eps && t->indices[p] >= 0) t->indices[p] = -t->indices[p] - 1;
    SHT_HIP(hipMemcpy(t->d_indices, t->indices.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice));
    if (compress) {
      sht_status s = sht_tensor_compress(t);
      if (s != SHT_OK) throw Error(s, sht_last_error());
    }
  });
}

sht_status sht_tensor_compress(sht_tensor* t)
{
  return guard([&] {
    SHT_REQUIRE(t && t->reserved, SHT_ERR_STATE, "tensor not reserved");
    const int64_t nb = static_cast<int64_t>(t->indices.size());
    std::vector<int64_t> keep_ids, keep_src;
    for (int64_t p = 0; p < nb; ++p) {
      if (t->indices[p] >= 0) {
        keep_ids.push_back(t->indices[p]);
        keep_src.push_back(p);
      }
    }
    if (static_cast<int64_t>(keep_ids.size()) == nb) return;
    // compact into a fresh allocation: one gather launch over all kept blocks
    const std::vector<int64_t> old_off = t->offsets;
    std::vector<int64_t> old_ids = t->indices;
    for (auto& x : old_ids) if (x < 0) x = -x - 1;
    void* old_data = t->d_data;
    t->d_data = nullptr;
    // keep old data alive across the reserve
    sht_ctx* c = t->ctx;
    t->release_storage();
    t->indices = std::move(keep_ids);
    const size_t es = t->elem_size();
    std::vector<int64_t> ids = t->indices;
    int64_t* d_recs = nullptr;
    try {
      sht::tensor_reserve_sorted(t, std::move(ids));
      const int64_t w = static_cast<int64_t>(es / 4);  // 4-byte words per element
      std::vector<int64_t> recs;
      recs.reserve(3 * keep_src.size());
      for (size_t q = 0; q < keep_src.size(); ++q) {
        const int64_t p = keep_src[q];
        recs.push_back(old_off[p] * w);
        recs.push_back(t->offsets[q] * w);
        recs.push_back(t->volume(old_ids[p]) * w);
      }
      if (!recs.empty()) {
        d_recs = static_cast<int64_t*>(c->arena.allocate(recs.size() * sizeof(int64_t)));
        SHT_HIP(hipMemcpy(d_recs, recs.data(), recs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        sht::launch_gather_blocks(old_data, t->d_data, d_recs, static_cast<int64_t>(keep_src.size()), nullptr);
        SHT_HIP(hipStreamSynchronize(nullptr));
      }
    } catch (...) {
      if (d_recs) c->arena.release(d_recs);
      c->arena.release(old_data);
      throw;
    }
    if (d_recs) c->arena.release(d_recs);
    c->arena.release(old_data);
  });
}

// ---- synthetic workloads ------------------------------------------------------------
static uint64_t splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static uint64_t tensor_key(uint64_t seed, int tid) { return splitmix64(seed + 0x632BE59BD9B4E019ULL * (uint64_t)(tid + 1)); }

sht_status sht_gen_block_sizes(uint64_t seed, int axis, int length, const int* choices, int nchoices, int* out,
                               int cap, int* count)
{
  return guard([&] {
    SHT_REQUIRE(choices && nchoices > 0 && out && count && length >= 0, SHT_ERR_INVALID, "bad argument");
    const uint64_t key = splitmix64(seed ^ (0xA0761D6478BD642FULL * (uint64_t)(axis + 1)));
    int n = 0, filled = 0;
    for (uint64_t ctr = 0; filled < length; ++ctr) {
      SHT_REQUIRE(n < cap, SHT_ERR_INVALID, "block size buffer too small");
      const int b = choices[splitmix64(key + ctr) % (uint64_t)nchoices];
      const int take = std::min(b, length - filled);
      out[n++] = take;
      filled += take;
    }
    *count = n;
  });
}

sht_status sht_gen_reserve(sht_tensor* t, uint64_t seed, int tid, double p)
{
  return guard([&] {
    SHT_REQUIRE(t, SHT_ERR_INVALID, "null tensor");
    const uint64_t key = tensor_key(seed, tid);
    std::vector<int64_t> ids;
    ids.reserve(static_cast<size_t>(t->nblocks_total * std::min(1.0, p * 1.1 + 0.01)));
    for (int64_t id = 0; id < t->nblocks_total; ++id) {
      const uint64_t h = splitmix64(key ^ splitmix64(static_cast<uint64_t>(id)));
      if (static_cast<double>(h >> 11) * 0x1.0p-53 < p) ids.push_back(id);
    }
    sht::tensor_reserve_sorted(t, std::move(ids));
  });
}

sht_status sht_gen_fill(sht_tensor* t, uint64_t seed, int tid, double lambda, void* stream)
{
  return guard([&] {
    SHT_REQUIRE(t && t->reserved, SHT_ERR_STATE, "tensor not reserved");
    if (t->indices.empty()) return;
    // decay factors exp(-lambda*d) computed on the host (std::exp), so device
    // values equal a CPU generator's bit for bit
    int nf = 1;
    for (int d = 0; d < std::min(2, t->N); ++d) nf = std::max(nf, t->nblk[d]);
    std::vector<double> f(nf, 1.0);
    if (lambda != 0.0)
      for (int d = 0; d < nf; ++d) f[d] = std::exp(-lambda * static_cast<double>(d));
    double* d_f = static_cast<double*>(t->ctx->arena.allocate(nf * sizeof(double)));
    try {
      SHT_HIP(hipMemcpy(d_f, f.data(), nf * sizeof(double), hipMemcpyHostToDevice));
      sht::launch_gen_fill(t->desc(), t->d_data, t->dtype == SHT_FLOAT64, tensor_key(seed, tid), d_f, nf,
                           sht::as_stream(stream));
      SHT_HIP(hipStreamSynchronize(sht::as_stream(stream)));
    } catch (...) {
      t->ctx->arena.release(d_f);
      throw;
    }
    t->ctx->arena.release(d_f);
  });
}

}  // extern "C"
