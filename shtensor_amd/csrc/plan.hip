// Contraction driver: GPU task list + C pattern, plan, execution, graphs.
//
// The reference has no contraction driver (src/Tensor/Tensor.cpp is empty);
// its building blocks define the rule implemented here:
//   * block-level expression analysis = ContractInfo on the block counts
//     (src/Basics/ContractInfo.h:68-346);
//   * task = (A block, B block) with equal contracted block coordinates,
//     keyed by the contracted coordinates rolled in A's axis order;
//   * tasks sorted by (C block id, key); C pattern = sorted unique C ids
//     (the storage order of BaseTensor::reserve, BaseTensor.cpp:270-312);
//   * per C element: beta*C (not read when beta == +0), then the Kernel's
//     fma chain (KernelImpl.cpp:594-710) over the tasks in key order.
// Built on the device: key extraction, radix sort of B by key, binary-search
// ranges, scan + expansion of the pairs, radix sort of (C id, key), head-flag
// scan for the unique C ids, and per-C-block task ranges.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <numeric>
#include <unordered_set>

#include "plan.hpp"

namespace sht {

namespace {

struct KeySpec {
  int nk;
  int kax[kMaxModes];
  long long kmul[kMaxModes];
  int nf;
  int fax[kMaxModes];
  long long fmul[kMaxModes];
};

__device__ inline void coords_of(const TensorDesc& d, int64_t id, int* c)
{
  for (int m = d.N - 1; m >= 0; --m) {
    c[m] = static_cast<int>(id / d.bstride[m]);
    id -= static_cast<int64_t>(c[m]) * d.bstride[m];
  }
}

__global__ void k_keys(TensorDesc d, KeySpec ks, int64_t* kkey, int64_t* cpart)
{
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= d.nnz) return;
  int c[kMaxModes];
  coords_of(d, d.indices[p], c);
  int64_t k = 0, f = 0;
  for (int r = 0; r < ks.nk; ++r) k += static_cast<int64_t>(c[ks.kax[r]]) * ks.kmul[r];
  for (int r = 0; r < ks.nf; ++r) f += static_cast<int64_t>(c[ks.fax[r]]) * ks.fmul[r];
  kkey[p] = k;
  cpart[p] = f;
}

__global__ void k_iota(int64_t* x, int64_t n)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = i;
}

__device__ inline int64_t lower_bound_dev(const int64_t* a, int64_t n, int64_t v)
{
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_ranges(const int64_t* kA, int64_t nA, const int64_t* kBs, int64_t nB, int64_t* lo, int64_t* cnt)
{
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= nA) return;
  const int64_t l = lower_bound_dev(kBs, nB, kA[p]);
  const int64_t h = lower_bound_dev(kBs, nB, kA[p] + 1);
  lo[p] = l;
  cnt[p] = h - l;
}

__global__ void k_expand(const int64_t* lo, const int64_t* cnt, const int64_t* toff, int64_t nA, const int64_t* permB,
                         const int64_t* cA, const int64_t* cB, const int64_t* kA, int64_t nK, int64_t* ta,
                         int64_t* tb, uint64_t* key)
{
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= nA) return;
  const int64_t base = toff[p], n = cnt[p], l = lo[p];
  for (int64_t q = 0; q < n; ++q) {
    const int64_t b = permB[l + q];
    ta[base + q] = p;
    tb[base + q] = b;
    key[base + q] = static_cast<uint64_t>((cA[p] + cB[b]) * nK + kA[p]);
  }
}

__global__ void k_gather(const int64_t* order, int64_t n, const int64_t* ta, const int64_t* tb, const uint64_t* key,
                         int64_t nK, int64_t* ta_s, int64_t* tb_s, int64_t* cid_s, int64_t* kk_s)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t o = order[t];
  ta_s[t] = ta[o];
  tb_s[t] = tb[o];
  const uint64_t k = key[t];
  cid_s[t] = static_cast<int64_t>(k / static_cast<uint64_t>(nK));
  kk_s[t] = static_cast<int64_t>(k % static_cast<uint64_t>(nK));
}

__global__ void k_headflags(const int64_t* cid, int64_t n, int64_t* flag)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < n) flag[t] = (t == 0 || cid[t] != cid[t - 1]) ? 1 : 0;
}

__global__ void k_unique(const int64_t* cid, const int64_t* flag, int64_t* incl, int64_t n, int64_t* uids)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t pos = incl[t] - 1;
  if (flag[t]) uids[pos] = cid[t];
  incl[t] = pos;  // becomes the task's C position
}

__global__ void k_lookup(const int64_t* cid, int64_t n, const int64_t* cind, int64_t nc, int64_t* tc, char* keep)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t pos = lower_bound_dev(cind, nc, cid[t]);
  const bool found = pos < nc && cind[pos] == cid[t];
  tc[t] = found ? pos : -1;
  keep[t] = found ? 1 : 0;
}

__global__ void k_cstart(const int64_t* tc, int64_t nt, int64_t ncb, int64_t* cstart)
{
  const int64_t pc = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pc <= ncb) cstart[pc] = lower_bound_dev(tc, nt, pc);
}

__device__ inline int64_t prod_axes(const TensorDesc& d, const int* c, const int* ax, int n)
{
  int64_t v = 1;
  for (int r = 0; r < n; ++r) v *= d.bsizes[d.boff[ax[r]] + c[ax[r]]];
  return v;
}

struct AxisList {
  int n;
  int ax[kMaxModes];
};

__global__ void k_task_recs(const int64_t* ta, const int64_t* tb, int64_t nt, TensorDesc A, TensorDesc B,
                            AxisList acol, TaskRec* out)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int64_t a = ta[t], b = tb[t];
  int c[kMaxModes];
  coords_of(A, A.indices[a], c);
  TaskRec r;
  r.a_off = A.offsets[a];
  r.b_off = B.offsets[b];
  r.k = static_cast<int32_t>(prod_axes(A, c, acol.ax, acol.n));
  r.pad = 0;
  out[t] = r;
}

__global__ void k_cblk_recs(TensorDesc C, AxisList crow, AxisList ccol, const int64_t* cstart, const TaskRec* tasks,
                            CBlkRec* out, int64_t* work)
{
  const int64_t pc = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (pc >= C.nnz) return;
  int c[kMaxModes];
  coords_of(C, C.indices[pc], c);
  CBlkRec r;
  r.c_off = C.offsets[pc];
  r.m = static_cast<int32_t>(prod_axes(C, c, crow.ax, crow.n));
  r.n = static_cast<int32_t>(prod_axes(C, c, ccol.ax, ccol.n));
  r.t0 = static_cast<int32_t>(cstart[pc]);
  r.t1 = static_cast<int32_t>(cstart[pc + 1]);
  int64_t w = 0;
  for (int64_t t = r.t0; t < r.t1; ++t) w += tasks[t].k;
  out[pc] = r;
  work[pc] = w;
}

inline unsigned blocks_for(int64_t n, int tpb = 256) { return static_cast<unsigned>((std::max<int64_t>(n, 1) + tpb - 1) / tpb); }

double wall_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int bits_for(uint64_t maxval)
{
  int b = 1;
  while (b < 64 && (maxval >> b)) ++b;
  return b;
}

// Scoped arena allocation for temporaries.
class Scratch {
 public:
  explicit Scratch(Arena& a) : a_(a) {}
  ~Scratch()
  {
    for (void* p : ptrs_) a_.release(p);
  }
  template <class T>
  T* get(int64_t n)
  {
    void* p = a_.allocate(static_cast<size_t>(std::max<int64_t>(n, 1)) * sizeof(T));
    ptrs_.push_back(p);
    return static_cast<T*>(p);
  }
  void* bytes(size_t n)
  {
    void* p = a_.allocate(std::max<size_t>(n, 1));
    ptrs_.push_back(p);
    return p;
  }

 private:
  Arena& a_;
  std::vector<void*> ptrs_;
};

bool canonical(const std::vector<int>& lead, const std::vector<int>& trail)
{
  for (size_t r = 0; r < lead.size(); ++r)
    if (lead[r] != static_cast<int>(r)) return false;
  for (size_t r = 0; r < trail.size(); ++r)
    if (trail[r] != static_cast<int>(lead.size() + r)) return false;
  return true;
}

AxisList axis_list(const std::vector<int>& v)
{
  AxisList a{};
  a.n = static_cast<int>(v.size());
  for (size_t i = 0; i < v.size(); ++i) a.ax[i] = v[i];
  return a;
}

}  // namespace

}  // namespace sht

namespace sht {

namespace {

void check_tensors(const sht_tensor* A, const sht_tensor* B, const sht_tensor* C, sht_ctx* ctx, sht_precision prec)
{
  SHT_REQUIRE(A && B && C, SHT_ERR_INVALID, "null tensor");
  SHT_REQUIRE(A->reserved && B->reserved, SHT_ERR_STATE, "A and B must be reserved");
  SHT_REQUIRE(A->ctx == ctx && B->ctx == ctx && C->ctx == ctx, SHT_ERR_INVALID, "tensors of another context");
  SHT_REQUIRE(C->dtype == SHT_FLOAT64, SHT_ERR_UNSUPPORTED,
              "C must be FLOAT64 (the sum over contracted blocks runs in FP64 in every form)");
  SHT_REQUIRE(A->dtype == B->dtype, SHT_ERR_UNSUPPORTED, "A and B must have the same float type");
  SHT_REQUIRE(A->dtype == SHT_FLOAT64 || prec == SHT_PREC_FP32 || prec == SHT_PREC_BF16X3, SHT_ERR_UNSUPPORTED,
              "FLOAT32 operands need the FP32 or BF16X3 form (the FP64 forms read FLOAT64 tensors)");
  SHT_REQUIRE(A != C && B != C, SHT_ERR_INVALID, "C must not alias A or B");
  // Tensor::filter without compress leaves negated ids (-id-1) in the index
  // list (Tensor.h:336-363); such a tensor has no valid block coordinates
  // until BaseTensor::compress (BaseTensor.cpp:253-354) runs.
  for (const sht_tensor* t : {A, B, C})
    SHT_REQUIRE(!t->reserved || std::all_of(t->indices.begin(), t->indices.end(), [](int64_t id) { return id >= 0; }),
                SHT_ERR_STATE, "tensor holds filtered blocks: compress the tensor first");
}

// Build the task list and C pattern on the device.
void build_tasks(sht_plan* P)
{
  const sht_tensor* A = P->A;
  const sht_tensor* B = P->B;
  sht_tensor* C = P->C;
  const ContractMaps& m = P->maps;
  Arena& arena = P->ctx->arena;
  Scratch S(arena);
  hipStream_t s = nullptr;

  // keys: contracted coords rolled in A's con order; C id split into the part
  // carried by A's free coords and the part carried by B's free coords
  KeySpec ka{}, kb{};
  int64_t nK = 1;
  ka.nk = kb.nk = static_cast<int>(m.a_col.size());
  for (size_t r = 0; r < m.a_col.size(); ++r) {
    ka.kax[r] = m.a_col[r];
    kb.kax[r] = m.b_row[r];
    ka.kmul[r] = kb.kmul[r] = nK;
    nK *= A->nblk[m.a_col[r]];
  }
  ka.nf = static_cast<int>(m.a_row.size());
  for (size_t r = 0; r < m.a_row.size(); ++r) {
    ka.fax[r] = m.a_row[r];
    ka.fmul[r] = C->bstride[m.c_row[r]];
  }
  kb.nf = static_cast<int>(m.b_col.size());
  for (size_t r = 0; r < m.b_col.size(); ++r) {
    kb.fax[r] = m.b_col[r];
    kb.fmul[r] = C->bstride[m.c_col[r]];
  }
  const double comp = static_cast<double>(C->nblocks_total) * static_cast<double>(nK);
  SHT_REQUIRE(comp < 9.0e18, SHT_ERR_UNSUPPORTED, "block index space too large for 64-bit task keys");

  const TensorDesc dA = A->desc(), dB = B->desc();
  const int64_t nA = dA.nnz, nB = dB.nnz;
  P->ntasks = 0;
  if (nA == 0 || nB == 0) {
    if (C->indices.empty() && !C->reserved) tensor_reserve_sorted(C, {});
    else if (!C->reserved) tensor_reserve_sorted(C, {});
    P->ncb = static_cast<int64_t>(C->indices.size());
    return;
  }
  int64_t* kA = S.get<int64_t>(nA);
  int64_t* cA = S.get<int64_t>(nA);
  int64_t* kB = S.get<int64_t>(nB);
  int64_t* cB = S.get<int64_t>(nB);
  hipLaunchKernelGGL(k_keys, dim3(blocks_for(nA)), dim3(256), 0, s, dA, ka, kA, cA);
  hipLaunchKernelGGL(k_keys, dim3(blocks_for(nB)), dim3(256), 0, s, dB, kb, kB, cB);
  SHT_HIP(hipGetLastError());

  // sort B by key
  int64_t* kBs = S.get<int64_t>(nB);
  int64_t* iotaB = S.get<int64_t>(nB);
  int64_t* permB = S.get<int64_t>(nB);
  hipLaunchKernelGGL(k_iota, dim3(blocks_for(nB)), dim3(256), 0, s, iotaB, nB);
  const int kbits = bits_for(static_cast<uint64_t>(nK));
  size_t tmp = 0;
  SHT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kB, kBs, iotaB, permB, static_cast<int>(nB), 0, kbits, s));
  SHT_HIP(hipcub::DeviceRadixSort::SortPairs(S.bytes(tmp), tmp, kB, kBs, iotaB, permB, static_cast<int>(nB), 0,
                                             kbits, s));

  // matching B range of every A block, then task offsets
  int64_t* lo = S.get<int64_t>(nA);
  int64_t* cnt = S.get<int64_t>(nA);
  int64_t* toff = S.get<int64_t>(nA);
  hipLaunchKernelGGL(k_ranges, dim3(blocks_for(nA)), dim3(256), 0, s, kA, nA, kBs, nB, lo, cnt);
  tmp = 0;
  SHT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt, toff, static_cast<int>(nA), s));
  SHT_HIP(hipcub::DeviceScan::ExclusiveSum(S.bytes(tmp), tmp, cnt, toff, static_cast<int>(nA), s));
  int64_t last_off = 0, last_cnt = 0;
  SHT_HIP(hipMemcpy(&last_off, toff + nA - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
  SHT_HIP(hipMemcpy(&last_cnt, cnt + nA - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
  int64_t nt = last_off + last_cnt;
  SHT_REQUIRE(nt < (int64_t(1) << 31), SHT_ERR_UNSUPPORTED, "more than 2^31 block products in one contraction");

  const bool keep = C->reserved && !C->indices.empty();
  if (nt == 0) {
    if (!keep) tensor_reserve_sorted(C, {});
    P->ncb = static_cast<int64_t>(C->indices.size());
    P->d_cblk = nullptr;
    return;
  }
  int64_t* ta = S.get<int64_t>(nt);
  int64_t* tb = S.get<int64_t>(nt);
  uint64_t* key = S.get<uint64_t>(nt);
  hipLaunchKernelGGL(k_expand, dim3(blocks_for(nA)), dim3(256), 0, s, lo, cnt, toff, nA, permB, cA, cB, kA, nK, ta, tb,
                     key);
  SHT_HIP(hipGetLastError());

  // sort tasks by (C id, key)
  uint64_t* key_s = S.get<uint64_t>(nt);
  int64_t* iotaT = S.get<int64_t>(nt);
  int64_t* order = S.get<int64_t>(nt);
  hipLaunchKernelGGL(k_iota, dim3(blocks_for(nt)), dim3(256), 0, s, iotaT, nt);
  const int cbits = bits_for(static_cast<uint64_t>(C->nblocks_total) * static_cast<uint64_t>(nK));
  tmp = 0;
  SHT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key_s, iotaT, order, static_cast<int>(nt), 0, cbits, s));
  SHT_HIP(hipcub::DeviceRadixSort::SortPairs(S.bytes(tmp), tmp, key, key_s, iotaT, order, static_cast<int>(nt), 0,
                                             cbits, s));
  int64_t* ta_s = S.get<int64_t>(nt);
  int64_t* tb_s = S.get<int64_t>(nt);
  int64_t* cid_s = S.get<int64_t>(nt);
  int64_t* kk_s = S.get<int64_t>(nt);
  hipLaunchKernelGGL(k_gather, dim3(blocks_for(nt)), dim3(256), 0, s, order, nt, ta, tb, key_s, nK, ta_s, tb_s, cid_s,
                     kk_s);
  SHT_HIP(hipGetLastError());

  int64_t* tc = S.get<int64_t>(nt);
  if (!keep) {
    int64_t* flag = S.get<int64_t>(nt);
    hipLaunchKernelGGL(k_headflags, dim3(blocks_for(nt)), dim3(256), 0, s, cid_s, nt, flag);
    tmp = 0;
    SHT_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, flag, tc, static_cast<int>(nt), s));
    SHT_HIP(hipcub::DeviceScan::InclusiveSum(S.bytes(tmp), tmp, flag, tc, static_cast<int>(nt), s));
    int64_t ncb = 0;
    SHT_HIP(hipMemcpy(&ncb, tc + nt - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
    int64_t* uids = S.get<int64_t>(ncb);
    hipLaunchKernelGGL(k_unique, dim3(blocks_for(nt)), dim3(256), 0, s, cid_s, flag, tc, nt, uids);
    SHT_HIP(hipGetLastError());
    std::vector<int64_t> ids(static_cast<size_t>(ncb));
    SHT_HIP(hipMemcpy(ids.data(), uids, ncb * sizeof(int64_t), hipMemcpyDeviceToHost));
    tensor_reserve_sorted(C, std::move(ids));
  } else {
    char* kf = S.get<char>(nt);
    int64_t* tcl = S.get<int64_t>(nt);
    hipLaunchKernelGGL(k_lookup, dim3(blocks_for(nt)), dim3(256), 0, s, cid_s, nt, C->d_indices,
                       static_cast<int64_t>(C->indices.size()), tcl, kf);
    SHT_HIP(hipGetLastError());
    // compact the four task arrays by the keep flags
    int* nsel = S.get<int>(1);
    int64_t* bufs_in[4] = {ta_s, tb_s, tcl, kk_s};
    int64_t* bufs_out[4] = {S.get<int64_t>(nt), S.get<int64_t>(nt), tc, S.get<int64_t>(nt)};
    for (int q = 0; q < 4; ++q) {
      tmp = 0;
      SHT_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp, bufs_in[q], kf, bufs_out[q], nsel, static_cast<int>(nt), s));
      SHT_HIP(hipcub::DeviceSelect::Flagged(S.bytes(tmp), tmp, bufs_in[q], kf, bufs_out[q], nsel, static_cast<int>(nt),
                                            s));
    }
    int ns = 0;
    SHT_HIP(hipMemcpy(&ns, nsel, sizeof(int), hipMemcpyDeviceToHost));
    nt = ns;
    ta_s = bufs_out[0];
    tb_s = bufs_out[1];
    kk_s = bufs_out[3];
  }
  P->ntasks = nt;
  P->ncb = static_cast<int64_t>(C->indices.size());

  // keep the task list
  P->d_ta = static_cast<int64_t*>(arena.allocate(std::max<int64_t>(nt, 1) * sizeof(int64_t)));
  P->d_tb = static_cast<int64_t*>(arena.allocate(std::max<int64_t>(nt, 1) * sizeof(int64_t)));
  P->d_tc = static_cast<int64_t*>(arena.allocate(std::max<int64_t>(nt, 1) * sizeof(int64_t)));
  P->d_kk = static_cast<int64_t*>(arena.allocate(std::max<int64_t>(nt, 1) * sizeof(int64_t)));
  if (nt) {
    SHT_HIP(hipMemcpyAsync(P->d_ta, ta_s, nt * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    SHT_HIP(hipMemcpyAsync(P->d_tb, tb_s, nt * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    SHT_HIP(hipMemcpyAsync(P->d_tc, tc, nt * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    SHT_HIP(hipMemcpyAsync(P->d_kk, kk_s, nt * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
  }
  SHT_HIP(hipStreamSynchronize(s));
}

// The 32-row form of a 64-row MFMA tile class.
KClass half_rows(KClass k)
{
  switch (k) {
    case KClass::Mfma16: return KClass::Mfma16h;
    case KClass::Mfma32: return KClass::Mfma64S;
    case KClass::Mfma64N: return KClass::Mfma16h;
    case KClass::F32N: return KClass::F32x16h;
    case KClass::Mfma64M: return KClass::Mfma64S;
    case KClass::F32M: return KClass::F32S;
    default: return k;
  }
}

// The 128-row form of a 64-row FP64 MFMA tile class (others: unchanged).
KClass quad_rows(KClass k)
{
  switch (k) {
    case KClass::Mfma16: return KClass::Mfma16q;
    case KClass::Mfma32: return KClass::Mfma32q;
    case KClass::F32N: return KClass::F32Nq;
    case KClass::F32M: return KClass::F32Mq;
    default: return k;
  }
}

void build_tiles(sht_plan* P);

// Tile placement from the host's plan of the tile list: tile (C block, first
// row, first column) goes to position pos; every other position holds a
// padding tile.
struct TilePlace {
  int64_t pos;
  int32_t cb, m0, n0, kc;  // kc: the tile's class
};

__global__ void k_tile_pads(TileRec* t, int64_t n)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TileRec r{};
  r.pad = 1;
  t[i] = r;
}

// Each placed tile's record: its C block record and its first task's operand
// offsets copied in.
__global__ void k_tile_place(TileRec* t, const TilePlace* items, int64_t nitems, const CBlkRec* cblk,
                             const TaskRec* tasks)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const TilePlace p = items[i];
  TileRec r{};
  r.cb = p.cb;
  r.m0 = p.m0;
  r.n0 = p.n0;
  r.kc = p.kc;
  const CBlkRec c = cblk[p.cb];
  r.c_off = c.c_off;
  r.m = c.m;
  r.n = c.n;
  r.t0 = c.t0;
  r.t1 = c.t1;
  if (c.t1 > c.t0) {
    const TaskRec k = tasks[c.t0];
    r.a0 = k.a_off;
    r.b0 = k.b_off;
    r.k0 = k.k;
  }
  t[p.pos] = r;
}

void build_schedule(sht_plan* P)
{
  Arena& arena = P->ctx->arena;
  const int64_t nt = P->ntasks, ncb = P->ncb;
  const ContractMaps& m = P->maps;
  hipStream_t s = nullptr;
  Scratch S(arena);
  P->d_tasks = static_cast<TaskRec*>(arena.allocate(std::max<int64_t>(nt, 1) * sizeof(TaskRec)));
  P->d_cblk = static_cast<CBlkRec*>(arena.allocate(std::max<int64_t>(ncb, 1) * sizeof(CBlkRec)));
  int64_t* cstart = S.get<int64_t>(ncb + 1);
  int64_t* work = S.get<int64_t>(ncb);
  if (nt)
    hipLaunchKernelGGL(k_task_recs, dim3(blocks_for(nt)), dim3(256), 0, s, P->d_ta, P->d_tb, nt, P->A->desc(),
                       P->B->desc(), axis_list(m.a_col), P->d_tasks);
  hipLaunchKernelGGL(k_cstart, dim3(blocks_for(ncb + 1)), dim3(256), 0, s, P->d_tc, nt, ncb, cstart);
  if (ncb)
    hipLaunchKernelGGL(k_cblk_recs, dim3(blocks_for(ncb)), dim3(256), 0, s, P->C->desc(), axis_list(m.c_row),
                       axis_list(m.c_col), cstart, P->d_tasks, P->d_cblk, work);
  SHT_HIP(hipGetLastError());
  std::vector<CBlkRec> cb(static_cast<size_t>(ncb));
  std::vector<int64_t> w(static_cast<size_t>(ncb));
  if (ncb) {
    SHT_HIP(hipMemcpy(cb.data(), P->d_cblk, ncb * sizeof(CBlkRec), hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(w.data(), work, ncb * sizeof(int64_t), hipMemcpyDeviceToHost));
  }
  // useful flops
  double fl = 0.0;
  for (int64_t c = 0; c < ncb; ++c) fl += 2.0 * cb[c].m * static_cast<double>(cb[c].n) * static_cast<double>(w[c]);
  P->flops = fl;
  double sk = 0.0;
  for (int64_t c = 0; c < ncb; ++c) sk += static_cast<double>(w[c]);
  P->sum_k = sk;
  // operands at their stored width (FP32 tensors for the FP32 and
  // split-BF16 forms), C in FP64
  const double ea = static_cast<double>(P->A->elem_size()), es = 8.0;
  P->bytes_in = ea * (static_cast<double>(P->A->nelem) + static_cast<double>(P->B->nelem)) +
                es * static_cast<double>(P->C->nelem);
  P->bytes_c = es * static_cast<double>(P->C->nelem);
  P->h_cblk = std::move(cb);
  P->h_work = std::move(w);
  build_tiles(P);
}

// Operand element reads of the C-stationary tiles (a tile reads its rows of
// every A block and its columns of every B block of its C block's tasks;
// 64-wide tiles assumed), over the elements A and B hold.
double cs_reuse(const sht_plan* P)
{
  double reads = 0.0;
  for (int64_t c = 0; c < P->ncb; ++c) {
    const CBlkRec& r = P->h_cblk[c];
    const double wk = static_cast<double>(P->h_work[c]);
    reads += wk * (static_cast<double>(r.m) * ((r.n + 63) / 64) + static_cast<double>(r.n) * ((r.m + 63) / 64));
  }
  const double held = static_cast<double>(P->A->nelem) + static_cast<double>(P->B->nelem);
  return held > 0.0 ? reads / held : 0.0;
}

void drop_stage32(sht_plan* P)
{
  Arena& arena = P->ctx->arena;
  if (P->d_A32) arena.release(P->d_A32);
  if (P->d_B32) arena.release(P->d_B32);
  P->d_A32 = P->d_B32 = nullptr;
  P->stage32 = false;
}

// The FP32 staging buffers of A and B; false (and nothing kept) when the
// arena cannot hold them.
bool make_stage32(sht_plan* P)
{
  if (P->stage32) return true;
  Arena& arena = P->ctx->arena;
  void* a = nullptr;
  try {
    a = arena.allocate(static_cast<size_t>(std::max<int64_t>(P->A->nelem, 1)) * 4 + kReadSlack);
    P->d_B32 = arena.allocate(static_cast<size_t>(std::max<int64_t>(P->B->nelem, 1)) * 4 + kReadSlack);
  } catch (const Error& e) {
    if (e.code != SHT_ERR_OOM) throw;
    if (a) arena.release(a);
    return false;
  }
  P->d_A32 = a;
  P->stage32 = true;
  return true;
}

// A new FP32 or split-BF16 plan over FP64 operands that starts on the
// C-stationary schedule stages them in FP32 when its tiles read each operand
// element at least kStage32Reuse times on average (the conversion then costs less than reading FP64 values
// at every use; SHT_FP32_STAGING=0 or 1 overrides) and the arena holds the
// copies.  Both forms round every operand to FP32 first, so results are
// bitwise the same with or without the staging.
constexpr double kStage32Reuse = 10.0;
void build_stage32(sht_plan* P)
{
  if ((P->prec != SHT_PREC_FP32 && P->prec != SHT_PREC_BF16X3) || P->A->dtype != SHT_FLOAT64 || P->ntasks == 0)
    return;
  static const int force = [] {
    const char* e = std::getenv("SHT_FP32_STAGING");
    return e ? std::atoi(e) : -1;
  }();
  if (force == 1 || (force < 0 && cs_reuse(P) >= kStage32Reuse)) make_stage32(P);
}

// C-stationary tile lists of the plan's current variant (P->cs_variant):
// 0 automatic (below), 1 without the 128-row tiles, 2 with 128-row tiles in
// any launch (FP64 / FP32), 3 tiles by descending cost (no XCD dealing),
// 4 uniform 64 x 64 four-wave tiles for every block (128 x 32 for the fma
// chain): the fixed-tile reference the autotuner is measured against;
// 5 every tile in one persistent launch (the four-wave FP64 / FP32 /
// split-BF16 classes by block width, XCD queues as for the grouped classes).
void build_tiles(sht_plan* P)
{
  Arena& arena = P->ctx->arena;
  const int64_t ncb = P->ncb;
  const ContractMaps& m = P->maps;
  const std::vector<CBlkRec>& cb = P->h_cblk;
  const std::vector<int64_t>& w = P->h_work;
  const int variant = P->cs_variant;
  if (P->d_tiles) {
    arena.release(P->d_tiles);
    P->d_tiles = nullptr;
  }

  // Tiles per kernel class.  When A is the larger operand (at least twice
  // B), an XCD-aware order: workgroup g runs on XCD g % 8 and every XCD has
  // its own L2, so the tiles of one C block row (which read the same A
  // blocks) go to one XCD, consecutively, and those blocks are shared in that
  // L2.  Rows are dealt to the XCD with the least accumulated cost (most
  // expensive first), the eight queues are interleaved, and short queues are
  // padded with no-work tiles.  Otherwise (A and B alike) the unit dealt is
  // a square of kSq x kSq C blocks (block rows and block columns), so the
  // tiles an XCD runs at one time read a few block rows of A and a few block
  // columns of B -- in launches of at least 32 waves per SIMD; shorter ones
  // (a few waves per SIMD, where the last wave to finish sets the time) order their
  // tiles by descending cost instead.
  const bool by_row = P->A->nelem >= 2 * P->B->nelem;
  auto axes_key = [&](int64_t id, const std::vector<int>& axes) {
    int64_t k = 0, mul = 1;
    for (int ax : axes) {
      k += ((id / P->C->bstride[ax]) % P->C->nblk[ax]) * mul;
      mul *= P->C->nblk[ax];
    }
    return k;
  };
  constexpr int64_t kSq = 16;
  int64_t ncolkeys = 1;
  for (int ax : m.c_col) ncolkeys *= P->C->nblk[ax];
  auto row_key = [&](int64_t id) {
    if (by_row) return axes_key(id, m.c_row);
    return (axes_key(id, m.c_row) / kSq) * (ncolkeys / kSq + 1) + axes_key(id, m.c_col) / kSq;
  };
  constexpr int NK = static_cast<int>(KClass::Count);
  constexpr int kXcd = 8;
  const bool exact = P->prec == SHT_PREC_FP64_EXACT;
  // 128-row FP64 / FP32 tiles only for large launches (more than 64 64-row tiles
  // per SIMD): there a wave's fixed cost (the dependent record loads before
  // its first MFMA) dominates, while small launches need every wave
  int64_t n64 = 0;
  for (int64_t c = 0; c < ncb; ++c) n64 += (cb[c].m + 63) / 64;
  const bool quad_ok = P->prec == SHT_PREC_FP64 || P->prec == SHT_PREC_FP32;
  // variant 5: one persistent launch over every tile (four-wave classes only)
  const bool persist = variant == 5;
  const bool quad = quad_ok && !persist && (variant == 2 || (variant != 1 && variant != 4 && n64 > 64 * 1024));
  const bool uniform = variant == 4;
  const KClass persist_k = P->prec == SHT_PREC_FP64   ? KClass::Persist64
                           : P->prec == SHT_PREC_FP32 ? KClass::Persist32
                                                      : KClass::PersistBf16;
  SHT_REQUIRE(!persist || !exact, SHT_ERR_INVALID, "the persistent variant needs the FP64, FP32 or BF16X3 form");
  // the tile class of C block c (its row tiles may take the 32- or 128-row
  // form of it)
  auto main_class = [&](int64_t c) {
    const int N = cb[c].n;
    if (uniform)
      return exact ? KClass::Exact32
                   : (P->prec == SHT_PREC_FP32 ? KClass::F32W
                                               : (P->prec == SHT_PREC_BF16X3 ? KClass::Bf16W : KClass::Mfma64W));
    if (exact) return N <= 8 ? KClass::Exact8 : (N <= 16 ? KClass::Exact16 : KClass::Exact32);
    if (P->prec == SHT_PREC_FP32) return N <= 16 ? KClass::F32N : (N <= 32 ? KClass::F32M : KClass::F32W);
    if (P->prec == SHT_PREC_BF16X3) return N <= 16 ? KClass::Bf16N : (N <= 32 ? KClass::Bf16S : KClass::Bf16W);
    if (N > 32) return KClass::Mfma64W;
    if (persist) return N <= 16 ? KClass::Mfma64N : KClass::Mfma64M;
    // (FP64 large launches: one-wave 64-row tiles beside the 128-row ones,
    // measured faster there than the four-wave 64-row forms)
    if (quad) return N <= 16 ? KClass::Mfma16 : KClass::Mfma32;
    return N <= 16 ? KClass::Mfma64N : KClass::Mfma64M;
  };
  // Calls f(launch, m0, n0, cost, class) for the tiles of C block c in
  // order: row tiles of 128 rows while more than 96 remain (FP64 and FP32 up
  // to 32 columns, large launches), then 64; the last row tile of an MFMA
  // class takes the 32-row form when it can (the persistent variant: when
  // that form is a four-wave class too).  The launch is the tile's class, or
  // the persistent launch that serves every class.
  auto each_tile = [&](int64_t c, auto&& f) {
    const int M = cb[c].m, N = cb[c].n;
    const KClass k = main_class(c);
    const int cols = kclass_cols(k);
    for (int m0 = 0; m0 < M;) {
      KClass kt = k;
      if (uniform) kt = k;
      else if (!exact && M - m0 <= 32 && (!persist || kclass_waves(half_rows(k)) == 4)) kt = half_rows(k);
      else if (quad && M - m0 > 96) kt = quad_rows(k);
      const int rows = kclass_rows(kt);
      for (int n0 = 0; n0 < N; n0 += cols)
        f(static_cast<int>(persist ? persist_k : kt), m0, n0,
          static_cast<double>(std::min(rows, M - m0)) * std::min(cols, N - n0) * (w[c] + 1), static_cast<int>(kt));
      m0 += rows;
    }
  };
  auto place = [](int64_t pos, int64_t c, int m0, int n0, int kt) {
    return TilePlace{pos, static_cast<int32_t>(c), m0, n0, kt};
  };
  std::vector<int64_t> count(NK, 0);
  for (int64_t c = 0; c < ncb; ++c) each_tile(c, [&](int k, int, int, double, int) { ++count[k]; });
  // Classes ordered by descending tile cost (variant 3, and small launches
  // when A and B are alike) and classes dealt to the XCD queues by group
  std::vector<char> by_cost(NK, 0);
  for (int k = 0; k < NK; ++k)
    by_cost[k] = count[k] > 0 && (variant == 3 || (!by_row && count[k] * kclass_waves(static_cast<KClass>(k)) < 32 * 1024));
  struct CostTile {
    int64_t c;
    int32_t m0, n0;
    double cost;
    int32_t kt;
  };
  std::vector<std::vector<CostTile>> ct(NK);
  for (int k = 0; k < NK; ++k)
    if (by_cost[k]) ct[k].reserve(static_cast<size_t>(count[k]));
  if (std::any_of(by_cost.begin(), by_cost.end(), [](char b) { return b != 0; }))
    for (int64_t c = 0; c < ncb; ++c)
      each_tile(c, [&](int k, int m0, int n0, double cost, int kt) {
        if (by_cost[k]) ct[k].push_back({c, m0, n0, cost, kt});
      });
  // The grouped classes: C blocks in group-key order (a stable counting sort;
  // keys are bounded by the number of C block rows or squares), so the tiles
  // of one group come in C block order.  A group's tiles are dealt to one
  // XCD queue, groups by descending total cost to the least-loaded queue,
  // and the eight queues interleaved with padding tiles.
  std::vector<int64_t> key(static_cast<size_t>(ncb)), corder(static_cast<size_t>(ncb));
  int64_t kmax = 0;
  for (int64_t c = 0; c < ncb; ++c) kmax = std::max(kmax, key[c] = row_key(P->C->indices[c]));
  if (kmax < 8 * ncb + 4096) {
    std::vector<int64_t> start(static_cast<size_t>(kmax) + 2, 0);
    for (int64_t c = 0; c < ncb; ++c) ++start[static_cast<size_t>(key[c]) + 1];
    for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
    for (int64_t c = 0; c < ncb; ++c) corder[start[static_cast<size_t>(key[c])]++] = c;
  } else {
    std::iota(corder.begin(), corder.end(), int64_t{0});
    std::stable_sort(corder.begin(), corder.end(), [&](int64_t x, int64_t y) { return key[x] < key[y]; });
  }
  struct Group {
    int64_t key, n;  // tiles of the class in the group
    double cost;
    int x;           // XCD queue
    int64_t slot0;   // the group's first slot in its queue
  };
  std::vector<std::vector<Group>> grp(NK);
  for (int64_t c : corder)
    each_tile(c, [&](int k, int, int, double cost, int) {
      if (by_cost[k]) return;
      auto& g = grp[k];
      if (g.empty() || g.back().key != key[c]) g.push_back({key[c], 0, 0.0, 0, 0});
      ++g.back().n;
      g.back().cost += cost;
    });
  // where each class's launch starts in the tile list, and its length
  std::vector<int64_t> first(NK, 0), len(NK, 0);
  int64_t total = 0;
  P->launches.clear();
  for (int k = 0; k < NK; ++k) {
    if (!count[k]) continue;
    double ck = 0.0;
    if (by_cost[k]) {
      len[k] = count[k];
      for (const CostTile& t : ct[k]) ck += t.cost;
    } else {
      auto& g = grp[k];
      std::vector<int64_t> go(g.size());
      std::iota(go.begin(), go.end(), int64_t{0});
      std::stable_sort(go.begin(), go.end(), [&](int64_t a, int64_t b) { return g[a].cost > g[b].cost; });
      std::vector<double> load(kXcd, 0.0);
      std::vector<int64_t> qlen(kXcd, 0);
      for (int64_t gi : go) {
        const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        g[gi].x = x;
        g[gi].slot0 = qlen[x];
        qlen[x] += g[gi].n;
        load[x] += g[gi].cost;
        ck += g[gi].cost;
      }
      // (the last slot's trailing padding tiles are left out)
      const int64_t longest = *std::max_element(qlen.begin(), qlen.end());
      int64_t last = 0;
      for (int x = 0; x < kXcd; ++x)
        if (qlen[x] == longest) last = x;
      len[k] = (longest - 1) * kXcd + last + 1;
    }
    first[k] = total;
    total += len[k];
    P->launches.push_back({static_cast<KClass>(k), first[k], len[k], ck});
  }
  // the tiles' positions, written in walk order (the device scatters them)
  int64_t nreal = 0;
  for (int k = 0; k < NK; ++k) nreal += count[k];
  std::vector<TilePlace> items;
  items.reserve(static_cast<size_t>(nreal));
  for (int k = 0; k < NK; ++k) {
    if (!by_cost[k]) continue;
    // descending cost, ties in walk order (the order of a stable sort)
    std::vector<std::pair<double, int64_t>> o(ct[k].size());
    for (size_t i = 0; i < o.size(); ++i) o[i] = {-ct[k][i].cost, static_cast<int64_t>(i)};
    std::sort(o.begin(), o.end());
    for (size_t i = 0; i < o.size(); ++i) {
      const CostTile& t = ct[k][o[i].second];
      items.push_back(place(first[k] + static_cast<int64_t>(i), t.c, t.m0, t.n0, t.kt));
    }
  }
  {
    // second walk in the same order: each group's tiles go to consecutive
    // slots of its queue (tile j of a group at slot slot0 + j)
    std::vector<int64_t> gi(NK, -1), pos(NK, 0);
    std::vector<int64_t> gkey(NK, -1);
    for (int64_t c : corder)
      each_tile(c, [&](int k, int m0, int n0, double, int kt) {
        if (by_cost[k]) return;
        if (gkey[k] != key[c] || gi[k] < 0) {
          ++gi[k];
          gkey[k] = key[c];
          pos[k] = 0;
        }
        const Group& g = grp[k][gi[k]];
        items.push_back(place(first[k] + (g.slot0 + pos[k]++) * kXcd + g.x, c, m0, n0, kt));
      });
  }
  if (total > 0) {
    P->d_tiles = static_cast<TileRec*>(arena.allocate(total * sizeof(TileRec)));
    Scratch S(arena);
    TilePlace* d_items = S.get<TilePlace>(std::max<int64_t>(nreal, 1));
    SHT_HIP(hipMemcpy(d_items, items.data(), items.size() * sizeof(TilePlace), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_tile_pads, dim3(blocks_for(total)), dim3(256), 0, nullptr, P->d_tiles, total);
    hipLaunchKernelGGL(k_tile_place, dim3(blocks_for(nreal)), dim3(256), 0, nullptr, P->d_tiles, d_items, nreal,
                       P->d_cblk, P->d_tasks);
    SHT_HIP(hipGetLastError());
    SHT_HIP(hipStreamSynchronize(nullptr));
  }
}

// ---- A-stationary schedule, built on the device --------------------------
// Per block of a tensor: the keys of its two axis groups (coordinates rolled
// in the group's axis order) and the element counts over each group.
__global__ void k_group_keys(TensorDesc d, AxisList r, AxisList c, int64_t* krow, int64_t* kcol, int32_t* erow,
                             int32_t* ecol)
{
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= d.nnz) return;
  int co[kMaxModes];
  coords_of(d, d.indices[p], co);
  int64_t k = 0, mul = 1, e = 1;
  for (int i = 0; i < r.n; ++i) {
    k += co[r.ax[i]] * mul;
    mul *= d.nblk[r.ax[i]];
    e *= d.bsizes[d.boff[r.ax[i]] + co[r.ax[i]]];
  }
  krow[p] = k;
  erow[p] = static_cast<int32_t>(e);
  k = 0, mul = 1, e = 1;
  for (int i = 0; i < c.n; ++i) {
    k += co[c.ax[i]] * mul;
    mul *= d.nblk[c.ax[i]];
    e *= d.bsizes[d.boff[c.ax[i]] + co[c.ax[i]]];
  }
  kcol[p] = k;
  ecol[p] = static_cast<int32_t>(e);
}

// Sort key of a task in the A-stationary order (block row I, contracted key
// K, block column J), and the flag of every B block that occurs in a task.
__global__ void k_as_taskkey(const int64_t* ta, const int64_t* tb, int64_t nt, const int64_t* aI, const int64_t* aK,
                             const int64_t* bJ, uint64_t nK, uint64_t nJ, uint64_t* key, char* used)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int64_t a = ta[t], b = tb[t];
  key[t] = (static_cast<uint64_t>(aI[a]) * nK + static_cast<uint64_t>(aK[a])) * nJ + static_cast<uint64_t>(bJ[b]);
  used[b] = 1;
}

// First task (in A-stationary order) of every A block.
__global__ void k_as_segflag(const int64_t* order, const int64_t* ta, int64_t nt, int64_t* flag)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  flag[t] = (t == 0 || ta[order[t]] != ta[order[t - 1]]) ? 1 : 0;
}

// Per A block (segment s of the sorted tasks): its panel and slot count, and
// its sweep counts at the two sweep widths.
__global__ void k_as_segs(const int64_t* order, const int64_t* ta, const int64_t* tb, const int64_t* flag,
                          const int64_t* incl, int64_t nt, const int32_t* b_panel, const int32_t* pn_nslots,
                          int64_t* seg_a, int32_t* seg_panel, int64_t* seg_nslots, int64_t* seg_nsw,
                          int64_t* seg_nsw4)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt || !flag[t]) return;
  const int64_t s = incl[t] - 1;
  const int32_t pn = b_panel[tb[order[t]]];
  const int64_t ns = pn_nslots[pn];
  seg_a[s] = ta[order[t]];
  seg_panel[s] = pn;
  seg_nslots[s] = ns;
  seg_nsw[s] = (ns + kAsSweepSlots - 1) / kAsSweepSlots;
  seg_nsw4[s] = (ns + kAsNarrowSlots - 1) / kAsNarrowSlots;
}

// The sweep records of every A block at one sweep width (cap slots).
__global__ void k_as_sweeps(int64_t nseg, const int64_t* seg_a, const int32_t* seg_panel, const int64_t* seg_nslots,
                            const int64_t* sweep_base, const int64_t* slot_base, const int64_t* a_off,
                            const int32_t* aM, const int32_t* aK, const int32_t* pn_pitch, const int64_t* pn_bt,
                            int bf3, int cap, AsSweep* out)
{
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const int64_t a = seg_a[s], ns = seg_nslots[s];
  const int32_t pn = seg_panel[s];
  const int32_t pitch = pn_pitch[pn];
  for (int64_t g = 0, w = sweep_base[s]; g < ns; g += cap, ++w) {
    AsSweep sw;
    sw.a_off = a_off[a];
    sw.b_off = bf3 ? pn_bt[pn] + 16 * g * pitch : pn_bt[pn] + 16 * g;
    sw.c0 = 16 * (slot_base[s] + g);
    sw.k = aK[a];
    sw.m = aM[a];
    sw.pitch = pitch;
    sw.ns = static_cast<int32_t>(min<int64_t>(cap, ns - g));
    out[w] = sw;
  }
}

__global__ void k_fill_i64(int64_t* x, int64_t n, int64_t v)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// The C column offsets of every sweep column a task covers: column j of the
// task's partner inside the panel -> C(0, j) of its C block; marked -(v + 2)
// when the task is the first (smallest contracted key) of its C block, i.e.
// the first sweep of the block row that touches the column.
__global__ void k_as_ccol(const int64_t* order, const int64_t* ta, const int64_t* tb, const int64_t* tc,
                          const int64_t* incl, int64_t nt, const int64_t* slot_base, const int32_t* aM,
                          const int32_t* bN, const int32_t* b_col0, const int64_t* c_off, const int64_t* cstart,
                          int64_t* ccol)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int64_t o = order[t];
  const int64_t a = ta[o], b = tb[o], c = tc[o];
  const int64_t base = 16 * slot_base[incl[t] - 1] + b_col0[b];
  const int64_t M = aM[a], v0 = c_off[c];
  const bool first = cstart[c] == o;
  for (int32_t j = 0; j < bN[b]; ++j) {
    const int64_t v = v0 + j * M;
    ccol[base + j] = first ? -v - 2 : v;
  }
}

__global__ void k_as_rowsegs(int64_t nseg, const int64_t* seg_a, const int64_t* aI, const int32_t* aM,
                             const int32_t* aK, const int64_t* seg_nslots, int64_t* out_I, int32_t* out_M,
                             double* out_work)
{
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const int64_t a = seg_a[s];
  out_I[s] = aI[a];
  out_M[s] = aM[a];
  out_work[s] = static_cast<double>(aK[a]) * 16.0 * static_cast<double>(seg_nslots[s]);
}

__global__ void k_touch(const int64_t* tc, int64_t nt, char* touched)
{
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < nt) touched[tc[t]] = 1;
}

// A-stationary schedule from the task list, on the device: tasks sorted by
// (block row I, contracted key K, block column J) with a radix sort; one panel
// per contracted key holding the key's partners (B blocks that occur in a
// task) side by side (panel table on the host: one entry per B block); one
// sweep list entry per A block and group of cap slots, built from segment
// scans; the C offset of every sweep column written by one thread per task.
// The sweep columns of an A block are the same at both sweep widths, so the
// 8-slot and 4-slot sweep lists share one column array.  Units (row chunks of
// every block row, dealt XCD-aware) are built on the host from the per-row
// totals.
void build_as_schedule_impl(sht_plan* P)
{
  const int64_t nt = P->ntasks;
  const ContractMaps& m = P->maps;
  const sht_tensor* A = P->A;
  const sht_tensor* B = P->B;
  const sht_tensor* C = P->C;
  Arena& arena = P->ctx->arena;
  Scratch S(arena);
  hipStream_t st = nullptr;
  const bool bf3 = P->prec == SHT_PREC_BF16X3;
  const bool exact = P->prec == SHT_PREC_FP64_EXACT;
  const bool narrow = !exact && !bf3;
  const int64_t na = static_cast<int64_t>(A->indices.size()), nbk = static_cast<int64_t>(B->indices.size());
  const TensorDesc dA = A->desc(), dB = B->desc();

  // group keys and element counts of every A and B block
  int64_t* aI = S.get<int64_t>(na);
  int64_t* aK = S.get<int64_t>(na);
  int32_t* aM = S.get<int32_t>(na);
  int32_t* aKe = S.get<int32_t>(na);
  int64_t* bK = S.get<int64_t>(nbk);
  int64_t* bJ = S.get<int64_t>(nbk);
  int32_t* bKe = S.get<int32_t>(nbk);
  int32_t* bN = S.get<int32_t>(nbk);
  if (na)
    hipLaunchKernelGGL(k_group_keys, dim3(blocks_for(na)), dim3(256), 0, st, dA, axis_list(m.a_row),
                       axis_list(m.a_col), aI, aK, aM, aKe);
  if (nbk)
    hipLaunchKernelGGL(k_group_keys, dim3(blocks_for(nbk)), dim3(256), 0, st, dB, axis_list(m.b_row),
                       axis_list(m.b_col), bK, bJ, bKe, bN);
  uint64_t nI = 1, nK = 1, nJ = 1;
  for (int ax : m.a_row) nI *= static_cast<uint64_t>(A->nblk[ax]);
  for (int ax : m.a_col) nK *= static_cast<uint64_t>(A->nblk[ax]);
  for (int ax : m.b_col) nJ *= static_cast<uint64_t>(B->nblk[ax]);
  SHT_REQUIRE(static_cast<double>(nI) * static_cast<double>(nK) * static_cast<double>(nJ) < 9.0e18,
              SHT_ERR_UNSUPPORTED, "block index space too large for 64-bit schedule keys");

  // tasks in A-stationary order
  uint64_t* key = S.get<uint64_t>(nt);
  uint64_t* key_s = S.get<uint64_t>(nt);
  int64_t* iota = S.get<int64_t>(nt);
  int64_t* order = S.get<int64_t>(nt);
  char* used = S.get<char>(nbk);
  SHT_HIP(hipMemsetAsync(used, 0, std::max<int64_t>(nbk, 1), st));
  if (nt) {
    hipLaunchKernelGGL(k_as_taskkey, dim3(blocks_for(nt)), dim3(256), 0, st, P->d_ta, P->d_tb, nt, aI, aK, bJ, nK, nJ,
                       key, used);
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(nt)), dim3(256), 0, st, iota, nt);
    size_t tmp = 0;
    const int bits = bits_for(nI * nK * nJ);
    SHT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key_s, iota, order, static_cast<int>(nt), 0, bits,
                                               st));
    SHT_HIP(hipcub::DeviceRadixSort::SortPairs(S.bytes(tmp), tmp, key, key_s, iota, order, static_cast<int>(nt), 0,
                                               bits, st));
  }
  SHT_HIP(hipGetLastError());

  // panels (host table over the B blocks): the used B blocks of each
  // contracted key in ascending block column, side by side
  std::vector<char> h_used(static_cast<size_t>(nbk));
  std::vector<int64_t> hbK(nbk), hbJ(nbk);
  std::vector<int32_t> hbKe(nbk), hbN(nbk);
  if (nbk) {
    SHT_HIP(hipMemcpy(h_used.data(), used, nbk, hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(hbK.data(), bK, nbk * 8, hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(hbJ.data(), bJ, nbk * 8, hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(hbKe.data(), bKe, nbk * 4, hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(hbN.data(), bN, nbk * 4, hipMemcpyDeviceToHost));
  }
  std::vector<int64_t> usedq;
  for (int64_t q = 0; q < nbk; ++q)
    if (h_used[q]) usedq.push_back(q);
  std::sort(usedq.begin(), usedq.end(), [&](int64_t x, int64_t y) {
    if (hbK[x] != hbK[y]) return hbK[x] < hbK[y];
    return hbJ[x] < hbJ[y];
  });
  std::vector<int32_t> b_panel(nbk, -1), b_col0(nbk, 0);
  std::vector<int32_t> pn_nslots, pn_pitch;
  std::vector<int64_t> pn_bt;
  std::vector<BtRec> bt;
  int64_t bt_total = 0;
  for (size_t u = 0; u < usedq.size();) {
    size_t e = u;
    while (e < usedq.size() && hbK[usedq[e]] == hbK[usedq[u]]) ++e;
    const int32_t pk = hbKe[usedq[u]];
    int64_t ncol = 0;
    for (size_t x = u; x < e; ++x) ncol += hbN[usedq[x]];
    SHT_REQUIRE(ncol < (int64_t(1) << 30), SHT_ERR_UNSUPPORTED, "panel too wide");
    const int64_t npad = (ncol + kBtPad - 1) / kBtPad * kBtPad;
    const int32_t pitch = bf3 ? (pk + kBfK - 1) / kBfK * kBfK : static_cast<int32_t>(npad);
    const int32_t id = static_cast<int32_t>(pn_nslots.size());
    int64_t c = 0;
    for (size_t x = u; x < e; ++x) {
      const int64_t q = usedq[x];
      b_panel[q] = id;
      b_col0[q] = static_cast<int32_t>(c);
      bt.push_back(BtRec{B->offsets[q], bf3 ? bt_total + c * pitch : bt_total + c, pk, hbN[q], pitch, 0});
      c += hbN[q];
    }
    pn_nslots.push_back(static_cast<int32_t>(npad / 16));
    pn_pitch.push_back(pitch);
    pn_bt.push_back(bt_total);
    bt_total += bf3 ? npad * pitch : static_cast<int64_t>((pk + 15) / 16 * 16) * npad;
    u = e;
  }
  const int64_t npanel = static_cast<int64_t>(pn_nslots.size());
  auto up = [&](const void* src, size_t bytes) {
    void* d = S.bytes(std::max<size_t>(bytes, 1));
    if (bytes) SHT_HIP(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
    return d;
  };
  const int32_t* d_b_panel = static_cast<const int32_t*>(up(b_panel.data(), nbk * 4));
  const int32_t* d_b_col0 = static_cast<const int32_t*>(up(b_col0.data(), nbk * 4));
  const int32_t* d_pn_nslots = static_cast<const int32_t*>(up(pn_nslots.data(), npanel * 4));
  const int32_t* d_pn_pitch = static_cast<const int32_t*>(up(pn_pitch.data(), npanel * 4));
  const int64_t* d_pn_bt = static_cast<const int64_t*>(up(pn_bt.data(), npanel * 8));

  // A blocks = segments of the sorted tasks; sweep and slot offsets by scans
  int64_t* flag = S.get<int64_t>(nt);
  int64_t* incl = S.get<int64_t>(nt);
  int64_t nseg = 0;
  if (nt) {
    hipLaunchKernelGGL(k_as_segflag, dim3(blocks_for(nt)), dim3(256), 0, st, order, P->d_ta, nt, flag);
    size_t tmp = 0;
    SHT_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, flag, incl, static_cast<int>(nt), st));
    SHT_HIP(hipcub::DeviceScan::InclusiveSum(S.bytes(tmp), tmp, flag, incl, static_cast<int>(nt), st));
    SHT_HIP(hipMemcpy(&nseg, incl + nt - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
  }
  int64_t* seg_a = S.get<int64_t>(nseg);
  int32_t* seg_panel = S.get<int32_t>(nseg);
  int64_t* seg_nslots = S.get<int64_t>(nseg);
  int64_t* seg_nsw = S.get<int64_t>(nseg);
  int64_t* seg_nsw4 = S.get<int64_t>(nseg);
  int64_t* slot_base = S.get<int64_t>(nseg);
  int64_t* sweep_base = S.get<int64_t>(nseg);
  int64_t* sweep_base4 = S.get<int64_t>(nseg);
  int64_t nsweep = 0, nsweep4 = 0, nslot = 0;
  if (nseg) {
    hipLaunchKernelGGL(k_as_segs, dim3(blocks_for(nt)), dim3(256), 0, st, order, P->d_ta, P->d_tb, flag, incl, nt,
                       d_b_panel, d_pn_nslots, seg_a, seg_panel, seg_nslots, seg_nsw, seg_nsw4);
    auto excl = [&](const int64_t* in, int64_t* out) {
      size_t tmp = 0;
      SHT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, static_cast<int>(nseg), st));
      SHT_HIP(hipcub::DeviceScan::ExclusiveSum(S.bytes(tmp), tmp, in, out, static_cast<int>(nseg), st));
      int64_t lo = 0, lc = 0;
      SHT_HIP(hipMemcpy(&lo, out + nseg - 1, 8, hipMemcpyDeviceToHost));
      SHT_HIP(hipMemcpy(&lc, in + nseg - 1, 8, hipMemcpyDeviceToHost));
      return lo + lc;
    };
    nslot = excl(seg_nslots, slot_base);
    nsweep = excl(seg_nsw, sweep_base);
    if (narrow) nsweep4 = excl(seg_nsw4, sweep_base4);
  }
  SHT_REQUIRE(nsweep < (int64_t(1) << 31) && nsweep4 < (int64_t(1) << 31), SHT_ERR_UNSUPPORTED, "schedule too large");

  // sweep lists and their shared column offsets (kept by the plan)
  int64_t* a_off = A->d_offsets;
  P->d_as_sweep = static_cast<AsSweep*>(arena.allocate(std::max<int64_t>(nsweep, 1) * sizeof(AsSweep)));
  P->d_as_ccol = static_cast<int64_t*>(arena.allocate(std::max<int64_t>(16 * nslot, 1) * sizeof(int64_t)));
  P->as_nccol = 16 * nslot;
  if (narrow) P->d_as_sweep4 = static_cast<AsSweep*>(arena.allocate(std::max<int64_t>(nsweep4, 1) * sizeof(AsSweep)));
  if (nseg) {
    hipLaunchKernelGGL(k_as_sweeps, dim3(blocks_for(nseg)), dim3(256), 0, st, nseg, seg_a, seg_panel, seg_nslots,
                       sweep_base, slot_base, a_off, aM, aKe, d_pn_pitch, d_pn_bt, bf3 ? 1 : 0, kAsSweepSlots,
                       P->d_as_sweep);
    if (narrow)
      hipLaunchKernelGGL(k_as_sweeps, dim3(blocks_for(nseg)), dim3(256), 0, st, nseg, seg_a, seg_panel, seg_nslots,
                         sweep_base4, slot_base, a_off, aM, aKe, d_pn_pitch, d_pn_bt, 0, kAsNarrowSlots,
                         P->d_as_sweep4);
    hipLaunchKernelGGL(k_fill_i64, dim3(blocks_for(16 * nslot)), dim3(256), 0, st, P->d_as_ccol, 16 * nslot,
                       int64_t(-1));
    // first task of every C block in the task list (sorted by C position)
    int64_t* cstart = S.get<int64_t>(P->ncb + 1);
    hipLaunchKernelGGL(k_cstart, dim3(blocks_for(P->ncb + 1)), dim3(256), 0, st, P->d_tc, nt, P->ncb, cstart);
    hipLaunchKernelGGL(k_as_ccol, dim3(blocks_for(nt)), dim3(256), 0, st, order, P->d_ta, P->d_tb, P->d_tc, incl, nt,
                       slot_base, aM, bN, d_b_col0, C->d_offsets, cstart, P->d_as_ccol);
  }
  SHT_HIP(hipGetLastError());

  // block rows: runs of A blocks with the same row key (host, per segment)
  std::vector<int64_t> sI(nseg), snsw(nseg), snsw4(nseg);
  std::vector<int32_t> sM(nseg);
  std::vector<double> swork(nseg);
  if (nseg) {
    int64_t* oI = S.get<int64_t>(nseg);
    int32_t* oM = S.get<int32_t>(nseg);
    double* ow = S.get<double>(nseg);
    hipLaunchKernelGGL(k_as_rowsegs, dim3(blocks_for(nseg)), dim3(256), 0, st, nseg, seg_a, aI, aM, aKe, seg_nslots,
                       oI, oM, ow);
    SHT_HIP(hipMemcpy(sI.data(), oI, nseg * 8, hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(sM.data(), oM, nseg * 4, hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(swork.data(), ow, nseg * 8, hipMemcpyDeviceToHost));
    SHT_HIP(hipMemcpy(snsw.data(), seg_nsw, nseg * 8, hipMemcpyDeviceToHost));
    if (narrow) SHT_HIP(hipMemcpy(snsw4.data(), seg_nsw4, nseg * 8, hipMemcpyDeviceToHost));
  }
  struct Row {
    int32_t sweep0, nsweep;
    int64_t rows;
    double work;  // sum over sweeps of k * columns
  };
  auto rows_of = [&](const std::vector<int64_t>& nsw) {
    std::vector<Row> out;
    int64_t w = 0;
    for (int64_t s = 0; s < nseg;) {
      Row r{static_cast<int32_t>(w), 0, sM[s], 0.0};
      const int64_t I = sI[s];
      for (; s < nseg && sI[s] == I; ++s) {
        r.nsweep += static_cast<int32_t>(nsw[s]);
        r.work += swork[s];
        w += nsw[s];
      }
      out.push_back(r);
    }
    return out;
  };
  const std::vector<Row> brows = rows_of(snsw);
  const std::vector<Row> brows4 = narrow ? rows_of(snsw4) : std::vector<Row>();
  // XCD-aware order.  Workgroup g runs on XCD g % 8 (round-robin dispatch),
  // and every XCD has its own L2: all units of a block row go to one XCD, so
  // the row's panels are shared in that L2.  Rows are dealt to the XCD with
  // the least accumulated work, then the eight queues are interleaved; short
  // queues are padded with empty units.  Rows are taken in descending order
  // of their sweep-list cost, which is the running time of each of the row's
  // workgroups (every one walks the whole list), not of the row's total
  // work: longest-processing-time-first keeps the last round of workgroups
  // short (C1: 52.8 -> 49.8 ms FP64, 38.4 -> 36.6 ms FP32 against ordering by
  // total row work).
  constexpr int kXcd = 8;
  // Units of one block row are grouped per_wg to a workgroup (padded with
  // ghost units, so a workgroup never mixes rows): the waves of a workgroup
  // run the same sweep list on neighbouring row chunks and share the panel
  // fragments (through LDS for the FP64/FP32 classes, L1 for split-BF16).
  auto build = [&](const std::vector<Row>& rows, int step, int per_wg) {
    std::vector<int64_t> ro(rows.size());
    std::iota(ro.begin(), ro.end(), 0);
    std::stable_sort(ro.begin(), ro.end(), [&](int64_t x, int64_t y) { return rows[x].work > rows[y].work; });
    std::vector<std::vector<AsUnit>> q(kXcd);
    std::vector<double> load(kXcd, 0.0);
    for (int64_t r : ro) {
      const Row& br = rows[r];
      const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      for (int64_t r0 = 0; r0 < br.rows; r0 += step)
        q[x].push_back(AsUnit{br.sweep0, br.nsweep, static_cast<int32_t>(r0), 0});
      // ghost units: pad the row's workgroup; they walk the same sweep list
      // (shared panel loads and barriers) without rows of their own
      while (q[x].size() % per_wg) q[x].push_back(AsUnit{br.sweep0, br.nsweep, 0, 1});
      load[x] += br.work * static_cast<double>(br.rows);
    }
    size_t longest = 0;
    for (const auto& v : q) longest = std::max(longest, v.size());
    std::vector<AsUnit> out;
    out.reserve(longest * kXcd);
    for (size_t slot = 0; slot < longest; slot += per_wg)
      for (int x = 0; x < kXcd; ++x)
        for (int w = 0; w < per_wg; ++w)
          out.push_back(slot + w < q[x].size() ? q[x][slot + w] : AsUnit{0, 0, 0, 0});
    while (!out.empty() && out.back().nsweep == 0) out.pop_back();
    while (out.size() % per_wg) out.push_back(AsUnit{0, 0, 0, 0});
    return out;
  };
  std::vector<AsUnit> us = build(brows, exact ? kAsUnitRowsW : kAsUnitRows, 1);
  std::vector<AsUnit> usw = exact ? std::vector<AsUnit>() : build(brows, kAsUnitRows, kAsGroupWaves);
  std::vector<AsUnit> us4 = narrow ? build(brows4, kAsUnitRowsN, 1) : std::vector<AsUnit>();
  auto keep = [&](const void* src, size_t bytes) {
    void* d = arena.allocate(std::max<size_t>(bytes, 1));
    if (bytes) SHT_HIP(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
    return d;
  };
  P->d_as_unit = static_cast<AsUnit*>(keep(us.data(), us.size() * sizeof(AsUnit)));
  P->d_as_unitw = static_cast<AsUnit*>(keep(usw.data(), usw.size() * sizeof(AsUnit)));
  if (narrow) P->d_as_unit4 = static_cast<AsUnit*>(keep(us4.data(), us4.size() * sizeof(AsUnit)));
  P->d_bt_rec = static_cast<BtRec*>(keep(bt.data(), bt.size() * sizeof(BtRec)));
  // (split-BF16: three bf16 planes, or one FP32 panel in the same layout for
  // the grouped class, whichever schedule runs)
  const size_t bt_bytes = bf3 ? 3 * sizeof(uint16_t) : (P->prec == SHT_PREC_FP32 ? sizeof(float) : sizeof(double));
  P->d_bt = arena.allocate(std::max<int64_t>(bt_total, 1) * bt_bytes);
  // pad columns and rows of the panels stay zero: the conversion writes only
  // the blocks' own columns
  SHT_HIP(hipMemsetAsync(P->d_bt, 0, std::max<int64_t>(bt_total, 1) * bt_bytes, st));
  {
    // every C block has a task: first sweeps may start from zeros
    const int64_t ncb = static_cast<int64_t>(C->indices.size());
    char* touched = S.get<char>(ncb);
    SHT_HIP(hipMemsetAsync(touched, 0, std::max<int64_t>(ncb, 1), st));
    if (nt) hipLaunchKernelGGL(k_touch, dim3(blocks_for(nt)), dim3(256), 0, st, P->d_tc, nt, touched);
    std::vector<char> h(static_cast<size_t>(ncb));
    if (ncb) SHT_HIP(hipMemcpy(h.data(), touched, ncb, hipMemcpyDeviceToHost));
    P->as_all_touched = std::all_of(h.begin(), h.end(), [](char c) { return c != 0; });
    P->d_as_zero = static_cast<double*>(arena.allocate(kAsZeroLen * sizeof(double)));
    SHT_HIP(hipMemsetAsync(P->d_as_zero, 0, kAsZeroLen * sizeof(double), st));
    if (!P->d_as_qctr) P->d_as_qctr = static_cast<unsigned long long*>(arena.allocate(8 * sizeof(unsigned long long)));
  }
  SHT_HIP(hipStreamSynchronize(st));
  P->bt_total = bt_total;
  P->bt_nrec = static_cast<int64_t>(bt.size());
  P->as_nunits = static_cast<int64_t>(us.size());
  P->as_nunitsw = static_cast<int64_t>(usw.size());
  P->as_nunits4 = static_cast<int64_t>(us4.size());
  P->as_nsweep = nsweep;
  P->as_npanel = npanel;
  P->as_built = true;
}

void build_as_schedule(sht_plan* P)
{
  if (P->as_built) return;
  const double t0 = wall_ms();
  build_as_schedule_impl(P);
  P->ms_astat = wall_ms() - t0;
}

// The A-stationary kernels address an A block with 32-bit element offsets:
// every A block must hold fewer than 2^28 elements (bound by the largest
// block size of each mode).
bool as_supported(const sht_plan* P)
{
  double vmax = 1.0;
  for (int ax = 0; ax < P->A->N; ++ax) {
    int b = 0;
    for (int x : P->A->bs[ax]) b = std::max(b, x);
    vmax *= b;
  }
  return vmax < static_cast<double>(1 << 28);
}

}  // namespace

// (also used by the serving loop's merged A-stationary launches)
BtFormat bt_format(const sht_plan* P)
{
  if (P->prec == SHT_PREC_BF16X3) return P->schedule == 2 ? BtFormat::F32K : BtFormat::BF16X3;
  return P->prec == SHT_PREC_FP32 ? BtFormat::F32T : BtFormat::F64T;
}

AsClass as_class(const sht_plan* P)
{
  const bool wide = P->schedule == 2, narrow = P->schedule == 3;
  switch (P->prec) {
    case SHT_PREC_FP64_EXACT: return AsClass::Exact;
    case SHT_PREC_FP32: return wide ? AsClass::Mfma32G : (narrow ? AsClass::Mfma32N : AsClass::Mfma32);
    case SHT_PREC_BF16X3: return wide ? AsClass::Bf16x3G : AsClass::Bf16x3;
    default: return wide ? AsClass::Mfma64G : (narrow ? AsClass::Mfma64N : AsClass::Mfma64);
  }
}

namespace {

// schedule 3 (narrow sweeps) exists for the FP64 and FP32 MFMA forms
bool narrow_supported(const sht_plan* P) { return P->prec == SHT_PREC_FP64 || P->prec == SHT_PREC_FP32; }

// The schedule a new plan starts with (sht_plan_autotune measures instead):
// the grouped A-stationary form (2) where it pays -- A at least 8x B, so A
// dominates the traffic and is read once, and tasks with contracted lengths
// of 64 or more on average, so each sweep's k loop is long against its C
// tile load and store (config 1: 186 on average, 52 ms against 94 ms
// C-stationary); otherwise C-stationary tiles (0), which win for short k
// (configs 0, 2: 16 and 14) and for A and B alike (config 3).  The fma-chain
// form stays C-stationary (its A-stationary kernel is slower at every
// measured size).
int default_schedule(const sht_plan* P)
{
  if (P->prec == SHT_PREC_FP64_EXACT || P->ntasks == 0 || !as_supported(P)) return 0;
  const double mean_k = P->sum_k / static_cast<double>(P->ntasks);
  const bool a_dominant = static_cast<double>(P->A->nelem) >= 8.0 * static_cast<double>(P->B->nelem);
  return a_dominant && mean_k >= 64.0 ? 2 : 0;
}

void setup_staging(sht_plan* P)
{
  const ContractMaps& m = P->maps;
  P->permA = !canonical(m.a_row, m.a_col);
  P->permB = !canonical(m.b_row, m.b_col);
  P->permC = !canonical(m.c_row, m.c_col);
  auto order = [](const std::vector<int>& a, const std::vector<int>& b, int* out) {
    int i = 0;
    for (int x : a) out[i++] = x;
    for (int x : b) out[i++] = x;
  };
  order(m.a_row, m.a_col, P->orderA);
  order(m.b_row, m.b_col, P->orderB);
  order(m.c_row, m.c_col, P->orderC);
  Arena& arena = P->ctx->arena;
  if (P->permA) P->d_Ap = arena.allocate(std::max<int64_t>(P->A->nelem, 1) * P->A->elem_size() + kReadSlack);
  if (P->permB) P->d_Bp = arena.allocate(std::max<int64_t>(P->B->nelem, 1) * P->B->elem_size() + kReadSlack);
  if (P->permC) P->d_Cp = arena.allocate(std::max<int64_t>(P->C->nelem, 1) * 8 + kReadSlack);
}

}  // namespace

namespace {

bool zero_bytes(double v)
{
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u == 0;
}

void permute(const sht_tensor* t, const void* in, void* out, const int* order, int inverse, hipStream_t s)
{
  PermArgs pa{};
  pa.d = t->desc();
  pa.in = in;
  pa.out = out;
  for (int i = 0; i < t->N; ++i) pa.order[i] = order[i];
  pa.inverse = inverse;
  launch_permute(pa, t->dtype == SHT_FLOAT64, s);
}

}  // namespace

void enqueue_stage_in(sht_plan* P, double alpha, double beta, hipStream_t s)
{
  if (P->permA && !zero_bytes(alpha)) permute(P->A, P->A->d_data, P->d_Ap, P->orderA, 0, s);
  if (P->permB && !zero_bytes(alpha)) permute(P->B, P->B->d_data, P->d_Bp, P->orderB, 0, s);
  if (P->permC && !zero_bytes(beta)) permute(P->C, P->C->d_data, P->d_Cp, P->orderC, 0, s);
}

void enqueue_stage_out(sht_plan* P, hipStream_t s)
{
  if (P->permC) permute(P->C, P->d_Cp, P->C->d_data, P->orderC, 1, s);
}

// Streams for the C-stationary classes of one execution: one per class, at
// most four (the caller's and three plan-owned ones; SHT_CS_STREAMS=1..4
// overrides the maximum).
int cs_streams(const sht_plan* P)
{
  static const int cap = [] {
    const char* e = std::getenv("SHT_CS_STREAMS");
    const int v = e ? std::atoi(e) : 1 + sht_plan::kSideStreams;
    return std::max(1, std::min(v, 1 + sht_plan::kSideStreams));
  }();
  return static_cast<int>(std::min<size_t>(P->launches.size(), static_cast<size_t>(cap)));
}

// Non-blocking streams of default priority, created on first use.
void ensure_side_streams(sht_plan* P, int n)
{
  if (!P->fork_ev) SHT_HIP(hipEventCreateWithFlags(&P->fork_ev, hipEventDisableTiming));
  for (int k = 0; k < n; ++k) {
    if (!P->side[k]) SHT_HIP(hipStreamCreateWithFlags(&P->side[k], hipStreamNonBlocking));
    if (!P->join_ev[k]) SHT_HIP(hipEventCreateWithFlags(&P->join_ev[k], hipEventDisableTiming));
  }
}

void enqueue(sht_plan* P, double alpha, double beta, hipStream_t s, std::vector<hipEvent_t>* ev)
{
  const bool az = zero_bytes(alpha), bz = zero_bytes(beta);
  enqueue_stage_in(P, alpha, beta, s);
  const void* Ad = P->permA ? P->d_Ap : P->A->d_data;
  const void* Bd = P->permB ? P->d_Bp : P->B->d_data;
  void* Cd = P->permC ? P->d_Cp : P->C->d_data;
  if (P->schedule >= 1) {
    // beta == 0 and every C block has a task: the first sweep of every C
    // column starts from zeros and no C = beta * C pass is needed
    const bool zinit = bz && !az && P->as_all_touched;
    if (!zinit) launch_scale(static_cast<double*>(Cd), P->C->nelem, beta, bz ? 1 : 0, s);
    if (!az) {
      launch_bt(bt_format(P), P->d_bt_rec, P->bt_nrec, Bd, P->B->dtype == SHT_FLOAT32, P->d_bt, P->bt_total, s);
      AsArgs g{};
      g.A = Ad;
      g.B = P->d_bt;
      g.bt_total = P->bt_total;
      g.C = Cd;
      const AsClass cls = as_class(P);
      const bool wide = as_group_waves(cls) > 1, narrow = P->schedule == 3;
      g.sweeps = narrow ? P->d_as_sweep4 : P->d_as_sweep;
      g.ccol = P->d_as_ccol;  // (both sweep widths share the column offsets)
      g.units = narrow ? P->d_as_unit4 : (wide ? P->d_as_unitw : P->d_as_unit);
      g.nunits = narrow ? P->as_nunits4 : (wide ? P->as_nunitsw : P->as_nunits);
      g.alpha = alpha;
      g.alpha_one = alpha == 1.0 ? 1 : 0;
      g.zero = zinit ? P->d_as_zero : nullptr;
      g.qctr = P->d_as_qctr;
      g.f32 = P->A->dtype == SHT_FLOAT32 ? 1 : 0;
      if (ev) SHT_HIP(hipEventRecord((*ev)[0], s));
      launch_as(cls, g, s);
      if (ev) SHT_HIP(hipEventRecord((*ev)[1], s));
    }
    enqueue_stage_out(P, s);
    return;
  }
  ExecArgs a{};
  a.A = Ad;
  a.B = Bd;
  a.C = Cd;
  a.cblk = P->d_cblk;
  a.tasks = P->d_tasks;
  a.alpha = alpha;
  a.beta = beta;
  a.alpha_zero = az ? 1 : 0;
  a.beta_zero = bz ? 1 : 0;
  a.alpha_one = alpha == 1.0 ? 1 : 0;
  a.f32 = P->A->dtype == SHT_FLOAT32 ? 1 : 0;
  if (P->stage32 && !az) {
    // FP64 operands read many times: converted once, the tiles read FP32
    launch_convert(Ad, true, P->d_A32, false, P->A->nelem, s);
    launch_convert(Bd, true, P->d_B32, false, P->B->nelem, s);
    a.A = P->d_A32;
    a.B = P->d_B32;
    a.f32 = 1;
  }
  // Tile classes are independent (every C element belongs to one tile of one
  // class), so with several classes they run concurrently: the caller's
  // stream forks to up to three plan-owned streams and joins them again, so
  // the execution stays ordered on s (and is captured whole into a graph).
  // Classes go to the least-loaded stream, largest first.  Timed executions
  // keep one stream so that each class's events bracket it alone.
  const int nstr = ev ? 1 : cs_streams(P);
  std::vector<int> on(P->launches.size(), 0);
  if (nstr > 1) {
    ensure_side_streams(P, nstr - 1);
    std::vector<size_t> ord(P->launches.size());
    std::iota(ord.begin(), ord.end(), size_t{0});
    std::stable_sort(ord.begin(), ord.end(),
                     [&](size_t x, size_t y) { return P->launches[x].cost > P->launches[y].cost; });
    std::vector<double> load(nstr, 0.0);
    for (size_t i : ord) {
      const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      on[i] = k;
      load[k] += P->launches[i].cost;
    }
    SHT_HIP(hipEventRecord(P->fork_ev, s));
    for (int k = 0; k + 1 < nstr; ++k) SHT_HIP(hipStreamWaitEvent(P->side[k], P->fork_ev, 0));
  }
  for (size_t i = 0; i < P->launches.size(); ++i) {
    const auto& L = P->launches[i];
    a.tiles = P->d_tiles + L.first;
    a.ntiles = L.count;
    hipStream_t si = on[i] == 0 ? s : P->side[on[i] - 1];
    if (ev) SHT_HIP(hipEventRecord((*ev)[2 * i], si));
    launch_exec(L.k, a, si);
    if (ev) SHT_HIP(hipEventRecord((*ev)[2 * i + 1], si));
  }
  if (nstr > 1)
    for (int k = 0; k + 1 < nstr; ++k) {
      SHT_HIP(hipEventRecord(P->join_ev[k], P->side[k]));
      SHT_HIP(hipStreamWaitEvent(s, P->join_ev[k], 0));
    }
  enqueue_stage_out(P, s);
}

}  // namespace sht

using sht::Error;
using sht::guard;

extern "C" {

sht_status sht_plan_create(sht_ctx* ctx, const char* expr, const sht_tensor* A, const sht_tensor* B, sht_tensor* C,
                           sht_precision prec, sht_plan** out)
{
  return guard([&] {
    SHT_REQUIRE(ctx && expr && out, SHT_ERR_INVALID, "null argument");
    sht::check_tensors(A, B, C, ctx, prec);
    SHT_REQUIRE(prec >= SHT_PREC_FP64 && prec <= SHT_PREC_BF16X3, SHT_ERR_INVALID, "unknown precision");
    auto P = std::make_unique<sht_plan>();
    P->ctx = ctx;
    P->expr = expr;
    P->A = A;
    P->B = B;
    P->C = C;
    P->prec = prec;
    // block-level analysis: the expression over block counts
    P->maps = sht::analyse_expression(P->expr, A->nblk, B->nblk, C->nblk);
    const sht::ContractMaps& m = P->maps;
    for (size_t r = 0; r < m.a_row.size(); ++r)
      SHT_REQUIRE(A->bs[m.a_row[r]] == C->bs[m.c_row[r]], SHT_ERR_INVALID, "Block sizes of A and C do not match");
    for (size_t r = 0; r < m.b_col.size(); ++r)
      SHT_REQUIRE(B->bs[m.b_col[r]] == C->bs[m.c_col[r]], SHT_ERR_INVALID, "Block sizes of B and C do not match");
    for (size_t r = 0; r < m.a_col.size(); ++r)
      SHT_REQUIRE(A->bs[m.a_col[r]] == B->bs[m.b_row[r]], SHT_ERR_INVALID, "Block sizes of A and B do not match");
    try {
      const double t0 = sht::wall_ms();
      sht::build_tasks(P.get());
      const double t1 = sht::wall_ms();
      sht::build_schedule(P.get());
      P->ms_cstat = sht::wall_ms() - t1;
      P->ms_tasks = t1 - t0;
      sht::setup_staging(P.get());
      P->schedule = sht::default_schedule(P.get());
      if (P->schedule >= 1) sht::build_as_schedule(P.get());
      else sht::build_stage32(P.get());  // (the copies serve the C-stationary tiles only)
    } catch (...) {
      P->release();
      throw;
    }
    *out = P.release();
  });
}

sht_status sht_plan_execute(sht_plan* P, double alpha, double beta, void* stream)
{
  return guard([&] {
    SHT_REQUIRE(P, SHT_ERR_INVALID, "null plan");
    SHT_REQUIRE(P->C->reserved && static_cast<int64_t>(P->C->indices.size()) == P->ncb, SHT_ERR_STATE,
                "C pattern changed since the plan was built");
    sht::enqueue(P, alpha, beta, sht::as_stream(stream));
  });
}

sht_status sht_plan_execute_timed(sht_plan* P, double alpha, double beta, void* stream, int* kclass_out,
                                  float* ms_out, int cap, int* nlaunch)
{
  return guard([&] {
    SHT_REQUIRE(P && nlaunch, SHT_ERR_INVALID, "null argument");
    const size_t n = P->schedule >= 1 ? 1 : P->launches.size();
    std::vector<hipEvent_t> ev(2 * std::max<size_t>(n, 1));
    for (auto& e : ev) SHT_HIP(hipEventCreate(&e));
    sht_status st = SHT_OK;
    std::string msg;
    try {
      hipStream_t s = sht::as_stream(stream);
      sht::enqueue(P, alpha, beta, s, &ev);
      SHT_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < n && static_cast<int>(i) < cap; ++i) {
        float ms = 0.f;
        SHT_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        if (ms_out) ms_out[i] = ms;
        if (kclass_out)
          kclass_out[i] = P->schedule >= 1 ? 100 + static_cast<int>(sht::as_class(P))
                                           : static_cast<int>(P->launches[i].k);
      }
      *nlaunch = static_cast<int>(n);
    } catch (const Error& e) {
      st = e.code;
      msg = e.what();
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (st != SHT_OK) throw Error(st, msg);
  });
}

const char* sht_kclass_name(int k)
{
  // (one name per class, in the order of the AsClass / KClass enumerations)
  static const char* as_names[] = {"as_mfma64",   "as_exact",    "as_mfma64_g",  "as_mfma32",   "as_bf16x3",
                                   "as_mfma32_g", "as_bf16x3_g", "as_mfma64_n4", "as_mfma32_n4"};
  static_assert(sizeof(as_names) / sizeof(as_names[0]) == static_cast<size_t>(sht::AsClass::Count),
                "a name per A-stationary class");
  if (k >= 100 && k < 100 + static_cast<int>(sizeof(as_names) / sizeof(as_names[0]))) return as_names[k - 100];
  static const char* names[] = {"exact_n8",         "exact_n16",        "exact_n32",        "mfma64_n16",
                                "mfma64_n32",       "mfma64_n16_r32",   "mfma32_n16_r32",   "mfma64_n64w",
                                "mfma32_n64w",      "bf16x3_n64w",      "mfma64_n32s",      "mfma32_n32s",
                                "bf16x3_n32s",      "mfma64_n16_r128",  "mfma64_n32_r128",  "mfma64_n16w",
                                "mfma32_n16w",      "bf16x3_n16w",      "mfma64_n32w",      "mfma32_n32w",
                                "mfma32_n16w_r128", "mfma32_n32w_r128", "mfma64_persist",   "mfma32_persist",
                                "bf16x3_persist"};
  static_assert(sizeof(names) / sizeof(names[0]) == static_cast<size_t>(sht::KClass::Count),
                "a name per C-stationary class");
  if (k < 0 || k >= static_cast<int>(sizeof(names) / sizeof(names[0]))) return "unknown";
  return names[k];
}

sht_status sht_plan_graph_replay(sht_plan* P, double alpha, double beta, int replays, void* stream)
{
  return guard([&] {
    SHT_REQUIRE(P && replays >= 0, SHT_ERR_INVALID, "bad argument");
    hipStream_t s = sht::as_stream(stream);
    SHT_REQUIRE(s != nullptr, SHT_ERR_INVALID, "graph replay needs a created stream");
    if (!P->graph || P->graph_alpha != alpha || P->graph_beta != beta) {
      if (P->graph) SHT_HIP(hipGraphExecDestroy(P->graph));
      if (P->graph_def) SHT_HIP(hipGraphDestroy(P->graph_def));
      P->graph = nullptr;
      P->graph_def = nullptr;
      // the side streams of concurrent C-stationary classes exist before the
      // capture starts (no object creation inside a capture)
      if (P->schedule == 0 && sht::cs_streams(P) > 1) sht::ensure_side_streams(P, sht::cs_streams(P) - 1);
      SHT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        sht::enqueue(P, alpha, beta, s);
      } catch (...) {
        hipGraph_t g;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      SHT_HIP(hipStreamEndCapture(s, &P->graph_def));
      SHT_HIP(hipGraphInstantiate(&P->graph, P->graph_def, nullptr, nullptr, 0));
      P->graph_alpha = alpha;
      P->graph_beta = beta;
    }
    for (int r = 0; r < replays; ++r) SHT_HIP(hipGraphLaunch(P->graph, s));
  });
}

sht_status sht_plan_set_schedule(sht_plan* P, int schedule)
{
  return guard([&] {
    SHT_REQUIRE(P && schedule >= 0 && schedule <= 3, SHT_ERR_INVALID, "schedule must be 0, 1, 2 or 3");
    if (schedule >= 1) {
      SHT_REQUIRE(sht::as_supported(P), SHT_ERR_UNSUPPORTED, "A-stationary schedule not available");
      SHT_REQUIRE(schedule != 3 || sht::narrow_supported(P), SHT_ERR_UNSUPPORTED,
                  "narrow A-stationary schedule not available for this precision");
      sht::build_as_schedule(P);
    }
    if (P->schedule != schedule && P->graph) {
      SHT_HIP(hipGraphExecDestroy(P->graph));
      SHT_HIP(hipGraphDestroy(P->graph_def));
      P->graph = nullptr;
      P->graph_def = nullptr;
    }
    P->schedule = schedule;
  });
}

sht_status sht_plan_get_schedule(const sht_plan* P, int* schedule)
{
  return guard([&] {
    SHT_REQUIRE(P && schedule, SHT_ERR_INVALID, "null argument");
    *schedule = P->schedule;
  });
}

namespace {

// Release a captured graph (its launches depend on the schedule and tiles).
void drop_graph(sht_plan* P)
{
  if (P->graph) SHT_HIP(hipGraphExecDestroy(P->graph));
  if (P->graph_def) SHT_HIP(hipGraphDestroy(P->graph_def));
  P->graph = nullptr;
  P->graph_def = nullptr;
}

// Every (schedule, C-stationary variant) the plan can run, timed with `reps`
// executions after one warm-up; the fastest is kept.
void autotune(sht_plan* P, double alpha, double beta, int reps, hipStream_t s, std::vector<int>& sched,
              std::vector<int>& variant, std::vector<float>& ms)
{
  std::vector<std::pair<int, int>> cand;
  const bool quad_ok = P->prec == SHT_PREC_FP64 || P->prec == SHT_PREC_FP32;
  for (int v = 0; v <= 4; ++v)
    if (quad_ok || (v != 1 && v != 2)) cand.push_back({0, v});
  if (P->prec != SHT_PREC_FP64_EXACT) cand.push_back({0, 5});
  if (sht::as_supported(P)) {
    cand.push_back({1, 0});
    // the fma-chain form has a single A-stationary kernel
    if (P->prec != SHT_PREC_FP64_EXACT) cand.push_back({2, 0});
    if (sht::narrow_supported(P)) cand.push_back({3, 0});
  }
  drop_graph(P);
  hipEvent_t e0, e1;
  SHT_HIP(hipEventCreate(&e0));
  SHT_HIP(hipEventCreate(&e1));
  int best_s = P->schedule, best_v = P->cs_variant;
  float best_ms = 1e30f;
  try {
    for (const auto& c : cand) {
      if (c.first >= 1) sht::build_as_schedule(P);
      if (c.first == 0 && P->cs_variant != c.second) {
        P->cs_variant = c.second;
        sht::build_tiles(P);
      }
      P->schedule = c.first;
      sht::enqueue(P, alpha, beta, s);
      SHT_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; ++r) sht::enqueue(P, alpha, beta, s);
      SHT_HIP(hipEventRecord(e1, s));
      SHT_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      SHT_HIP(hipEventElapsedTime(&t, e0, e1));
      t /= reps;
      sched.push_back(c.first);
      variant.push_back(c.second);
      ms.push_back(t);
      if (t < best_ms) {
        best_ms = t;
        best_s = c.first;
        best_v = c.second;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (P->cs_variant != best_v) {
    P->cs_variant = best_v;
    sht::build_tiles(P);
  }
  P->schedule = best_s;
}

}  // namespace

// Time every applicable schedule and C-stationary tile variant (reps
// executions each, after one warm-up, on `stream`) and keep the fastest.
// C is overwritten.
sht_status sht_plan_autotune(sht_plan* P, double alpha, double beta, int reps, void* stream, float* ms_out)
{
  return guard([&] {
    SHT_REQUIRE(P && reps > 0, SHT_ERR_INVALID, "bad argument");
    std::vector<int> sc, va;
    std::vector<float> ms;
    autotune(P, alpha, beta, reps, sht::as_stream(stream), sc, va, ms);
    if (ms_out) {
      for (int c = 0; c < 4; ++c) ms_out[c] = -1.f;
      for (size_t i = 0; i < ms.size(); ++i)
        if (ms_out[sc[i]] < 0.f || ms[i] < ms_out[sc[i]]) ms_out[sc[i]] = ms[i];
    }
  });
}

sht_status sht_plan_autotune_ex(sht_plan* P, double alpha, double beta, int reps, void* stream, int cap, int* n_out,
                                int* sched_out, int* variant_out, float* ms_out)
{
  return guard([&] {
    SHT_REQUIRE(P && reps > 0 && n_out, SHT_ERR_INVALID, "bad argument");
    std::vector<int> sc, va;
    std::vector<float> ms;
    autotune(P, alpha, beta, reps, sht::as_stream(stream), sc, va, ms);
    *n_out = static_cast<int>(ms.size());
    for (int i = 0; i < cap && i < static_cast<int>(ms.size()); ++i) {
      if (sched_out) sched_out[i] = sc[i];
      if (variant_out) variant_out[i] = va[i];
      if (ms_out) ms_out[i] = ms[i];
    }
  });
}

sht_status sht_plan_set_variant(sht_plan* P, int variant)
{
  return guard([&] {
    SHT_REQUIRE(P && variant >= 0 && variant <= 5, SHT_ERR_INVALID, "C-stationary variant must be 0..5");
    SHT_REQUIRE(variant != 5 || P->prec != SHT_PREC_FP64_EXACT, SHT_ERR_INVALID,
                "the persistent variant needs the FP64, FP32 or BF16X3 form");
    if (variant != P->cs_variant) {
      drop_graph(P);
      P->cs_variant = variant;
      sht::build_tiles(P);
    }
  });
}

sht_status sht_plan_get_variant(const sht_plan* P, int* variant)
{
  return guard([&] {
    SHT_REQUIRE(P && variant, SHT_ERR_INVALID, "null argument");
    *variant = P->cs_variant;
  });
}

sht_status sht_plan_set_fp32_staging(sht_plan* P, int on)
{
  return guard([&] {
    SHT_REQUIRE(P && (on == 0 || on == 1), SHT_ERR_INVALID, "staging must be 0 or 1");
    SHT_REQUIRE(!on || ((P->prec == SHT_PREC_FP32 || P->prec == SHT_PREC_BF16X3) && P->A->dtype == SHT_FLOAT64),
                SHT_ERR_INVALID, "FP32 staging needs FLOAT64 operands and the FP32 or BF16X3 form");
    if ((on != 0) == P->stage32) return;
    drop_graph(P);
    if (!on) sht::drop_stage32(P);
    else SHT_REQUIRE(sht::make_stage32(P), SHT_ERR_OOM, "Out of memory");
  });
}

sht_status sht_plan_get_fp32_staging(const sht_plan* P, int* on)
{
  return guard([&] {
    SHT_REQUIRE(P && on, SHT_ERR_INVALID, "null argument");
    *on = P->stage32 ? 1 : 0;
  });
}

const char* sht_cs_variant_name(int v)
{
  static const char* names[] = {"auto", "no_128row", "all_128row", "cost_order", "uniform_64x64", "persistent"};
  return v >= 0 && v <= 5 ? names[v] : "unknown";
}

sht_status sht_plan_tasks(sht_plan* P, const int64_t** ta, const int64_t** tb, const int64_t** tc,
                          const int64_t** kkey, int64_t* ntasks)
{
  return guard([&] {
    SHT_REQUIRE(P, SHT_ERR_INVALID, "null plan");
    if (!P->host_tasks) {
      const int64_t n = P->ntasks;
      P->h_ta.resize(n);
      P->h_tb.resize(n);
      P->h_tc.resize(n);
      P->h_kk.resize(n);
      if (n) {
        SHT_HIP(hipMemcpy(P->h_ta.data(), P->d_ta, n * 8, hipMemcpyDeviceToHost));
        SHT_HIP(hipMemcpy(P->h_tb.data(), P->d_tb, n * 8, hipMemcpyDeviceToHost));
        SHT_HIP(hipMemcpy(P->h_tc.data(), P->d_tc, n * 8, hipMemcpyDeviceToHost));
        SHT_HIP(hipMemcpy(P->h_kk.data(), P->d_kk, n * 8, hipMemcpyDeviceToHost));
      }
      P->host_tasks = true;
    }
    if (ta) *ta = P->h_ta.data();
    if (tb) *tb = P->h_tb.data();
    if (tc) *tc = P->h_tc.data();
    if (kkey) *kkey = P->h_kk.data();
    if (ntasks) *ntasks = P->ntasks;
  });
}

sht_status sht_plan_stats(const sht_plan* P, double* flops, double* bytes, int64_t* ntasks, int64_t* ncblocks)
{
  return guard([&] {
    SHT_REQUIRE(P, SHT_ERR_INVALID, "null plan");
    if (flops) *flops = P->flops;
    if (bytes) *bytes = P->bytes_in;
    if (ntasks) *ntasks = P->ntasks;
    if (ncblocks) *ncblocks = P->ncb;
  });
}

sht_status sht_plan_build_ms(const sht_plan* P, double* tasks_ms, double* cstat_ms, double* astat_ms)
{
  return guard([&] {
    SHT_REQUIRE(P, SHT_ERR_INVALID, "null plan");
    if (tasks_ms) *tasks_ms = P->ms_tasks;
    if (cstat_ms) *cstat_ms = P->ms_cstat;
    if (astat_ms) *astat_ms = P->as_built ? P->ms_astat : -1.0;
  });
}

sht_status sht_plan_destroy(sht_plan* P)
{
  return guard([&] {
    if (!P) return;
    P->release();
    delete P;
  });
}

sht_status sht_contract(sht_ctx* ctx, const char* expr, double alpha, const sht_tensor* A, const sht_tensor* B,
                        double beta, sht_tensor* C, sht_precision prec, void* stream)
{
  return guard([&] {
    sht_plan* P = nullptr;
    sht_status st = sht_plan_create(ctx, expr, A, B, C, prec, &P);
    if (st != SHT_OK) throw Error(st, sht_last_error());
    st = sht_plan_execute(P, alpha, beta, stream);
    if (st == SHT_OK) {
      hipError_t e = hipStreamSynchronize(sht::as_stream(stream));
      if (e != hipSuccess) {
        sht_plan_destroy(P);
        throw Error(SHT_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
      }
    }
    const std::string msg = st == SHT_OK ? "" : sht_last_error();
    sht_plan_destroy(P);
    if (st != SHT_OK) throw Error(st, msg);
  });
}

}  // extern "C"
