// Serving front-end: continuous batching of contraction requests.
//
// The reference has no serving layer; this is the request loop built around
// the contraction plan.  Requests name operand tensors that already live in
// the context's HBM arena.  A dispatcher thread
//   1. admits queued requests in arrival order while the summed memory
//      footprint of admitted-but-unfinished requests stays within the budget
//      (a request larger than the whole budget is rejected at submit);
//   2. builds each admitted request's plan (GPU task list + C pattern);
//   3. merges the C-stationary tile lists of every ready request with the
//      same (precision, alpha, beta) into one grouped launch per kernel
//      class: task / C-block records are copied with their element offsets
//      rebased to the arena base, so one launch walks tiles of many
//      contractions of different orders, shapes and sparsities;
//   4. records an event per batch and completes requests when it fires,
//      while the next batch is admitted, planned and launched.
// A request's C blocks are computed by exactly the kernels and task order of
// a stand-alone execution, so its result is bit-identical alone or in a mix.

#include <condition_variable>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "plan.hpp"

namespace sht {

namespace {

__global__ void k_rebase_tasks(TaskRec* dst, const TaskRec* src, int64_t n, int64_t da, int64_t db)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TaskRec r = src[i];
  r.a_off += da;
  r.b_off += db;
  dst[i] = r;
}

__global__ void k_rebase_cblk(CBlkRec* dst, const CBlkRec* src, int64_t n, int64_t dc, int32_t dt)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  CBlkRec r = src[i];
  r.c_off += dc;
  r.t0 += dt;
  r.t1 += dt;
  dst[i] = r;
}

__global__ void k_rebase_tiles(TileRec* dst, const TileRec* src, int64_t n, int32_t dcb)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TileRec r = src[i];
  r.cb += dcb;
  dst[i] = r;
}

inline unsigned nblk(int64_t n) { return static_cast<unsigned>((std::max<int64_t>(n, 1) + 255) / 256); }

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int64_t id = 0;
  std::string expr;
  double alpha = 1.0, beta = 0.0;
  const sht_tensor* A = nullptr;
  const sht_tensor* B = nullptr;
  sht_tensor* C = nullptr;
  sht_precision prec = SHT_PREC_FP64;
  size_t footprint = 0;
  double t_submit = 0, t_done = 0;
  sht_plan* plan = nullptr;
  bool done = false;
  sht_status status = SHT_OK;
  std::string err;
};

struct Batch {
  std::vector<Request*> reqs;
  hipEvent_t ev = nullptr;
  std::vector<void*> scratch;  // merged record arrays (arena)
};

size_t tensor_bytes(const sht_tensor* t) { return static_cast<size_t>(std::max<int64_t>(t->nelem, 0)) * 8; }

}  // namespace

}  // namespace sht

struct sht_server {
  sht_ctx* ctx = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  int max_batch = 64;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<sht::Request*> pending;
  std::map<int64_t, sht::Request*> reqs;
  std::deque<sht::Batch> inflight;
  int64_t next_id = 1;
  bool stop = false;
  bool head_blocked = false;  // queue head waits for memory to be released
  std::thread worker;
  hipStream_t stream = nullptr;
  // statistics
  int64_t nbatches = 0, nreq_done = 0, max_merged = 0, nlaunches = 0;
};

namespace sht {

namespace {

// Launch one merged batch of planned requests (all share prec/alpha/beta).
void launch_group(sht_server* S, const std::vector<Request*>& g, Batch& b)
{
  Arena& arena = S->ctx->arena;
  const char* base = static_cast<const char*>(S->ctx->d_base);
  auto rel = [base](const void* p) { return static_cast<int64_t>((static_cast<const char*>(p) - base) / 8); };
  int64_t nt = 0, ncb = 0;
  std::vector<int64_t> ntiles(static_cast<int>(KClass::Count), 0);
  for (Request* r : g) {
    nt += r->plan->ntasks;
    ncb += r->plan->ncb;
    for (const auto& L : r->plan->launches) ntiles[static_cast<int>(L.k)] += L.count;
  }
  TaskRec* tasks = static_cast<TaskRec*>(arena.allocate(std::max<int64_t>(nt, 1) * sizeof(TaskRec)));
  b.scratch.push_back(tasks);
  CBlkRec* cblk = static_cast<CBlkRec*>(arena.allocate(std::max<int64_t>(ncb, 1) * sizeof(CBlkRec)));
  b.scratch.push_back(cblk);
  std::vector<TileRec*> tiles(ntiles.size(), nullptr);
  for (size_t k = 0; k < ntiles.size(); ++k) {
    if (!ntiles[k]) continue;
    tiles[k] = static_cast<TileRec*>(arena.allocate(ntiles[k] * sizeof(TileRec)));
    b.scratch.push_back(tiles[k]);
  }
  std::vector<int64_t> tfill(ntiles.size(), 0);
  int64_t t0 = 0, c0 = 0;
  const Request* first = g.front();
  for (Request* r : g) {
    sht_plan* P = r->plan;
    enqueue_stage_in(P, first->alpha, first->beta, S->stream);
    const void* Ad = P->permA ? P->d_Ap : P->A->d_data;
    const void* Bd = P->permB ? P->d_Bp : P->B->d_data;
    const void* Cd = P->permC ? P->d_Cp : P->C->d_data;
    if (P->ntasks)
      hipLaunchKernelGGL(k_rebase_tasks, dim3(nblk(P->ntasks)), dim3(256), 0, S->stream, tasks + t0, P->d_tasks,
                         P->ntasks, rel(Ad), rel(Bd));
    if (P->ncb)
      hipLaunchKernelGGL(k_rebase_cblk, dim3(nblk(P->ncb)), dim3(256), 0, S->stream, cblk + c0, P->d_cblk, P->ncb,
                         rel(Cd), static_cast<int32_t>(t0));
    for (const auto& L : P->launches) {
      const int k = static_cast<int>(L.k);
      hipLaunchKernelGGL(k_rebase_tiles, dim3(nblk(L.count)), dim3(256), 0, S->stream, tiles[k] + tfill[k],
                         P->d_tiles + L.first, L.count, static_cast<int32_t>(c0));
      tfill[k] += L.count;
    }
    SHT_HIP(hipGetLastError());
    t0 += P->ntasks;
    c0 += P->ncb;
  }
  SHT_REQUIRE(t0 < (int64_t(1) << 31) && c0 < (int64_t(1) << 31), SHT_ERR_UNSUPPORTED, "batch too large");
  ExecArgs a{};
  a.A = base;
  a.B = base;
  a.C = const_cast<char*>(base);
  a.cblk = cblk;
  a.tasks = tasks;
  a.alpha = first->alpha;
  a.beta = first->beta;
  uint64_t ua, ub;
  std::memcpy(&ua, &first->alpha, 8);
  std::memcpy(&ub, &first->beta, 8);
  a.alpha_zero = ua == 0;
  a.beta_zero = ub == 0;
  a.alpha_one = first->alpha == 1.0;
  for (size_t k = 0; k < ntiles.size(); ++k) {
    if (!ntiles[k]) continue;
    a.tiles = tiles[k];
    a.ntiles = ntiles[k];
    launch_exec(static_cast<KClass>(k), a, S->stream);
    S->nlaunches++;
  }
  for (Request* r : g) enqueue_stage_out(r->plan, S->stream);
}

void finish(sht_server* S, Batch& b)
{
  const double t = now_ms();
  for (void* p : b.scratch) S->ctx->arena.release(p);
  if (b.ev) (void)hipEventDestroy(b.ev);
  std::lock_guard<std::mutex> lk(S->mu);
  for (Request* r : b.reqs) {
    if (r->plan) {
      sht_plan_destroy(r->plan);
      r->plan = nullptr;
    }
    r->t_done = t;
    r->done = true;
    S->in_use -= r->footprint;
    S->nreq_done++;
  }
  S->head_blocked = false;
  S->cv_done.notify_all();
  S->cv_work.notify_all();
}

void worker_loop(sht_server* S)
{
  // (the context's device is already current in the creating thread; a
  // failure here surfaces as the error of the first batch's HIP calls)
  (void)hipSetDevice(S->ctx->device);
  for (;;) {
    std::vector<Request*> admitted;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait_for(lk, std::chrono::microseconds(200),
                          [S] { return S->stop || (!S->pending.empty() && !S->head_blocked); });
      if (S->stop && S->pending.empty() && S->inflight.empty()) return;
      while (!S->pending.empty() && static_cast<int>(admitted.size()) < S->max_batch) {
        Request* r = S->pending.front();
        if (S->in_use + r->footprint > S->budget && S->in_use > 0) {
          S->head_blocked = true;  // wait for memory
          break;
        }
        S->pending.pop_front();
        S->in_use += r->footprint;
        admitted.push_back(r);
      }
    }
    // retire finished batches
    while (!S->inflight.empty() && hipEventQuery(S->inflight.front().ev) == hipSuccess) {
      finish(S, S->inflight.front());
      S->inflight.pop_front();
    }
    if (admitted.empty()) continue;
    // plan every admitted request
    std::vector<Request*> ready;
    Batch failed;
    for (Request* r : admitted) {
      sht_plan* P = nullptr;
      const sht_status st = sht_plan_create(S->ctx, r->expr.c_str(), r->A, r->B, r->C, r->prec, &P);
      if (st != SHT_OK) {
        r->status = st;
        r->err = sht_last_error();
        failed.reqs.push_back(r);
        continue;
      }
      P->schedule = 0;  // merged launches use the C-stationary tile lists
      r->plan = P;
      ready.push_back(r);
    }
    if (!failed.reqs.empty()) finish(S, failed);
    // group by (precision, alpha, beta) and launch
    std::map<std::tuple<int, uint64_t, uint64_t>, std::vector<Request*>> groups;
    for (Request* r : ready) {
      uint64_t ua, ub;
      std::memcpy(&ua, &r->alpha, 8);
      std::memcpy(&ub, &r->beta, 8);
      groups[std::make_tuple(static_cast<int>(r->prec), ua, ub)].push_back(r);
    }
    for (auto& kv : groups) {
      Batch b;
      b.reqs = kv.second;
      try {
        launch_group(S, kv.second, b);
        SHT_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
        SHT_HIP(hipEventRecord(b.ev, S->stream));
      } catch (const Error& e) {
        (void)hipStreamSynchronize(S->stream);
        for (Request* r : b.reqs) {
          r->status = e.code;
          r->err = e.what();
        }
        if (b.ev) {
          (void)hipEventDestroy(b.ev);
          b.ev = nullptr;
        }
        finish(S, b);
        continue;
      }
      S->nbatches++;
      S->max_merged = std::max<int64_t>(S->max_merged, static_cast<int64_t>(b.reqs.size()));
      S->inflight.push_back(std::move(b));
    }
  }
}

}  // namespace

}  // namespace sht

using sht::Error;
using sht::guard;

extern "C" {

sht_status sht_server_create(sht_ctx* ctx, size_t budget_bytes, int max_batch, sht_server** out)
{
  return guard([&] {
    SHT_REQUIRE(ctx && out && !ctx->host, SHT_ERR_INVALID, "server needs a device context");
    auto S = std::make_unique<sht_server>();
    S->ctx = ctx;
    S->budget = budget_bytes;
    S->max_batch = max_batch > 0 ? max_batch : 64;
    SHT_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    sht_server* raw = S.get();
    S->worker = std::thread([raw] { sht::worker_loop(raw); });
    *out = S.release();
  });
}

sht_status sht_server_submit(sht_server* S, const char* expr, double alpha, const sht_tensor* A, const sht_tensor* B,
                             double beta, sht_tensor* C, sht_precision prec, int64_t* req_id)
{
  return guard([&] {
    SHT_REQUIRE(S && expr && A && B && C && req_id, SHT_ERR_INVALID, "null argument");
    auto r = std::make_unique<sht::Request>();
    r->expr = expr;
    r->alpha = alpha;
    r->beta = beta;
    r->A = A;
    r->B = B;
    r->C = C;
    r->prec = prec;
    // footprint: operands, C (when already reserved) and the plan's task
    // records (bounded by the pair count estimate nnzA * nnzB / |K| ~ counted
    // at plan time; here A + B + C is the admitted amount)
    r->footprint = sht::tensor_bytes(A) + sht::tensor_bytes(B) + (C->reserved ? sht::tensor_bytes(C) : 0);
    if (r->footprint > S->budget)
      throw Error(SHT_ERR_ADMISSION, "request footprint " + std::to_string(r->footprint) +
                                         " bytes exceeds the admission budget of " + std::to_string(S->budget));
    r->t_submit = sht::now_ms();
    std::lock_guard<std::mutex> lk(S->mu);
    sht::Request* raw = r.release();
    raw->id = S->next_id++;
    *req_id = raw->id;
    S->reqs[raw->id] = raw;
    S->pending.push_back(raw);
    S->cv_work.notify_all();
  });
}

sht_status sht_server_wait(sht_server* S, int64_t req_id, double* latency_ms)
{
  return guard([&] {
    SHT_REQUIRE(S, SHT_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    SHT_REQUIRE(it != S->reqs.end(), SHT_ERR_INVALID, "unknown request id");
    sht::Request* r = it->second;
    S->cv_done.wait(lk, [r] { return r->done; });
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    const sht_status st = r->status;
    const std::string err = r->err;
    S->reqs.erase(it);
    delete r;
    if (st != SHT_OK) throw Error(st, err);
  });
}

sht_status sht_server_poll(sht_server* S, int64_t req_id, int* done)
{
  return guard([&] {
    SHT_REQUIRE(S && done, SHT_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    SHT_REQUIRE(it != S->reqs.end(), SHT_ERR_INVALID, "unknown request id");
    *done = it->second->done ? 1 : 0;
  });
}

sht_status sht_server_stats(sht_server* S, int64_t* nreq_done, int64_t* nbatches, int64_t* max_merged,
                            int64_t* nlaunches)
{
  return guard([&] {
    SHT_REQUIRE(S, SHT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (nreq_done) *nreq_done = S->nreq_done;
    if (nbatches) *nbatches = S->nbatches;
    if (max_merged) *max_merged = S->max_merged;
    if (nlaunches) *nlaunches = S->nlaunches;
  });
}

sht_status sht_server_destroy(sht_server* S)
{
  return guard([&] {
    if (!S) return;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
      S->cv_work.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    for (auto& kv : S->reqs) delete kv.second;
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
  });
}

}  // extern "C"
