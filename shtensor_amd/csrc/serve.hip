// Serving front-end: continuous batching of contraction requests.
//
// The reference has no serving layer; this is the request loop built around
// the contraction plan.  Requests name operand tensors that already live in
// the context's HBM arena.
//   * Admission (dispatcher thread, in arrival order): a request is admitted
//     while the summed memory footprint of admitted-but-unfinished requests
//     stays within the budget (a request larger than the whole budget is
//     rejected at submit).  A request that touches a tensor an earlier
//     unfinished request writes (or writes one it reads) waits for it; later
//     independent requests pass it, so dependent requests keep their
//     submission order and independent ones never share a race.
//   * Planning (planner threads, off the dispatcher): each admitted request's
//     plan (GPU task list + C pattern + schedule).  Its footprint is then
//     corrected to the memory the plan and C really hold; a plan that runs
//     out of arena while other requests hold memory goes back to the queue.
//   * Launch: ready requests whose plan runs C-stationary tiles are merged
//     per (precision, alpha, beta) into one grouped launch per kernel class
//     (task / C-block records rebased to the arena base, so one launch walks
//     tiles of many contractions of different orders, shapes and
//     sparsities); requests whose plans picked the grouped A-stationary
//     FP64/FP32 schedule are merged the same way into one launch of that
//     kernel (sweep, column and unit records rebased); other A-stationary
//     plans run their own launch.  Batches go round-robin to kServeStreams streams, so a
//     batch admitted while others run overlaps them on the device.
//   * Completion: an event per batch; requests complete when it fires.
//   * Cancellation: a request not yet launched can be cancelled.
// A request's C blocks are computed by exactly the kernels and task order of
// a stand-alone execution, so its result is bit-identical alone or in a mix.

#include <algorithm>
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

// (a tile record carries its C block's record and its first task's operand
// offsets: rebased like those records)
__global__ void k_rebase_tiles(TileRec* dst, const TileRec* src, int64_t n, int32_t dcb, int64_t da, int64_t db,
                               int64_t dc, int32_t dt)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TileRec r = src[i];
  if (!r.pad) {
    r.cb += dcb;
    r.c_off += dc;
    r.t0 += dt;
    r.t1 += dt;
    r.a0 += da;
    r.b0 += db;
  }
  dst[i] = r;
}

// A-stationary records of one request moved into a merged batch: element
// offsets rebased to the arena base (A and C in doubles, the panel in its
// own element type), sweep column entries to the batch's column array.
__global__ void k_rebase_sweeps(AsSweep* dst, const AsSweep* src, int64_t n, int64_t da, int64_t db, int64_t dc0)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  AsSweep r = src[i];
  r.a_off += da;
  r.b_off += db;
  r.c0 += dc0;
  dst[i] = r;
}

// Column entries: an offset v becomes v + dc, -1 stays; a first-touch
// marker -(v + 2) keeps its meaning (rebased) when the request's first
// sweeps start from zeros, and becomes the plain offset otherwise (its C was
// scaled by beta before the launch and is read like any other column).
__global__ void k_rebase_ccol(int64_t* dst, const int64_t* src, int64_t n, int64_t dc, int keep_first)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t e = src[i];
  dst[i] = e >= 0 ? e + dc : (e == -1 ? e : (keep_first ? -((-e - 2) + dc) - 2 : (-e - 2) + dc));
}

__global__ void k_rebase_units(AsUnit* dst, const AsUnit* src, int64_t n, int64_t npad, int32_t dsw)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  AsUnit r{0, 0, 0, 0};  // (past the request's units: empty filler units)
  if (i < n) {
    r = src[i];
    r.sweep0 += dsw;
  }
  dst[i] = r;
}

inline unsigned nblk(int64_t n) { return static_cast<unsigned>((std::max<int64_t>(n, 1) + 255) / 256); }

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum class RState { Pending, Planning, Ready, Launched, Done };

struct Request {
  int64_t id = 0;
  std::string expr;
  double alpha = 1.0, beta = 0.0;
  const sht_tensor* A = nullptr;
  const sht_tensor* B = nullptr;
  sht_tensor* C = nullptr;
  sht_precision prec = SHT_PREC_FP64;
  size_t footprint = 0;  // bytes counted against the budget while admitted
  double t_submit = 0, t_ready = 0, t_done = 0;
  sht_plan* plan = nullptr;
  RState state = RState::Pending;
  bool cancel = false;
  sht_status status = SHT_OK;
  std::string err;
};

struct Batch {
  std::vector<Request*> reqs;
  hipEvent_t ev = nullptr;
  std::vector<void*> scratch;  // merged record arrays (arena)
};

constexpr int kServeStreams = 4;  // device streams batches are dealt to
constexpr int kPlanners = 2;      // planner threads
// Ready requests wait for the ones still being planned (to share launches)
// at most this long.
constexpr double kBatchWindowMs = 2.0;

// (operands may be FP32 tensors for the FP32 and split-BF16 forms)
size_t tensor_bytes(const sht_tensor* t) { return static_cast<size_t>(std::max<int64_t>(t->nelem, 0)) * t->elem_size(); }

// Arena bytes a plan holds (task list, records, tiles, schedules, staging).
size_t plan_bytes(const sht_plan* P)
{
  const Arena& a = P->ctx->arena;
  size_t b = 0;
  for (const void* p : {static_cast<const void*>(P->d_ta), static_cast<const void*>(P->d_tb),
                        static_cast<const void*>(P->d_tc), static_cast<const void*>(P->d_kk),
                        static_cast<const void*>(P->d_tasks), static_cast<const void*>(P->d_cblk),
                        static_cast<const void*>(P->d_tiles), static_cast<const void*>(P->d_Ap),
                        static_cast<const void*>(P->d_Bp), static_cast<const void*>(P->d_Cp),
                        static_cast<const void*>(P->d_as_sweep), static_cast<const void*>(P->d_as_ccol),
                        static_cast<const void*>(P->d_as_unit), static_cast<const void*>(P->d_as_unitw),
                        static_cast<const void*>(P->d_as_sweep4), static_cast<const void*>(P->d_as_unit4),
                        static_cast<const void*>(P->d_bt_rec), static_cast<const void*>(P->d_bt),
                        static_cast<const void*>(P->d_as_zero)})
    if (p) b += a.size_of(p);
  return b;
}

// r must wait for q (or for a request skipped ahead of it): r writes a
// tensor q touches, or reads one q writes.
bool depends(const Request* r, const Request* q)
{
  return r->C == q->C || static_cast<const sht_tensor*>(r->C) == q->A ||
         static_cast<const sht_tensor*>(r->C) == q->B || r->A == q->C || r->B == q->C;
}

}  // namespace

}  // namespace sht

struct sht_server {
  sht_ctx* ctx = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  int max_batch = 64;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_plan;
  std::deque<sht::Request*> pending;  // submitted, not admitted
  std::deque<sht::Request*> to_plan;  // admitted, waiting for a planner
  std::vector<sht::Request*> ready;   // planned (or failed / cancelled), not launched
  std::vector<sht::Request*> active;  // admitted and unfinished
  std::map<int64_t, sht::Request*> reqs;
  std::deque<sht::Batch> inflight;    // dispatcher thread only
  int64_t next_id = 1;
  bool stop = false;
  bool stop_planners = false;
  int planning = 0;                   // requests a planner holds
  std::thread dispatcher;
  std::vector<std::thread> planners;
  hipStream_t streams[sht::kServeStreams] = {};
  int next_stream = 0;
  // statistics
  int64_t nbatches = 0, nreq_done = 0, max_merged = 0, nlaunches = 0, ncancelled = 0, nrequeued = 0;
};

namespace sht {

namespace {

// Launch one merged batch of planned C-stationary requests (all share
// prec/alpha/beta) on stream s.
void launch_group(sht_server* S, const std::vector<Request*>& g, Batch& b, hipStream_t s)
{
  Arena& arena = S->ctx->arena;
  const char* base = static_cast<const char*>(S->ctx->d_base);
  // element offsets from the arena base (arena blocks are 256-byte aligned,
  // so FP32 operands have whole FP32 offsets)
  auto rel = [base](const void* p, int64_t w = 8) {
    return static_cast<int64_t>((static_cast<const char*>(p) - base) / w);
  };
  const int64_t ew = g.front()->A->dtype == SHT_FLOAT32 ? 4 : 8;  // operand element bytes (the group shares them)
  int64_t nt = 0, ncb = 0;
  std::vector<int64_t> ntiles(static_cast<int>(KClass::Count), 0);
  for (Request* r : g) {
    nt += r->plan->ntasks;
    ncb += r->plan->ncb;
    for (const auto& L : r->plan->launches) ntiles[static_cast<int>(L.k)] += L.count;
  }
  SHT_REQUIRE(nt < (int64_t(1) << 31) && ncb < (int64_t(1) << 31), SHT_ERR_UNSUPPORTED, "batch too large");
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
    enqueue_stage_in(P, first->alpha, first->beta, s);
    const void* Ad = P->permA ? P->d_Ap : P->A->d_data;
    const void* Bd = P->permB ? P->d_Bp : P->B->d_data;
    const void* Cd = P->permC ? P->d_Cp : P->C->d_data;
    if (P->ntasks)
      hipLaunchKernelGGL(k_rebase_tasks, dim3(nblk(P->ntasks)), dim3(256), 0, s, tasks + t0, P->d_tasks, P->ntasks,
                         rel(Ad, ew), rel(Bd, ew));
    if (P->ncb)
      hipLaunchKernelGGL(k_rebase_cblk, dim3(nblk(P->ncb)), dim3(256), 0, s, cblk + c0, P->d_cblk, P->ncb, rel(Cd),
                         static_cast<int32_t>(t0));
    for (const auto& L : P->launches) {
      const int k = static_cast<int>(L.k);
      hipLaunchKernelGGL(k_rebase_tiles, dim3(nblk(L.count)), dim3(256), 0, s, tiles[k] + tfill[k],
                         P->d_tiles + L.first, L.count, static_cast<int32_t>(c0), rel(Ad, ew), rel(Bd, ew), rel(Cd),
                         static_cast<int32_t>(t0));
      tfill[k] += L.count;
    }
    SHT_HIP(hipGetLastError());
    t0 += P->ntasks;
    c0 += P->ncb;
  }
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
  a.f32 = ew == 4 ? 1 : 0;
  for (size_t k = 0; k < ntiles.size(); ++k) {
    if (!ntiles[k]) continue;
    a.tiles = tiles[k];
    a.ntiles = ntiles[k];
    launch_exec(static_cast<KClass>(k), a, s);
    S->nlaunches++;
  }
  for (Request* r : g) enqueue_stage_out(r->plan, s);
}

// Requests whose plans run the grouped A-stationary FP64 or FP32 kernel can
// share one launch.
bool as_mergeable(const Request* r)
{
  const sht_plan* P = r->plan;
  return P->schedule == 2 && (r->prec == SHT_PREC_FP64 || r->prec == SHT_PREC_FP32);
}

// Launch one merged batch of grouped A-stationary requests (all share
// prec/alpha/beta) on stream s: each request's C = beta * C pass (unless its
// first sweeps start from zeros) and panel conversion, then one launch over
// the concatenated unit lists (each request's list padded to whole rounds of
// the eight XCD queues, so its XCD-aware order is kept).
void launch_group_as(sht_server* S, const std::vector<Request*>& g, Batch& b, hipStream_t s)
{
  Arena& arena = S->ctx->arena;
  const char* base = static_cast<const char*>(S->ctx->d_base);
  const Request* first = g.front();
  const bool f32 = first->prec == SHT_PREC_FP32;
  const int64_t tb = f32 ? 4 : 8;  // panel element bytes
  const bool op32 = first->A->dtype == SHT_FLOAT32;  // FP32 operands (the group shares the type)
  auto rel = [base](const void* p, int64_t w) {
    return static_cast<int64_t>((static_cast<const char*>(p) - base) / w);
  };
  const AsClass cls = as_class(first->plan);
  const int64_t w = as_group_waves(cls);
  const int64_t round = 8 * w;  // units per round of the eight XCD queues
  uint64_t ua, ub;
  std::memcpy(&ua, &first->alpha, 8);
  std::memcpy(&ub, &first->beta, 8);
  const bool az = ua == 0, bz = ub == 0;
  int64_t nsw = 0, ncol = 0, nun = 0;
  for (Request* r : g) {
    nsw += r->plan->as_nsweep;
    ncol += r->plan->as_nccol;
    nun += (r->plan->as_nunitsw + round - 1) / round * round;
  }
  SHT_REQUIRE(nsw < (int64_t(1) << 31) && nun < (int64_t(1) << 31), SHT_ERR_UNSUPPORTED, "batch too large");
  AsSweep* sweeps = static_cast<AsSweep*>(arena.allocate(std::max<int64_t>(nsw, 1) * sizeof(AsSweep)));
  b.scratch.push_back(sweeps);
  int64_t* ccol = static_cast<int64_t*>(arena.allocate(std::max<int64_t>(ncol, 1) * sizeof(int64_t)));
  b.scratch.push_back(ccol);
  AsUnit* units = static_cast<AsUnit*>(arena.allocate(std::max<int64_t>(nun, 1) * sizeof(AsUnit)));
  b.scratch.push_back(units);
  auto* qctr = static_cast<unsigned long long*>(arena.allocate(8 * sizeof(unsigned long long)));
  b.scratch.push_back(qctr);
  const double* zero = nullptr;
  int64_t s0 = 0, c0 = 0, u0 = 0;
  for (Request* r : g) {
    sht_plan* P = r->plan;
    enqueue_stage_in(P, first->alpha, first->beta, s);
    const void* Ad = P->permA ? P->d_Ap : P->A->d_data;
    const void* Bd = P->permB ? P->d_Bp : P->B->d_data;
    void* Cd = P->permC ? P->d_Cp : P->C->d_data;
    const bool zinit = bz && !az && P->as_all_touched;
    if (zinit) zero = P->d_as_zero;
    if (!zinit) launch_scale(static_cast<double*>(Cd), P->C->nelem, first->beta, bz ? 1 : 0, s);
    if (!az) {
      launch_bt(bt_format(P), P->d_bt_rec, P->bt_nrec, Bd, op32, P->d_bt, P->bt_total, s);
      if (P->as_nsweep)
        hipLaunchKernelGGL(k_rebase_sweeps, dim3(nblk(P->as_nsweep)), dim3(256), 0, s, sweeps + s0, P->d_as_sweep,
                           P->as_nsweep, rel(Ad, op32 ? 4 : 8), rel(P->d_bt, tb), c0);
      if (P->as_nccol)
        hipLaunchKernelGGL(k_rebase_ccol, dim3(nblk(P->as_nccol)), dim3(256), 0, s, ccol + c0, P->d_as_ccol,
                           P->as_nccol, rel(Cd, 8), zinit ? 1 : 0);
      const int64_t npad = (P->as_nunitsw + round - 1) / round * round;
      if (npad)
        hipLaunchKernelGGL(k_rebase_units, dim3(nblk(npad)), dim3(256), 0, s, units + u0, P->d_as_unitw,
                           P->as_nunitsw, npad, static_cast<int32_t>(s0));
      SHT_HIP(hipGetLastError());
      s0 += P->as_nsweep;
      c0 += P->as_nccol;
      u0 += npad;
    }
  }
  if (!az && u0 > 0) {
    AsArgs a{};
    a.A = base;
    a.B = base;
    a.C = const_cast<char*>(base);
    a.bt_total = 0;
    a.sweeps = sweeps;
    a.ccol = ccol;
    a.units = units;
    a.nunits = u0;
    a.alpha = first->alpha;
    a.alpha_one = first->alpha == 1.0 ? 1 : 0;
    a.zero = zero;
    a.qctr = qctr;
    a.f32 = op32 ? 1 : 0;
    launch_as(cls, a, s);
    S->nlaunches++;
  }
  for (Request* r : g) enqueue_stage_out(r->plan, s);
}

// Complete the requests of b (status already set on failures): release
// plans, merged records and footprints, wake waiters.
void finish(sht_server* S, Batch& b)
{
  for (void* p : b.scratch) S->ctx->arena.release(p);
  b.scratch.clear();
  if (b.ev) (void)hipEventDestroy(b.ev);
  b.ev = nullptr;
  for (Request* r : b.reqs)
    if (r->plan) {
      sht_plan_destroy(r->plan);
      r->plan = nullptr;
    }
  const double t = now_ms();
  std::lock_guard<std::mutex> lk(S->mu);
  for (Request* r : b.reqs) {
    r->t_done = t;
    r->state = RState::Done;
    S->in_use -= r->footprint;
    S->active.erase(std::remove(S->active.begin(), S->active.end(), r), S->active.end());
    if (r->status == SHT_ERR_CANCELLED) S->ncancelled++;
    S->nreq_done++;
  }
  S->cv_done.notify_all();
  S->cv_work.notify_all();
}

// Admission scan over the pending queue (lock held).
void admit(sht_server* S)
{
  std::vector<const Request*> skipped;
  int n = 0;
  for (auto it = S->pending.begin(); it != S->pending.end() && n < S->max_batch;) {
    Request* r = *it;
    bool wait = false;
    for (const Request* q : S->active) wait = wait || depends(r, q) || depends(q, r);
    for (const Request* q : skipped) wait = wait || depends(r, q) || depends(q, r);
    if (wait) {
      skipped.push_back(r);
      ++it;
      continue;
    }
    if (S->in_use + r->footprint > S->budget && S->in_use > 0) break;  // memory: arrival order
    S->in_use += r->footprint;
    S->active.push_back(r);
    r->state = RState::Planning;
    S->to_plan.push_back(r);
    it = S->pending.erase(it);
    ++n;
  }
  if (n) S->cv_plan.notify_all();
}

void planner_loop(sht_server* S)
{
  (void)hipSetDevice(S->ctx->device);
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_plan.wait(lk, [S] { return S->stop_planners || !S->to_plan.empty(); });
      if (S->to_plan.empty()) return;
      r = S->to_plan.front();
      S->to_plan.pop_front();
      if (r->cancel) {
        r->status = SHT_ERR_CANCELLED;
        r->err = "request cancelled";
        r->state = RState::Ready;
        r->t_ready = now_ms();
        S->ready.push_back(r);
        S->cv_work.notify_all();
        continue;
      }
      S->planning++;
    }
    const size_t c_before = r->C->reserved ? tensor_bytes(r->C) : 0;
    sht_plan* P = nullptr;
    const sht_status st = sht_plan_create(S->ctx, r->expr.c_str(), r->A, r->B, r->C, r->prec, &P);
    const std::string msg = st == SHT_OK ? std::string() : std::string(sht_last_error());
    std::lock_guard<std::mutex> lk(S->mu);
    S->planning--;
    if (st == SHT_ERR_OOM && S->active.size() > 1) {
      // out of arena while others hold memory: back to the head of the
      // queue, admitted again when memory is released
      S->in_use -= r->footprint;
      S->active.erase(std::remove(S->active.begin(), S->active.end(), r), S->active.end());
      r->state = RState::Pending;
      S->pending.push_front(r);
      S->nrequeued++;
      S->cv_work.notify_all();
      continue;
    }
    if (st != SHT_OK) {
      r->status = st;
      r->err = msg;
    } else {
      // the footprint becomes what the request really holds: operands, C
      // (reserved by the plan when it had no pattern) and the plan's memory
      const size_t real = tensor_bytes(r->A) + tensor_bytes(r->B) + std::max(c_before, tensor_bytes(r->C)) +
                          plan_bytes(P);
      S->in_use = S->in_use - r->footprint + real;
      r->footprint = real;
      r->plan = P;
    }
    r->state = RState::Ready;
    r->t_ready = now_ms();
    S->ready.push_back(r);
    S->cv_work.notify_all();
  }
}

void dispatcher_loop(sht_server* S)
{
  // (the context's device is already current in the creating thread; a
  // failure here surfaces as the error of the first batch's HIP calls)
  (void)hipSetDevice(S->ctx->device);
  for (;;) {
    std::vector<Request*> launch;
    bool drained = false;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait_for(lk, std::chrono::microseconds(200), [S] { return S->stop || !S->ready.empty(); });
      admit(S);
      // launch what is ready once nothing admitted is still being planned,
      // or once the oldest ready request has waited the batching window
      double oldest = 1e300;
      for (const Request* r : S->ready) oldest = std::min(oldest, r->t_ready);
      const bool planning = !S->to_plan.empty() || S->planning > 0;
      if (!S->ready.empty() && (!planning || S->stop || now_ms() - oldest >= kBatchWindowMs)) launch.swap(S->ready);
      drained = S->stop && S->pending.empty() && S->to_plan.empty() && S->planning == 0 && S->ready.empty() &&
                launch.empty();
    }
    // retire finished batches (in any order: streams run independently)
    for (auto it = S->inflight.begin(); it != S->inflight.end();) {
      if (hipEventQuery(it->ev) == hipSuccess) {
        finish(S, *it);
        it = S->inflight.erase(it);
      } else {
        ++it;
      }
    }
    if (drained && S->inflight.empty()) return;
    if (launch.empty()) continue;
    // failed and cancelled requests complete at once
    Batch done;
    std::vector<Request*> runnable;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (Request* r : launch) {
        if (r->cancel && r->status == SHT_OK) {
          r->status = SHT_ERR_CANCELLED;
          r->err = "request cancelled";
        }
        if (r->status != SHT_OK) done.reqs.push_back(r);
        else {
          r->state = RState::Launched;
          runnable.push_back(r);
        }
      }
    }
    if (!done.reqs.empty()) finish(S, done);
    // C-stationary plans merge per (precision, operand type, alpha, beta),
    // and so do plans that run the grouped A-stationary FP64/FP32 kernel;
    // other A-stationary plans run alone
    std::vector<std::vector<Request*>> groups;
    {
      std::map<std::tuple<int, int, int, uint64_t, uint64_t>, std::vector<Request*>> merged;
      for (Request* r : runnable) {
        const bool as = r->plan->schedule >= 1;
        if (as && !as_mergeable(r)) {
          groups.push_back({r});
          continue;
        }
        uint64_t ua, ub;
        std::memcpy(&ua, &r->alpha, 8);
        std::memcpy(&ub, &r->beta, 8);
        merged[std::make_tuple(as ? 1 : 0, static_cast<int>(r->prec), static_cast<int>(r->A->dtype), ua, ub)]
            .push_back(r);
      }
      for (auto& kv : merged) groups.push_back(std::move(kv.second));
    }
    for (auto& g : groups) {
      Batch b;
      b.reqs = g;
      hipStream_t s = S->streams[S->next_stream];
      S->next_stream = (S->next_stream + 1) % kServeStreams;
      try {
        if (g.size() == 1 && g[0]->plan->schedule >= 1) {
          enqueue(g[0]->plan, g[0]->alpha, g[0]->beta, s);
          S->nlaunches++;
        } else if (g[0]->plan->schedule >= 1) {
          launch_group_as(S, g, b, s);
        } else {
          launch_group(S, g, b, s);
        }
        SHT_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
        SHT_HIP(hipEventRecord(b.ev, s));
      } catch (const Error& e) {
        (void)hipStreamSynchronize(s);
        for (Request* r : b.reqs) {
          r->status = e.code;
          r->err = e.what();
        }
        finish(S, b);
        continue;
      }
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->nbatches++;
        S->max_merged = std::max<int64_t>(S->max_merged, static_cast<int64_t>(b.reqs.size()));
      }
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
    for (auto& st : S->streams) SHT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    sht_server* raw = S.get();
    S->dispatcher = std::thread([raw] { sht::dispatcher_loop(raw); });
    for (int p = 0; p < sht::kPlanners; ++p) S->planners.emplace_back([raw] { sht::planner_loop(raw); });
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
    // footprint at admission: operands, C when already reserved, and the
    // permuted staging copies the expression needs (corrected to the plan's
    // real memory once it is built)
    size_t fp = sht::tensor_bytes(A) + sht::tensor_bytes(B) + (C->reserved ? sht::tensor_bytes(C) : 0);
    try {
      const sht::ContractMaps m = sht::analyse_expression(r->expr, A->nblk, B->nblk, C->nblk);
      auto lead_trail = [](const std::vector<int>& l, const std::vector<int>& t) {
        for (size_t i = 0; i < l.size(); ++i)
          if (l[i] != static_cast<int>(i)) return false;
        for (size_t i = 0; i < t.size(); ++i)
          if (t[i] != static_cast<int>(l.size() + i)) return false;
        return true;
      };
      if (!lead_trail(m.a_row, m.a_col)) fp += sht::tensor_bytes(A);
      if (!lead_trail(m.b_row, m.b_col)) fp += sht::tensor_bytes(B);
      if (!lead_trail(m.c_row, m.c_col) && C->reserved) fp += sht::tensor_bytes(C);
    } catch (const Error&) {
      // the plan reports the expression error when the request runs
    }
    r->footprint = fp;
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

sht_status sht_server_cancel(sht_server* S, int64_t req_id)
{
  return guard([&] {
    SHT_REQUIRE(S, SHT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    SHT_REQUIRE(it != S->reqs.end(), SHT_ERR_INVALID, "unknown request id");
    sht::Request* r = it->second;
    switch (r->state) {
      case sht::RState::Pending:
        S->pending.erase(std::remove(S->pending.begin(), S->pending.end(), r), S->pending.end());
        r->status = SHT_ERR_CANCELLED;
        r->err = "request cancelled";
        r->state = sht::RState::Done;
        r->t_done = sht::now_ms();
        S->ncancelled++;
        S->nreq_done++;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
        break;
      case sht::RState::Planning:
      case sht::RState::Ready:
        r->cancel = true;  // completes as cancelled instead of launching
        break;
      default:
        throw Error(SHT_ERR_STATE, "request already launched: it cannot be cancelled");
    }
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
    S->cv_done.wait(lk, [r] { return r->state == sht::RState::Done; });
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
    *done = it->second->state == sht::RState::Done ? 1 : 0;
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

sht_status sht_server_counters(sht_server* S, int64_t* ncancelled, int64_t* nrequeued, size_t* in_use)
{
  return guard([&] {
    SHT_REQUIRE(S, SHT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (ncancelled) *ncancelled = S->ncancelled;
    if (nrequeued) *nrequeued = S->nrequeued;
    if (in_use) *in_use = S->in_use;
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
    if (S->dispatcher.joinable()) S->dispatcher.join();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop_planners = true;
      S->cv_plan.notify_all();
    }
    for (auto& t : S->planners)
      if (t.joinable()) t.join();
    for (auto& kv : S->reqs) delete kv.second;
    for (auto& st : S->streams)
      if (st) (void)hipStreamDestroy(st);
    delete S;
  });
}

}  // extern "C"
