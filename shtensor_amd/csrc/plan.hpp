// Contraction plan: task list, C pattern and execution schedule on the GPU.
#pragma once

#include <string>
#include <vector>

#include "internal.hpp"

namespace sht {

#ifdef __HIPCC__
// The A value of a padded contracted position (k >= K in the last 4-k step
// of an MFMA): a zero whose product with alpha, and then with the zero
// padded B value, is -0.  x + (-0) = x for every x, -0 included, so the
// padded products of an MFMA leave each fma chain's value and sign exactly
// where the reference's chain over k < K ends (a +0 product would turn a
// chain that ends at -0 into +0).
__device__ inline double pad_a(double alpha, bool alpha_one)
{
  return alpha_one ? -0.0 : __builtin_copysign(0.0, -alpha);
}
#endif

// One C-stationary tile: rows [m0, m0+rows) and columns [n0, n0+cols) of
// C block `cb`; the kernel class fixes rows/cols.  pad = 1: a padding tile
// (no work) that keeps the XCD-interleaved tile order aligned.
// The tile record carries its C block's record and the first task's
// operand offsets (filled on the device once the tile list is built), so a
// tile's first loads wait for one record, not a chain of three.
struct TileRec {
  int32_t cb;
  int32_t m0;
  int32_t n0;
  int32_t pad;
  int64_t c_off;   // C block (as CBlkRec)
  int32_t m, n;
  int32_t t0, t1;
  int64_t a0, b0;  // first task (t0): A and B block offsets and its K
  int32_t k0;
  int32_t kc;      // the tile's KClass (read by the persistent launch, whose tiles mix classes)
};

// Per C block: element offset, shape and task range (tasks sorted by K key).
struct CBlkRec {
  int64_t c_off;
  int32_t m, n;
  int32_t t0, t1;
};

// Per task: element offsets of the A and B blocks and the contracted length.
struct TaskRec {
  int64_t a_off;
  int64_t b_off;
  int32_t k;
  int32_t pad;
};

struct ExecArgs {
  const void* A;
  const void* B;
  void* C;
  const TileRec* tiles;
  int64_t ntiles;
  const CBlkRec* cblk;
  const TaskRec* tasks;
  double alpha, beta;
  int alpha_zero, beta_zero, alpha_one;
  // A and B hold FP32 (the FP32 and split-BF16 forms read FP32 tensors);
  // C is always FP64
  int f32;
  // persistent launches (the Persist classes): tiles per XCD queue (set by
  // launch_exec)
  int64_t qlen;
};

// kernel classes of the C-stationary launches
enum class KClass : int {
  Exact8 = 0, Exact16, Exact32,      // fma chain, 128 rows x {8,16,32} cols
  Mfma16, Mfma32,                    // FP64 MFMA, one wave, 64 rows x {16,32} cols (large
                                     // launches, beside the 128-row forms)
  Mfma16h, F32x16h,                  // one wave per 32 x 16 tile (a block's last row tile
                                     // when at most 32 rows remain, at most 16 columns)
  Mfma64W, F32W, Bf16W,              // four waves per 64 x 64 tile (more than 32 columns)
  Mfma64S, F32S, Bf16S,              // four waves per 32 x 32 tile (17-32 columns: FP64 / FP32
                                     // for a block's last row tile of at most 32 rows,
                                     // split-BF16 for every row tile)
  Mfma16q, Mfma32q,                  // FP64 one-wave 128-row tiles (more than 96 rows
                                     // remain, large launches)
  Mfma64N, F32N, Bf16N,              // four waves per 64 x 16 tile (at most 16 columns)
  Mfma64M, F32M,                     // FP64 / FP32, four waves per 64 x 32 tile (17-32 columns)
  F32Nq, F32Mq,                      // FP32, four waves per 128 x 16 / 128 x 32 tile (more
                                     // than 96 rows remain, large launches)
  Persist64, Persist32, PersistBf16, // one persistent launch over every tile of the plan
                                     // (FP64 / FP32 / split-BF16 four-wave classes, the
                                     // class of each tile in its record)
  Count
};

int kclass_rows(KClass k);
int kclass_cols(KClass k);
int kclass_waves(KClass k);  // waves per workgroup (one tile per workgroup)

void launch_exec(KClass k, const ExecArgs& a, hipStream_t s);

// Grid of a persistent kernel f of `threads` threads per workgroup: the
// workgroups the device keeps resident (a multiple of eight, one set per XCD
// queue), at most one per work item (queried once per device and kernel).
unsigned resident_grid(const void* f, int threads, int64_t nwork);

// ---- A-stationary row-panel schedule --------------------------------------
// A panel = the B blocks of one contracted key K (the partners of every A
// block (I, K)) in ascending block column J, converted once per execution
// into one packed matrix whose columns are the partners' columns side by
// side.  The sweep list of block row I: for every A block (I, K), in
// ascending K, the panel of K cut into sweeps of at most kAsSweepSlots
// 16-column slots.  A unit = one chunk of block row I (32 or 64 rows)
// walking the sweep list of I, so each A element is read once per sweep of
// its block and every C element is updated in place in the task order
// (C id, K) of the task list.
constexpr int kAsSweepSlots = 8;
constexpr int kAsNarrowSlots = 4;  // slots per sweep of the narrow schedule (3)
struct AsSweep {
  int64_t a_off;   // A block (m x k, column-major)
  int64_t b_off;   // first sweep column in the panel, in elements of the B format
  int64_t c0;      // first entry of the sweep's 16*ns column offsets (AsArgs::ccol)
  int32_t k;       // contracted length of the A block
  int32_t m;       // rows of block row I (= rows of every C block (I, J))
  int32_t pitch;   // F64T/F32T: panel row pitch (columns rounded up to 16); BF16X3: contracted pad (k rounded up to 32)
  int32_t ns;      // 16-column slots of the sweep (1..kAsSweepSlots)
};
constexpr int kBtPad = 16;  // panel columns are padded (zeros) to a multiple of 16
constexpr int kBfK = 32;    // k granularity of the bf16 MFMA (16x16x32)
// One record per B block: where its columns go in its panel.
//  F64T / F32T: element (k, j) -> bt[bt_off + k * pitch + j]; rows up to
//               roundup(k, 16) are written (zeros past k);
//  BF16X3: three bf16 planes (hi, mid, lo) of x = hi + mid + lo, bt_total
//          elements apart; element (k, j) -> plane[bt_off + j * pitch + k],
//          pitch = roundup(k, 32), zeros past k (so one 16-byte load gives the
//          8 consecutive k of one column an MFMA lane needs).
// Pad columns of a panel are zeroed once when the schedule is built.
struct BtRec {
  int64_t b_off;   // canonical B block (k x n, column-major)
  int64_t bt_off;  // first column of the block inside its panel
  int32_t k, n;
  int32_t pitch;
  int32_t pad;
};
// Panel formats: FP64 / FP32 k-major rows (the FP64 and FP32 MFMA classes),
// three bf16 planes with contracted indices contiguous per column (the
// one-wave split-BF16 class), FP32 in that column layout (the grouped
// split-BF16 class, which splits the panel as it stages it in LDS).
enum class BtFormat : int { F64T = 0, F32T = 1, BF16X3 = 2, F32K = 3 };
void launch_bt(BtFormat f, const BtRec* recs, int64_t nrec, const void* B, bool b_f32, void* BT, int64_t bt_total,
               hipStream_t s);
struct AsUnit {
  int32_t sweep0;  // first sweep of the block row
  int32_t nsweep;  // sweeps of the block row (0: empty unit)
  int32_t r0;      // first row of the chunk
  int32_t pad;     // 1: ghost unit (pads a grouped workgroup; no rows of its own)
};
struct AsArgs {
  const void* A;
  const void* B;      // panels (BtFormat of the plan)
  void* C;
  int64_t bt_total;   // elements per bf16 plane
  const AsSweep* sweeps;
  // per sweep column: element offset v of C(0, column); -1 for pad columns /
  // dropped blocks; -(v + 2) when this sweep is the first of its block row
  // that touches the column
  const int64_t* ccol;
  const AsUnit* units;
  int64_t nunits;
  double alpha;
  int alpha_one;
  // non-null (beta == 0, every C block has a task): the first sweep of a
  // column starts from zeros read here (kAsZeroLen doubles) instead of C,
  // and no C = beta * C pass runs before the sweeps
  const double* zero;
  // the grouped FP64/FP32 kernel is persistent: its workgroups take
  // workgroup tasks from the eight XCD queues of the unit list through these
  // eight counters (cleared by launch_as before every launch)
  unsigned long long* qctr;
  // A (and the B the panels are made from) hold FP32: the FP32 and
  // split-BF16 forms read FP32 tensors; C is always FP64
  int f32;
};
constexpr int kAsZeroLen = 256;
// Kernel classes.  The MFMA classes run 32-row units (two 16-row MFMA tiles
// per wave): Mfma64/Mfma32/Bf16x3 one wave per workgroup, the *G variants
// kAsGroupWaves waves (units of one block row) per workgroup sharing panel
// fragments (through LDS for FP64/FP32, through L1 for split-BF16).  Exact runs 64-row units with one row per lane.
// The *N classes (schedule 3) run sweeps of at most kAsNarrowSlots slots at
// three waves per SIMD.
enum class AsClass : int { Mfma64 = 0, Exact = 1, Mfma64G = 2, Mfma32 = 3, Bf16x3 = 4, Mfma32G = 5, Bf16x3G = 6,
                           Mfma64N = 7, Mfma32N = 8, Count };
constexpr int kAsUnitRows = 32;    // rows of a unit of the MFMA kernels
constexpr int kAsUnitRowsW = 64;   // rows of a unit of the fma-chain kernel
constexpr int kAsUnitRowsN = 64;   // rows of a unit of the narrow schedule (four MFMA tiles)
constexpr int kAsGroupWaves = 4;   // waves per workgroup of the *G classes
constexpr int kAsxCols = 8;        // columns per fma-chain slot
constexpr int kAsxSlots = 8;       // slots per fma-chain sweep
void launch_as(AsClass k, const AsArgs& a, hipStream_t s);
void launch_as_bf3(AsClass k, const AsArgs& a, hipStream_t s);  // the split-BF16 classes
// waves per workgroup of a class (its units come grouped that way)
int as_group_waves(AsClass k);
// The A-stationary kernel class and panel format of a plan's schedule.
AsClass as_class(const sht_plan* P);
BtFormat bt_format(const sht_plan* P);
// C = beta ? beta*C : 0 over nelem doubles (the A-stationary kernel accumulates)
void launch_scale(double* c, int64_t nelem, double beta, int beta_zero, hipStream_t s);

// Block permutation (Utils::reshape semantics per block, Utils.h:315-359):
// out block = in block with axes reordered by `order`.
struct PermArgs {
  TensorDesc d;
  const void* in;
  void* out;
  int order[kMaxModes];
  int inverse;  // 0: src layout -> canonical, 1: canonical -> src layout
};
void launch_permute(const PermArgs& p, bool f64, hipStream_t s);

}  // namespace sht

struct sht_plan {
  sht_ctx* ctx = nullptr;
  std::string expr;
  const sht_tensor* A = nullptr;
  const sht_tensor* B = nullptr;
  sht_tensor* C = nullptr;
  sht_precision prec = SHT_PREC_FP64;
  sht::ContractMaps maps;
  int64_t ntasks = 0, ncb = 0;
  // device (arena)
  int64_t* d_ta = nullptr;
  int64_t* d_tb = nullptr;
  int64_t* d_tc = nullptr;
  int64_t* d_kk = nullptr;
  sht::TaskRec* d_tasks = nullptr;
  // FP32 staging of FP64 operands for the C-stationary FP32 and split-BF16
  // classes: each execution converts A and B once and the tiles read the
  // FP32 copies (chosen when the tiles read each operand element many times)
  bool stage32 = false;
  void* d_A32 = nullptr;
  void* d_B32 = nullptr;
  sht::CBlkRec* d_cblk = nullptr;
  sht::TileRec* d_tiles = nullptr;
  struct Launch {
    sht::KClass k;
    int64_t first, count;
    double cost;  // summed tile cost estimates (stream dealing)
  };
  std::vector<Launch> launches;
  // plan-owned streams for concurrent C-stationary classes, with the events
  // that fork them from and join them to the caller's stream
  static constexpr int kSideStreams = 3;
  hipStream_t side[kSideStreams] = {nullptr, nullptr, nullptr};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kSideStreams] = {nullptr, nullptr, nullptr};
  // C-stationary tile variant (0 automatic .. 4 uniform 64 x 64) and the host
  // copies of the C block records and per-block work it is built from
  int cs_variant = 0;
  std::vector<sht::CBlkRec> h_cblk;
  std::vector<int64_t> h_work;
  // staging for non-canonical layouts
  bool permA = false, permB = false, permC = false;
  void* d_Ap = nullptr;
  void* d_Bp = nullptr;
  void* d_Cp = nullptr;
  int orderA[sht::kMaxModes], orderB[sht::kMaxModes], orderC[sht::kMaxModes];
  // host copies
  std::vector<int64_t> h_ta, h_tb, h_tc, h_kk;
  bool host_tasks = false;
  double flops = 0.0;
  double sum_k = 0.0;     // contracted length summed over tasks
  double bytes_in = 0.0;  // A + B + C write
  double bytes_c = 0.0;   // C read (beta != 0)
  // A-stationary schedule (built on demand)
  // 0: C-stationary tiles, 1: A-stationary one-wave units, 2: A-stationary
  // grouped units, 3: A-stationary one-wave units over narrow sweeps
  int schedule = 0;
  bool as_built = false;
  sht::AsSweep* d_as_sweep = nullptr;
  int64_t* d_as_ccol = nullptr;
  sht::AsUnit* d_as_unit = nullptr;   // one unit per workgroup (32 rows; 64 for the fma chain)
  sht::AsUnit* d_as_unitw = nullptr;  // kAsGroupWaves units of one block row per workgroup
  sht::AsSweep* d_as_sweep4 = nullptr;  // the narrow sweep list (kAsNarrowSlots; columns shared) and units
  sht::AsUnit* d_as_unit4 = nullptr;
  sht::BtRec* d_bt_rec = nullptr;
  void* d_bt = nullptr;
  int64_t bt_nrec = 0, bt_total = 0;
  int64_t as_nunits = 0, as_nunitsw = 0, as_nunits4 = 0, as_nsweep = 0, as_npanel = 0, as_nccol = 0;
  bool as_all_touched = false;  // every C block has a task (first sweeps may initialise C)
  double* d_as_zero = nullptr;  // kAsZeroLen zeros
  unsigned long long* d_as_qctr = nullptr;  // work counters of the persistent grouped kernel
  // host wall time of the build steps (task list + C pattern, C-stationary
  // tiles, A-stationary schedule), ms
  double ms_tasks = 0.0, ms_cstat = 0.0, ms_astat = 0.0;
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  double graph_alpha = 0.0, graph_beta = 0.0;

  void release()
  {
    sht::Arena& a = ctx->arena;
    for (void* p : {static_cast<void*>(d_ta), static_cast<void*>(d_tb), static_cast<void*>(d_tc),
                    static_cast<void*>(d_kk), static_cast<void*>(d_tasks), static_cast<void*>(d_cblk),
                    static_cast<void*>(d_tiles), d_Ap, d_Bp, d_Cp, static_cast<void*>(d_as_sweep),
                    static_cast<void*>(d_as_ccol), static_cast<void*>(d_as_unit),
                    static_cast<void*>(d_as_unitw), static_cast<void*>(d_as_sweep4),
                    static_cast<void*>(d_as_unit4), static_cast<void*>(d_bt_rec), d_bt,
                    static_cast<void*>(d_as_zero), static_cast<void*>(d_as_qctr), d_A32, d_B32})
      if (p) a.release(p);
    if (graph) (void)hipGraphExecDestroy(graph);
    if (graph_def) (void)hipGraphDestroy(graph_def);
    for (int i = 0; i < kSideStreams; ++i) {
      if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
      if (side[i]) (void)hipStreamDestroy(side[i]);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
  }
};

namespace sht {
// Enqueue one execution of P on s (events around grouped launches when ev).
void enqueue(sht_plan* P, double alpha, double beta, hipStream_t s, std::vector<hipEvent_t>* ev = nullptr);
// The permutation staging parts of an execution (non-canonical layouts).
void enqueue_stage_in(sht_plan* P, double alpha, double beta, hipStream_t s);
void enqueue_stage_out(sht_plan* P, hipStream_t s);
}  // namespace sht
