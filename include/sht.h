/*
 * sht.h — C ABI of shtensor_amd, an MI355X-native block-sparse tensor
 * contraction with the contract of shtensor (Kernel<T>, Tensor<T,N>).
 *
 * All entry points return sht_status; on failure sht_last_error() holds the
 * message (thread-local), using the reference's message text where the
 * reference raises std::runtime_error.  No exception crosses this boundary.
 * Device pointers are HIP device addresses; `stream` is a hipStream_t passed
 * as void* (NULL = the null stream).  Layouts are column-major, first index
 * fastest, as in the reference (src/Basics/Utils.h:80-94).
 */
#ifndef SHT_H
#define SHT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SHT_OK = 0,
  SHT_ERR_INVALID = 1,     /* bad argument                                      */
  SHT_ERR_EXPR = 2,        /* contraction expression rejected (ContractInfo.h)   */
  SHT_ERR_OOM = 3,         /* arena exhausted ("Out of memory", MemoryPool.h:148) */
  SHT_ERR_HIP = 4,         /* HIP runtime or hipRTC failure                     */
  SHT_ERR_STATE = 5,       /* object in the wrong state (e.g. not reserved)     */
  SHT_ERR_UNSUPPORTED = 6, /* valid request this build does not implement       */
  SHT_ERR_ADMISSION = 7,   /* serving: request over the admission budget        */
  SHT_ERR_CANCELLED = 8    /* serving: request cancelled before it was launched  */
} sht_status;

/* src/Basics/Definitions.h:45-49 */
typedef enum { SHT_FLOAT32 = 0, SHT_FLOAT64 = 1 } sht_dtype;
/* src/Kernel/KernelDefinitions.h:11-15 — accepted, single HIP backend */
typedef enum { SHT_KERNEL_LAPACK = 0, SHT_KERNEL_XMM = 1 } sht_kernel_type;
/* src/Basics/BlockSpan.h:87-91 */
typedef enum { SHT_NORM_FROBENIUS = 0, SHT_NORM_MAXABS = 1 } sht_block_norm;

/* Arithmetic of the block products of a contraction.                        */
typedef enum {
  SHT_PREC_FP64 = 0,       /* FP64 MFMA; <=1e-12 normwise vs the fma chain      */
  SHT_PREC_FP64_EXACT = 1, /* FP64 fma chain in the reference's order: bitwise  */
  SHT_PREC_FP32 = 2,       /* FP32 MFMA per block, FP64 across blocks           */
  SHT_PREC_BF16X3 = 3      /* split-BF16 x3 MFMA per block, FP64 across blocks  */
} sht_precision;

typedef struct sht_ctx sht_ctx;
typedef struct sht_kernel sht_kernel;
typedef struct sht_tensor sht_tensor;
typedef struct sht_plan sht_plan;

const char* sht_last_error(void);
const char* sht_version(void);

/* ---- context + HBM arena (replaces Context + MemoryPool + Window:
 *      src/Shmem/Context.cpp:9-48, MemoryPool.h:126-354, Window.h:11-71) ---- */
sht_status sht_ctx_create(size_t arena_bytes, int device, sht_ctx** out);
/* Host-backed context: same allocator over host memory, no device use.
 * Lets hosts (and tests on machines without a GPU) use the arena
 * accounting; tensors cannot be created in it.                              */
sht_status sht_ctx_create_host(size_t arena_bytes, sht_ctx** out);
sht_status sht_ctx_destroy(sht_ctx* ctx);
sht_status sht_ctx_arena_info(const sht_ctx* ctx, size_t* total, size_t* free_bytes, size_t* largest_free);
/* MemoryPool::allocate/reallocate/free (first fit, coalescing); 256-B aligned */
sht_status sht_arena_alloc(sht_ctx* ctx, size_t bytes, void** dptr);
sht_status sht_arena_realloc(sht_ctx* ctx, void* dptr, size_t bytes, void** out);
sht_status sht_arena_free(sht_ctx* ctx, void* dptr);
sht_status sht_arena_size(const sht_ctx* ctx, const void* dptr, size_t* bytes);

/* ---- device helpers (plain HIP, for hosts without their own HIP code) ----- */
sht_status sht_dev_alloc(size_t bytes, void** dptr);
sht_status sht_dev_free(void* dptr);
sht_status sht_copy_h2d(void* dst, const void* src, size_t bytes);
sht_status sht_copy_d2h(void* dst, const void* src, size_t bytes);
sht_status sht_device_synchronize(void);
sht_status sht_stream_create(void** stream);
sht_status sht_stream_destroy(void* stream);
sht_status sht_stream_synchronize(void* stream);

/* ---- diagnostics (no counterpart in the reference) ------------------------
 * FP64 MFMA rate (TFLOP/s) of a register-resident loop of independent
 * v_mfma_f64_16x16x4f64 (two waves per SIMD), run on `stream`: the matrix-core
 * rate this device sustains, for comparison with the contraction kernels.   */
sht_status sht_diag_mfma_f64_rate(void* stream, double* tflops);
/* Rate (TFLOP/s, 2048 flop per MFMA) of the grouped A-stationary kernel's
 * step shape with every operand in cache: per 4-k step one 16-byte A load,
 * `slots` panel fragments from an LDS ring and 2 x slots MFMAs, a workgroup
 * barrier every two steps, two 4-wave workgroups per CU (prec SHT_PREC_FP64
 * or SHT_PREC_FP32, slots 1..8): the ceiling of that loop without memory
 * misses, sweep boundaries, ghost waves or a tail.                          */
sht_status sht_diag_step_rate(sht_precision prec, int slots, void* stream, double* tflops);

/* ---- block-product kernel (replaces Kernel<T>, src/Kernel/Kernel.h:37-93,
 *      KernelImpl.cpp:9-73 ctor, 210-770 XMM JIT, 142-208 LAPACK) ----------
 * sht_kernel_create: validates the expression exactly as ContractInfo.h:68-346
 * and specialises a HIP kernel for the (expr, sizes, dtype, alpha, beta) key
 * with hipRTC (cached per process and on disk).  alpha/beta are converted to
 * dtype.  kernel_type is accepted for source compatibility only.          */
sht_status sht_kernel_create(const char* expr, const int* sizes_a, int na, const int* sizes_b, int nb,
                             const int* sizes_c, int nc, double alpha, double beta, sht_dtype dtype,
                             sht_kernel_type kernel_type, sht_kernel** out);
/* Kernel<T>::call (Kernel.h:88-91): nb_ops packed same-shape products,
 * C_op = beta*C_op + sum_k (alpha*A_op) B_op.  nb_ops <= 0 is a no-op.
 * Device pointers: asynchronous on `stream`.  Host pointers (pageable or
 * pinned, the reference's std::vector data) are staged through device
 * buffers the kernel allocates and keeps; the call then returns after C is
 * written back.                                                             */
sht_status sht_kernel_call(sht_kernel* k, const void* a, const void* b, void* c, int64_t nb_ops, void* stream);
/* No counterpart in the reference: sht_kernel_call for callers that cannot
 * allocate device memory (code under HIP graph capture is not meant: the
 * call synchronizes; tight serving loops, harnesses that forbid hipMalloc).
 * Host operands are staged through one fixed 64 MiB device array per GPU
 * instead of buffers the kernel allocates; device operands are used in
 * place.  Fails with SHT_ERR_INVALID when the staged operands of nb_ops ops
 * (A, B and C, each rounded up to 256 bytes) exceed 64 MiB.  Returns after C
 * is written.  Calls that overlap in time on one GPU share the array and
 * race.  No hipMalloc or hipFree.                                           */
sht_status sht_kernel_call_static(sht_kernel* k, const void* a, const void* b, void* c, int64_t nb_ops, void* stream);
/* KernelImpl::get_info (KernelImpl.cpp:773-794) text into buf.             */
sht_status sht_kernel_info(const sht_kernel* k, char* buf, size_t len);
sht_status sht_kernel_destroy(sht_kernel* k);
/* Compile a kernel's code object into the disk cache without a GPU.        */
sht_status sht_kernel_prewarm(const char* expr, const int* sizes_a, int na, const int* sizes_b, int nb,
                              const int* sizes_c, int nc, double alpha, double beta, sht_dtype dtype);
/* compute_contract_info (src/Basics/ContractInfo.h:68-346), host only.
 * maps_out: 6 records of 17 ints (len, v0..v15): A.row, A.col, B.row, B.col,
 * C.row, C.col.  Fails with SHT_ERR_EXPR and the reference's message.      */
sht_status sht_contract_info(const char* expr, const int* sizes_a, int na, const int* sizes_b, int nb,
                             const int* sizes_c, int nc, int* maps_out);

/* ---- block-sparse tensor (replaces Tensor<T,N>/BaseTensor,
 *      src/Tensor/Tensor.h:27-294, BaseTensor.h:27-53, BaseTensor.cpp) ----- */
/* block_sizes: N pointers to nblk[d] block lengths.                        */
sht_status sht_tensor_create(sht_ctx* ctx, sht_dtype dtype, int N, const int* const* block_sizes,
                             const int* nblk, sht_tensor** out);
sht_status sht_tensor_destroy(sht_tensor* t);
/* BaseTensor::reserve (BaseTensor.cpp:221-355): N pointers to nb block
 * coordinates; ids sorted ascending, packed zeroed data in the arena.
 * Errors: "reserve: Indices do not have the same size in each dimension"
 * cannot arise through this signature; duplicates and out-of-range
 * coordinates are rejected.                                                 */
sht_status sht_tensor_reserve(sht_tensor* t, const int* const* idx_per_mode, int64_t nb);
/* reserve by rolled block ids (any order)                                  */
sht_status sht_tensor_reserve_ids(sht_tensor* t, const int64_t* ids, int64_t nb);
sht_status sht_tensor_reserve_all(sht_tensor* t);
/* host views valid until the next reserve/filter/compress of t             */
sht_status sht_tensor_sparse_info(const sht_tensor* t, const int64_t** indices, const int64_t** offsets,
                                  const int64_t** rows, int64_t* nnz);
sht_status sht_tensor_data(const sht_tensor* t, void** dptr, int64_t* nelem);
sht_status sht_tensor_nb_nzblocks(const sht_tensor* t, int64_t* local, int64_t* global);
sht_status sht_tensor_occupation(const sht_tensor* t, double* occ);
/* copy the whole packed data from/to host memory                           */
sht_status sht_tensor_upload(sht_tensor* t, const void* host, int64_t nelem);
sht_status sht_tensor_download(const sht_tensor* t, void* host, int64_t nelem);
/* Per-block norms into a device array of nnz doubles (FP64 accumulation).  */
/* No counterpart in the reference (Tensor<T,N> fixes T per tensor): dst (same
 * block sizes, same context; unreserved, or reserved with src's pattern)
 * takes src's block pattern and src's values converted to dst's float type
 * (FP64 -> FP32 rounds to nearest even).  Asynchronous on `stream` after the
 * reservation.  The FP32 and BF16X3 contraction forms read FLOAT32 A and B. */
sht_status sht_tensor_convert(const sht_tensor* src, sht_tensor* dst, void* stream);
sht_status sht_tensor_block_norms(const sht_tensor* t, sht_block_norm norm, double* d_out, void* stream);
/* Tensor::filter (Tensor.h:336-363): block ids with norm < eps are negated
 * (-id-1); compress!=0 then compacts (BaseTensor::compress, fixed).        */
sht_status sht_tensor_filter(sht_tensor* t, double eps, sht_block_norm norm, int compress);
sht_status sht_tensor_compress(sht_tensor* t);

/* ---- synthetic workloads (counter-based SplitMix64; identical on CPU) ---- */
sht_status sht_gen_block_sizes(uint64_t seed, int axis, int length, const int* choices, int nchoices, int* out,
                               int cap, int* count);
/* reserve the Bernoulli(p) pattern of tensor `tid`                          */
sht_status sht_gen_reserve(sht_tensor* t, uint64_t seed, int tid, double p);
/* fill values u(-1,1) * exp(-lambda*|c0-c1|) on the device                  */
sht_status sht_gen_fill(sht_tensor* t, uint64_t seed, int tid, double lambda, void* stream);

/* ---- contraction (the driver the reference lacks: src/Tensor/Tensor.cpp is
 *      empty; semantics follow the Kernel's per-element fma chain) ----------
 * C = beta*C + alpha * contract(expr, A, B).  If C holds no blocks its
 * pattern is built from the product (sorted unique product blocks) and
 * reserved; otherwise C's pattern is kept and products outside it dropped.
 * The task list and pattern are built on the GPU.
 * Tensor types: C is FLOAT64 in every form (the sum over contracted blocks
 * runs in FP64).  A and B share one type: FLOAT64 for every form, or
 * FLOAT32 for SHT_PREC_FP32 / SHT_PREC_BF16X3 (both round each operand to
 * FP32 first, so FLOAT32 operands give the same C bit for bit as the FLOAT64
 * tensors they were rounded from, with half the operand bytes).  Other
 * combinations fail with SHT_ERR_UNSUPPORTED.                              */
sht_status sht_contract(sht_ctx* ctx, const char* expr, double alpha, const sht_tensor* A, const sht_tensor* B,
                        double beta, sht_tensor* C, sht_precision prec, void* stream);

/* Plans: build once (task list, C pattern, schedule), execute many times.  */
sht_status sht_plan_create(sht_ctx* ctx, const char* expr, const sht_tensor* A, const sht_tensor* B,
                           sht_tensor* C, sht_precision prec, sht_plan** out);
sht_status sht_plan_execute(sht_plan* p, double alpha, double beta, void* stream);
/* Execute once with a HIP event pair around every grouped launch (on
 * `stream`), wait, and report each launch's kernel class (kclass_out) and
 * duration in ms.  *nlaunch receives the launch count (<= cap written).     */
sht_status sht_plan_execute_timed(sht_plan* p, double alpha, double beta, void* stream, int* kclass_out,
                                  float* ms_out, int cap, int* nlaunch);
/* Name of a kernel class reported by sht_plan_execute_timed.               */
const char* sht_kclass_name(int kclass);
/* Execution schedule: 0 = C-stationary tiles (one tile per workgroup, C
 * block accumulated in registers over its tasks); 1 = A-stationary row
 * panels, one wave (one 32-row chunk of a block row) per workgroup, each A
 * element read once per sweep of at most 8 16-column slots; 2 = the same with
 * the chunks of one block row grouped four to a workgroup, sharing the panel
 * (not for SHT_PREC_FP64_EXACT); 3 = schedule 1 over sweeps of at most 4
 * slots at three waves per SIMD (SHT_PREC_FP64 and SHT_PREC_FP32 only).
 * Schedules 1-3 need every A block below 2^28 elements.  Results do not
 * depend on the schedule (FP64 forms: bit for bit).  A new plan starts with
 * schedule 2 when A holds at least 8x B's elements and its tasks average a
 * contracted length of at least 64 (not SHT_PREC_FP64_EXACT), else 0.     */
sht_status sht_plan_set_schedule(sht_plan* p, int schedule);
sht_status sht_plan_get_schedule(const sht_plan* p, int* schedule);
/* Autotuner: time each applicable schedule and C-stationary tile variant
 * (`reps` executions on `stream`) and keep the fastest.  ms_out (4 entries,
 * may be NULL) receives the best time of schedules 0-3 (-1 where not tried).
 * C is overwritten by the trial executions.                                 */
sht_status sht_plan_autotune(sht_plan* p, double alpha, double beta, int reps, void* stream, float* ms_out);
/* The autotuner over schedules and C-stationary tile variants: every
 * applicable (schedule, variant) pair is timed (`reps` executions after a
 * warm-up) and the fastest kept.  *n_out receives the candidate count; the
 * first `cap` candidates' schedule, variant and ms go to the arrays (each
 * may be NULL).  C is overwritten by the trials.                            */
sht_status sht_plan_autotune_ex(sht_plan* p, double alpha, double beta, int reps, void* stream, int cap, int* n_out,
                                int* sched_out, int* variant_out, float* ms_out);
/* C-stationary tile variant: 0 automatic (tile class by block width, 128-row
 * tiles in large FP64/FP32 launches, XCD-aware order), 1 without 128-row
 * tiles, 2 with 128-row tiles in any launch, 3 tiles by descending cost,
 * 4 uniform 64 x 64 four-wave tiles for every block (128 x 32 for the fma
 * chain): a fixed-tile reference, 5 one persistent launch serving every tile
 * of the plan (four-wave classes; not for SHT_PREC_FP64_EXACT).  Results do
 * not depend on it.                                                         */
sht_status sht_plan_set_variant(sht_plan* p, int variant);
sht_status sht_plan_get_variant(const sht_plan* p, int* variant);
const char* sht_cs_variant_name(int variant);
/* FP32 staging of FLOAT64 operands (FP32 and BF16X3 plans; no counterpart in
 * the reference, which has no reduced-precision form): with 1, each
 * C-stationary execution first converts A and B once to FP32 copies and the
 * tiles read those; with 0, the tiles read and round the FP64 values at every
 * use.  A new plan that starts on the C-stationary schedule stages when its
 * tiles read each operand element at least ten times on average and the
 * arena holds the copies (4 bytes per element).
 * Both forms round every operand to FP32 first: results are bitwise the same
 * either way.  SHT_ERR_OOM when the arena cannot hold the copies.            */
sht_status sht_plan_set_fp32_staging(sht_plan* p, int on);
sht_status sht_plan_get_fp32_staging(const sht_plan* p, int* on);
/* capture one execution into a hipGraph and replay it `replays` times       */
sht_status sht_plan_graph_replay(sht_plan* p, double alpha, double beta, int replays, void* stream);
/* task list (host copies): positions into A, B, C and the contracted-block
 * key, sorted by (C id, key).  Arrays valid until the plan is destroyed.     */
sht_status sht_plan_tasks(sht_plan* p, const int64_t** ta, const int64_t** tb, const int64_t** tc,
                          const int64_t** kkey, int64_t* ntasks);
/* useful flops (sum 2*m*n*k) and algorithmic HBM bytes of one execution     */
sht_status sht_plan_stats(const sht_plan* p, double* flops, double* bytes, int64_t* ntasks, int64_t* ncblocks);
/* Host wall time (ms) of the plan's build steps: task list and C pattern,
 * C-stationary tiles, A-stationary schedule (-1 if not built yet).         */
sht_status sht_plan_build_ms(const sht_plan* p, double* tasks_ms, double* cstat_ms, double* astat_ms);
sht_status sht_plan_destroy(sht_plan* p);

/* ---- serving: continuous batching of contractions ----------------------
 * Requests name operand tensors resident in the context's arena.  A
 * dispatcher thread admits requests in arrival order while the summed
 * footprint of unfinished requests fits `budget_bytes` (A + B + reserved C +
 * the permuted staging copies the expression needs at submit; corrected to
 * the operands, C and the plan's own memory once planned); a request that
 * writes a tensor an earlier unfinished request touches, or reads one it
 * writes, waits for it, while later independent requests pass it.  Planner
 * threads build the plans (a plan that runs out of arena while other
 * requests hold memory is queued again).  Ready requests with C-stationary
 * plans and equal (precision, alpha, beta) are merged into shared grouped
 * launches; a request whose plan picked the A-stationary schedule runs its
 * own launch.  Batches are dealt to 4 device streams, so work admitted while
 * other batches run overlaps them on the GPU.  A request larger than the
 * budget is rejected with SHT_ERR_ADMISSION.  C must not be touched until
 * sht_server_wait returns.  Results are bit-identical to a stand-alone
 * sht_contract.                                                             */
typedef struct sht_server sht_server;
sht_status sht_server_create(sht_ctx* ctx, size_t budget_bytes, int max_batch, sht_server** out);
sht_status sht_server_submit(sht_server* s, const char* expr, double alpha, const sht_tensor* A, const sht_tensor* B,
                             double beta, sht_tensor* C, sht_precision prec, int64_t* req_id);
/* blocks until the request finished; latency = completion - submission.
 * A cancelled request returns SHT_ERR_CANCELLED.                            */
sht_status sht_server_wait(sht_server* s, int64_t req_id, double* latency_ms);
sht_status sht_server_poll(sht_server* s, int64_t req_id, int* done);
/* Cancel a request that has not been launched (waiting for admission, being
 * planned or ready): it completes with SHT_ERR_CANCELLED (a request being
 * planned may already have reserved C's pattern).  SHT_ERR_STATE once it is
 * on the device or finished.                                                */
sht_status sht_server_cancel(sht_server* s, int64_t req_id);
sht_status sht_server_stats(sht_server* s, int64_t* nreq_done, int64_t* nbatches, int64_t* max_merged,
                            int64_t* nlaunches);
/* cancelled and re-queued (arena exhausted at plan time) request counts, and
 * the bytes currently counted against the budget                            */
sht_status sht_server_counters(sht_server* s, int64_t* ncancelled, int64_t* nrequeued, size_t* in_use);
sht_status sht_server_destroy(sht_server* s);

#ifdef __cplusplus
}
#endif

#endif
