/*
 * ORACLE — CPU restatement of shtensor's block-sparse contraction semantics.
 *
 * This is test infrastructure.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker (or as the
 * timed CPU baseline).  The product library (shtensor_amd) never links or
 * calls anything here.
 *
 * Each function cites the reference source it restates (paths relative to
 * the reference tree).  Pinning of each function (golden vectors, closed
 * forms, brute force) is done in tests/test_oracle_*.py.
 */
#ifndef SHT_ORACLE_H
#define SHT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Block-sparse tensor view (storage format of src/Tensor/BaseTensor.cpp:221-355):
 * indices = rolled block ids (mode 0 fastest) sorted ascending,
 * offsets = exclusive prefix sum of block volumes in that order,
 * data    = packed column-major blocks.                                     */
typedef struct {
  int N;
  const int* nblk;         /* [N] number of blocks per mode                  */
  const int* bsizes;       /* concatenation of the N block-size vectors       */
  int64_t nnz;
  const int64_t* indices;  /* [nnz]                                           */
  const int64_t* offsets;  /* [nnz]                                           */
  void* data;              /* double* or float*                               */
} orc_bst;

/* src/Basics/ContractInfo.h:68-346.  maps_out receives 6 records of 17 ints
 * (len, v0..v15) in the order A.row, A.col, B.row, B.col, C.row, C.col.
 * Returns 1 on success, 0 on failure (message in err).                     */
int orc_contract_info(const char* expr, const int* sa, int na, const int* sb, int nb,
                      const int* sc, int nc, int* maps_out, char* err, int errlen);

/* src/Basics/ContractInfo.h:38-66.  Returns the vector length.             */
int orc_scatter(const int* map, int nmap, const int* sizes, int nsizes, int* out);

/* XMM JIT semantics, src/Kernel/KernelImpl.cpp:553-735 with the fma of
 * src/Kernel/AvxHelper.cpp:277-285: per C element c = beta?beta*C:0, then
 * c = fma(alpha*A, B, c) for k ascending.  nb_ops packed same-shape triples.
 * Returns 1 on success, 0 on expression error.                             */
int orc_kernel_xmm_f64(const char* expr, const int* sa, int na, const int* sb, int nb,
                       const int* sc, int nc, double alpha, double beta,
                       const double* a, const double* b, double* c, int64_t nb_ops,
                       char* err, int errlen);
int orc_kernel_xmm_f32(const char* expr, const int* sa, int na, const int* sb, int nb,
                       const int* sc, int nc, float alpha, float beta,
                       const float* a, const float* b, float* c, int64_t nb_ops,
                       char* err, int errlen);

/* LAPACK variant, src/Kernel/KernelImpl.cpp:142-208: reshape (Utils.h:315-359)
 * then netlib dgemm('N','N') (reference BLAS 3.9 loop order) then inverse
 * reshape.                                                                 */
int orc_kernel_lapack_f64(const char* expr, const int* sa, int na, const int* sb, int nb,
                          const int* sc, int nc, double alpha, double beta,
                          const double* a, const double* b, double* c, int64_t nb_ops,
                          char* err, int errlen);

/* Utils::reshape, src/Basics/Utils.h:315-359 (double).                     */
void orc_reshape_f64(const double* in, const int* sizes, int n, const int* order, double* out);

/* BaseTensor::reserve, src/Tensor/BaseTensor.cpp:221-355.  idx is N rows of
 * nb block coordinates.  Outputs sorted ids, offsets and the per-mode-0 count
 * prefix (rows).  Returns the element count, or -1 on error.               */
int64_t orc_reserve(int N, const int* nblk, const int* bsizes, const int* idx, int64_t nb,
                    int64_t* indices_out, int64_t* offsets_out, int64_t* rows_out,
                    char* err, int errlen);

/* Task list and C block pattern of a block-sparse contraction (the rule the
 * reference's missing driver needs): tasks (posA, posB, posC)
 * sorted by (C block id, contracted-block key).  If C is NULL or has nnz==0
 * the C pattern is the sorted unique set of product blocks; otherwise the
 * pattern of C is kept and tasks outside it are dropped.
 * Call with cap=0 to count.  Returns the task count, or -1 on error.
 * c_ids_out (may be NULL) receives the pattern, *nc_out its length.        */
int64_t orc_tasks(const char* expr, const orc_bst* A, const orc_bst* B, const orc_bst* C,
                  int64_t* tasks_out, int64_t cap, int64_t* c_ids_out, int64_t* nc_out,
                  int64_t* kkey_out, char* err, int errlen);

/* Block-sparse contraction values.  C's pattern (indices/offsets) must be the
 * one orc_tasks produced (or the kept one).  mode 0: the XMM fma chain
 * continued over tasks in ascending K key (bit-exact definition).
 * mode 1: long double accumulation (error oracle).  subset: list of C
 * positions to compute (NULL = all).  nthreads > 1 runs a work-stealing pool
 * (src/Basics/ThreadPool.cpp:57-196 semantics) over C blocks.
 * Returns number of tasks processed, -1 on error.                          */
int64_t orc_contract_f64(const char* expr, double alpha, const orc_bst* A, const orc_bst* B,
                         double beta, orc_bst* C, int mode, const int64_t* subset,
                         int64_t nsubset, int nthreads, char* err, int errlen);

/* The same, reporting the wall time of building the task list and C
 * pattern (plan_s) and of the block products (compute_s); either may be NULL. */
int64_t orc_contract_f64_t(const char* expr, double alpha, const orc_bst* A, const orc_bst* B, double beta,
                           orc_bst* C, int mode, const int64_t* subset, int64_t nsubset, int nthreads,
                           char* err, int errlen, double* plan_s, double* compute_s);

/* Block norms (FROBENIUS=0 accumulated in double, MAXABS=1) of every block,
 * src/Basics/BlockSpan.h:140-164 with the int accumulator fixed.           */
void orc_block_norms_f64(const orc_bst* T, int norm, double* out);

/* ---- synthetic workload generator (counter-based, SplitMix64) ---------- */
uint64_t orc_splitmix64(uint64_t z);
/* Block sizes of one mode: draws from choices until length is filled, last
 * block clipped.  Returns number of blocks (<= cap).                       */
int orc_gen_block_sizes(uint64_t seed, int axis, int length, const int* choices, int nchoices,
                        int* out, int cap);
/* Present rolled block ids (ascending) with Bernoulli(p).  Returns count.  */
int64_t orc_gen_presence(uint64_t seed, int tid, int N, const int* nblk, double p,
                         int64_t* out, int64_t cap);
/* Values of selected blocks: u(-1,1) * exp(-lambda*|c0-c1|).              */
void orc_gen_values_f64(uint64_t seed, int tid, const orc_bst* T, double lambda,
                        const int64_t* positions, int64_t npos);

/* XMM chain timing kernel for the CPU baseline is orc_contract_f64 mode 0.  */

#ifdef __cplusplus
}
#endif

#endif
