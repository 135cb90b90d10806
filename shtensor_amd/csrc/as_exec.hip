// A-stationary row-panel kernels of a contraction plan.
//
// A unit (one wave, one workgroup) owns a chunk of rows of block row I and
// walks the passes of I: the A blocks (I, K) in ascending contracted key K.
// Per pass it sweeps the packed panel of K (the partners (K, J) side by side)
// one group of 16-column slots at a time: the slots' C tiles are loaded into
// accumulators, the A block's k loop runs with one B fragment per slot and
// step shared by every 16-row tile of the chunk, and the tiles are written
// back.  A C element belongs to exactly one unit, and its contributions
// arrive in the task order (C id, K): no atomics, one fixed order.
//
// MFMA operand layouts (lane l of 64; measured on gfx950):
//   v_mfma_f64_16x16x4f64   A[l%16][l/16], B[l/16][l%16], D[4v + l/16][l%16]
//   v_mfma_f32_16x16x4f32   A[l%16][l/16], B[l/16][l%16], D[4(l/16) + v][l%16]
//   v_mfma_f32_16x16x32bf16 A[l%16][8(l/16)+e], B[8(l/16)+e][l%16], D as f32
//
//  * sweep_f64<R, NS>: FP64 MFMA.  Each C element accumulates as one fused
//    multiply-add chain in k order, c = fma(alpha*a, b, c), continued across
//    passes — bitwise equal to the reference JIT's chain
//    (src/Kernel/KernelImpl.cpp:594-710) continued over the task list.
//  * sweep_f32<R, NS>: FP32 MFMA per block product, added into C in FP64.
//  * sweep_bf3<R, NS>: three-term split-BF16 (x = hi + mid + lo, about 24
//    significant bits), six bf16 products of order <= 2 per 32-k step, FP32
//    per block product, FP64 across blocks.
//  * k_as_exact: the fma chain with one row per lane (bitwise as well).

#include "plan.hpp"

namespace sht {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));

namespace {

// wave-uniform copies (the pass and unit records are the same for every lane)
__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline int64_t uni(int64_t x)
{
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(x) >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

struct PassU {
  int64_t a_off, bt_off;
  int k, m, ncol, pitch, col0, cof0;
};

__device__ inline PassU load_pass(const AsPass* passes, int i)
{
  const AsPass v = passes[i];
  PassU u;
  u.a_off = uni(v.a_off);
  u.bt_off = uni(v.bt_off);
  u.k = uni(v.k);
  u.m = uni(v.m);
  u.ncol = uni(v.ncol);
  u.pitch = uni(v.pitch);
  u.col0 = uni(v.col0);
  u.cof0 = uni(v.cof0);
  return u;
}

// Element offset of C(0, c) for panel column c of the pass, -1 when c is a
// pad column or its C block is not kept.
__device__ inline int64_t col_base(const AsArgs& p, const PassU& ps, int c)
{
  if (c >= ps.ncol) return -1;
  const AsCol cm = p.cols[ps.col0 + c];
  const int64_t cb = p.cofs[ps.cof0 + cm.p];
  return cb < 0 ? -1 : cb + static_cast<int64_t>(cm.j) * ps.m;
}

constexpr int cmin(int a, int b) { return a < b ? a : b; }

// ---------------------------------------------------------------------------
// FP64 MFMA sweep: R row tiles x NS slots.  The k loop is unrolled by two
// with two register sets, so the loads of step s+1 are in flight while the
// MFMAs of step s issue; every branch in it is wave-uniform.  Step indexes
// are clamped to the last step (every load in bounds and unconditional), A
// rows past M read row M-1 (their outputs are never stored), A columns past K
// read column K-1 masked to zero, and panel rows past K are zero.
// ---------------------------------------------------------------------------
template <int R, int NS>
__device__ inline void sweep_f64(const AsArgs& p, const PassU& ps, int g, int r0, int nt, int li, int lk)
{
  const double* A = static_cast<const double*>(p.A) + ps.a_off;
  const double* Bl = static_cast<const double*>(p.B) + ps.bt_off + 16 * g + li + static_cast<int64_t>(lk) * ps.pitch;
  double* C = static_cast<double*>(p.C);
  const int K = ps.k, M = ps.m;
  int64_t cb[NS];
  v4d acc[R][NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) cb[q] = col_base(p, ps, 16 * (g + q) + li);
  // Accumulators start from C.  Pad columns read column 0 of C and rows past
  // M read on into the following memory (tensor buffers carry kReadSlack
  // bytes of slack); those results are never stored.  Each slot's values
  // pass through a 0*0 + C MFMA as soon as they arrive, which moves them into
  // accumulation registers (D = C exactly, except that -0 becomes +0), so the
  // staging registers of only two slots are live at a time.
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const double* cq = C + (cb[q] >= 0 ? cb[q] : 0) + r0 + lk;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      v4d c;
#pragma unroll
      for (int v = 0; v < 4; ++v) c[v] = cq[16 * t + 4 * v];
      acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, c, 0, 0, 0);
    }
    if (q % 2 == 1) __builtin_amdgcn_sched_barrier(0);
  }
  int arow[R];
#pragma unroll
  for (int t = 0; t < R; ++t) arow[t] = min(r0 + 16 * t + li, M - 1);
  const int last = ((K + 3) >> 2) - 1;
  const int64_t bstep = 4 * static_cast<int64_t>(ps.pitch);
  // Loads fill raw registers; the mask and alpha are applied by the consumer
  // (after the scheduling barrier), so nothing waits for a load before the
  // MFMAs that precede its use.
  auto load = [&](int st, double(&a)[R], double(&b)[NS]) {
    const int kk = 4 * st + lk;
    const double* col = A + static_cast<int64_t>(min(kk, K - 1)) * M;
#pragma unroll
    for (int t = 0; t < R; ++t) a[t] = col[arow[t]];
    const double* bs = Bl + st * bstep;
#pragma unroll
    for (int q = 0; q < NS; ++q) b[q] = bs[16 * q];
  };
  auto mma = [&](int st, const double(&araw)[R], const double(&b)[NS]) {
    const bool kin = 4 * st + lk < K;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const double y = kin ? araw[t] : 0.0;
      const double a = p.alpha_one ? y : y * p.alpha;
#pragma unroll
      for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[q], acc[t][q], 0, 0, 0);
    }
  };
  double a0[R], a1[R], b0[NS], b1[NS];
  load(0, a0, b0);
  int st = 0;
  for (; st + 1 <= last; st += 2) {
    load(st + 1, a1, b1);
    __builtin_amdgcn_sched_barrier(0);
    mma(st, a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    load(min(st + 2, last), a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 1, a1, b1);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (st == last) mma(st, a0, b0);  // odd step count: a0/b0 hold the last step
#pragma unroll
  for (int q = 0; q < NS; ++q) {
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + 16 * t + 4 * v + lk;
        if (cb[q] >= 0 && row < M) C[cb[q] + row] = acc[t][q][v];
      }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ---------------------------------------------------------------------------
// C += alpha * (FP32 block product) in FP64, one slot at a time.  The f32
// accumulator layout puts rows 4*(l/16) + v of column l%16 in lane l.  Loads
// are unconditional (rows past M read into the buffer slack), stores are
// guarded.
// ---------------------------------------------------------------------------
template <int R, int NS>
__device__ inline void as_lowp_epilogue(const AsArgs& p, const PassU& ps, int g, int r0, int lk, int li,
                                        const v4f (&acc)[R][NS])
{
  double* C = static_cast<double*>(p.C);
  const int M = ps.m;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const int64_t cb = col_base(p, ps, 16 * (g + q) + li);
    double* cq = C + (cb >= 0 ? cb : 0) + r0 + 4 * lk;
    double cv[R][4];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) cv[t][v] = cq[16 * t + v];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (cb >= 0 && r0 + 16 * t + 4 * lk + v < M) cq[16 * t + v] = cv[t][v] + p.alpha * static_cast<double>(acc[t][q][v]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ---------------------------------------------------------------------------
// FP32 MFMA sweep: same pipeline; A converted to FP32 on load, the panel is
// FP32; the block product accumulates in FP32 and is added into C in FP64.
// ---------------------------------------------------------------------------
template <int R, int NS>
__device__ inline void sweep_f32(const AsArgs& p, const PassU& ps, int g, int r0, int nt, int li, int lk)
{
  const double* A = static_cast<const double*>(p.A) + ps.a_off;
  const float* Bl = static_cast<const float*>(p.B) + ps.bt_off + 16 * g + li + static_cast<int64_t>(lk) * ps.pitch;
  double* C = static_cast<double*>(p.C);
  const int K = ps.k, M = ps.m;
  v4f acc[R][NS];
#pragma unroll
  for (int t = 0; t < R; ++t)
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[t][q] = v4f{0.f, 0.f, 0.f, 0.f};
  int arow[R];
#pragma unroll
  for (int t = 0; t < R; ++t) arow[t] = min(r0 + 16 * t + li, M - 1);
  const int last = ((K + 3) >> 2) - 1;
  const int64_t bstep = 4 * static_cast<int64_t>(ps.pitch);
  auto load = [&](int st, double(&a)[R], float(&b)[NS]) {
    const int kk = 4 * st + lk;
    const double* col = A + static_cast<int64_t>(min(kk, K - 1)) * M;
#pragma unroll
    for (int t = 0; t < R; ++t) a[t] = col[arow[t]];
    const float* bs = Bl + st * bstep;
#pragma unroll
    for (int q = 0; q < NS; ++q) b[q] = bs[16 * q];
  };
  auto mma = [&](int st, const double(&araw)[R], const float(&b)[NS]) {
    const bool kin = 4 * st + lk < K;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const float a = kin ? static_cast<float>(araw[t]) : 0.f;
#pragma unroll
      for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[q], acc[t][q], 0, 0, 0);
    }
  };
  double a0[R], a1[R];
  float b0[NS], b1[NS];
  load(0, a0, b0);
  int st = 0;
  for (; st + 1 <= last; st += 2) {
    load(st + 1, a1, b1);
    __builtin_amdgcn_sched_barrier(0);
    mma(st, a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    load(min(st + 2, last), a0, b0);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 1, a1, b1);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (st == last) mma(st, a0, b0);  // odd step count: a0/b0 hold the last step
  as_lowp_epilogue<R, NS>(p, ps, g, r0, lk, li, acc);
}

// ---------------------------------------------------------------------------
// Three-term split-BF16 sweep.  x (rounded to FP32) = hi + mid + lo, each a
// bf16 (round to nearest even); per 32-k step six products of order <= 2,
// smallest first: lo*hi, mid*mid, hi*lo, mid*hi, hi*mid, hi*hi.
// ---------------------------------------------------------------------------
__device__ inline unsigned short bf16_bits(float x)
{
  const unsigned int u = __float_as_uint(x);
  return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ inline void split3(double xd, unsigned short& h, unsigned short& m, unsigned short& l)
{
  const float x = static_cast<float>(xd);
  h = bf16_bits(x);
  const float r1 = x - __uint_as_float(static_cast<unsigned int>(h) << 16);
  m = bf16_bits(r1);
  const float r2 = r1 - __uint_as_float(static_cast<unsigned int>(m) << 16);
  l = bf16_bits(r2);
}

template <int R, int NS>
__device__ inline void sweep_bf3(const AsArgs& p, const PassU& ps, int g, int r0, int nt, int li, int lk)
{
  const double* A = static_cast<const double*>(p.A) + ps.a_off;
  const int kp = ps.pitch;
  const unsigned short* Bl = static_cast<const unsigned short*>(p.B) + ps.bt_off +
                             static_cast<int64_t>(16 * g + li) * kp + 8 * lk;
  const int64_t plane = p.bt_total;
  const int64_t qstep = 16 * static_cast<int64_t>(kp);
  double* C = static_cast<double*>(p.C);
  const int K = ps.k, M = ps.m;
  v4f acc[R][NS];
#pragma unroll
  for (int t = 0; t < R; ++t)
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[t][q] = v4f{0.f, 0.f, 0.f, 0.f};
  int arow[R];
#pragma unroll
  for (int t = 0; t < R; ++t) arow[t] = min(r0 + 16 * t + li, M - 1);
  const int nsteps = (K + kBfK - 1) / kBfK;
  double x_nx[R][8];
  auto load_a = [&](int st) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int kk = kBfK * st + 8 * lk + e;
      const double* col = A + static_cast<int64_t>(min(kk, K - 1)) * M;
#pragma unroll
      for (int t = 0; t < R; ++t) {
        const double x = col[arow[t]];
        x_nx[t][e] = kk < K ? x : 0.0;
      }
    }
  };
  load_a(0);
  for (int st = 0; st < nsteps; ++st) {
    v8s ah[R], am[R], al[R];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        unsigned short h, m, l;
        split3(x_nx[t][e], h, m, l);
        ah[t][e] = static_cast<short>(h);
        am[t][e] = static_cast<short>(m);
        al[t][e] = static_cast<short>(l);
      }
    load_a(min(st + 1, nsteps - 1));
    const int k0 = kBfK * st;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const unsigned short* b = Bl + q * qstep + k0;
      const v8s bh = *reinterpret_cast<const v8s*>(b);
      const v8s bm = *reinterpret_cast<const v8s*>(b + plane);
      const v8s bl = *reinterpret_cast<const v8s*>(b + 2 * plane);
#pragma unroll
      for (int t = 0; t < R; ++t) {
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[t], bh, acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[t], bm, acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bl, acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[t], bh, acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bm, acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[t], bh, acc[t][q], 0, 0, 0);
      }
    }
  }
  as_lowp_epilogue<R, NS>(p, ps, g, r0, lk, li, acc);
}

template <int R, int F, int NS>
__device__ inline void run_sweep(const AsArgs& p, const PassU& ps, int g, int r0, int nt, int li, int lk)
{
  if constexpr (F == 0) sweep_f64<R, NS>(p, ps, g, r0, nt, li, lk);
  else if constexpr (F == 1) sweep_f32<R, NS>(p, ps, g, r0, nt, li, lk);
  else sweep_bf3<R, NS>(p, ps, g, r0, nt, li, lk);
}

// F: 0 = FP64 MFMA, 1 = FP32 MFMA, 2 = split-BF16.  One wave per unit of
// 16*R rows; sweeps of up to NSMAX slots.
template <int R, int NSMAX, int F>
__global__ __launch_bounds__(64) void k_as(AsArgs p)
{
  if (static_cast<int64_t>(blockIdx.x) >= p.nunits) return;
  const AsUnit u = p.units[blockIdx.x];
  const int npass = uni(u.npass), pass0 = uni(u.pass0), r0 = uni(u.r0);
  const int li = threadIdx.x & 15, lk = threadIdx.x >> 4;
  for (int pi = pass0; pi < pass0 + npass; ++pi) {
    const PassU ps = load_pass(p.passes, pi);
    const int nt = min(R, (ps.m - r0 + 15) >> 4);
    const int nslots = (ps.ncol + 15) >> 4;
    for (int g = 0; g < nslots; g += NSMAX) {
      switch (min(NSMAX, nslots - g)) {
        case 1: run_sweep<R, F, 1>(p, ps, g, r0, nt, li, lk); break;
        case 2: run_sweep<R, F, cmin(2, NSMAX)>(p, ps, g, r0, nt, li, lk); break;
        case 3: run_sweep<R, F, cmin(3, NSMAX)>(p, ps, g, r0, nt, li, lk); break;
        case 4: run_sweep<R, F, cmin(4, NSMAX)>(p, ps, g, r0, nt, li, lk); break;
        case 5: run_sweep<R, F, cmin(5, NSMAX)>(p, ps, g, r0, nt, li, lk); break;
        case 6: run_sweep<R, F, cmin(6, NSMAX)>(p, ps, g, r0, nt, li, lk); break;
        case 7: run_sweep<R, F, cmin(7, NSMAX)>(p, ps, g, r0, nt, li, lk); break;
        default: run_sweep<R, F, cmin(8, NSMAX)>(p, ps, g, r0, nt, li, lk); break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// A-stationary fma chain (bitwise equal to the oracle).  One wave per 64-row
// unit, lane = one C row; per sweep up to kAsxSlots slots of kAsxCols panel
// columns keep their C values in registers while k runs ascending:
// c = fma(alpha*A(i,k), B(k,j), c).  B(k, c0..c0+7) is wave-uniform and
// contiguous in the panel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_as_exact(AsArgs p)
{
  constexpr int S = kAsxSlots, W = kAsxCols;
  if (static_cast<int64_t>(blockIdx.x) >= p.nunits) return;
  const AsUnit u = p.units[blockIdx.x];
  const int npass = uni(u.npass), pass0 = uni(u.pass0), r0 = uni(u.r0);
  const double* A = static_cast<const double*>(p.A);
  const double* BT = static_cast<const double*>(p.B);
  double* C = static_cast<double*>(p.C);
  const int row = r0 + static_cast<int>(threadIdx.x);
  for (int pi = pass0; pi < pass0 + npass; ++pi) {
    const PassU ps = load_pass(p.passes, pi);
    const int K = ps.k, M = ps.m;
    const bool rowok = row < M;
    const double* Ab = A + ps.a_off + min(row, M - 1);
    for (int c0 = 0; c0 < ps.ncol; c0 += S * W) {
      const int ns = min(S, (ps.ncol - c0 + W - 1) / W);
      double acc[S][W];
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int j = 0; j < W; ++j) {
          const int64_t cb = q < ns ? col_base(p, ps, c0 + q * W + j) : -1;
          acc[q][j] = (cb >= 0 && rowok) ? C[cb + row] : 0.0;
        }
      const double* bt = BT + ps.bt_off + c0;
      for (int k = 0; k < K; ++k) {
        double a = Ab[static_cast<int64_t>(k) * M];
        if (!p.alpha_one) a = a * p.alpha;
        const double* bk = bt + static_cast<int64_t>(k) * ps.pitch;
#pragma unroll
        for (int q = 0; q < S; ++q) {
          if (q < ns) {
#pragma unroll
            for (int j = 0; j < W; ++j) acc[q][j] = __fma_rn(a, bk[q * W + j], acc[q][j]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int j = 0; j < W; ++j) {
          const int64_t cb = q < ns ? col_base(p, ps, c0 + q * W + j) : -1;
          if (cb >= 0 && rowok) C[cb + row] = acc[q][j];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Panel conversion, once per execution (one workgroup per B block record).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_bt(const BtRec* recs, int64_t nrec, const double* B, T* BT)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const BtRec rec = recs[r];
    const int krows = (rec.k + 3) / 4 * 4;  // zero rows up to a multiple of 4
    const int64_t tot = static_cast<int64_t>(krows) * rec.n;
    const double* b = B + rec.b_off;
    T* o = BT + rec.bt_off;
    for (int64_t x = threadIdx.x; x < tot; x += blockDim.x) {
      const int k = static_cast<int>(x / rec.n), j = static_cast<int>(x - static_cast<int64_t>(k) * rec.n);
      o[static_cast<int64_t>(k) * rec.pitch + j] =
          k < rec.k ? static_cast<T>(b[k + static_cast<int64_t>(j) * rec.k]) : T(0);
    }
  }
}

__global__ __launch_bounds__(256) void k_bt_bf3(const BtRec* recs, int64_t nrec, const double* B, unsigned short* BT,
                                                int64_t plane)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const BtRec rec = recs[r];
    const int kp = rec.pitch;
    const int64_t tot = static_cast<int64_t>(kp) * rec.n;
    const double* b = B + rec.b_off;
    unsigned short* o = BT + rec.bt_off;
    for (int64_t x = threadIdx.x; x < tot; x += blockDim.x) {
      const int j = static_cast<int>(x / kp), k = static_cast<int>(x - static_cast<int64_t>(j) * kp);
      unsigned short h = 0, m = 0, l = 0;
      if (k < rec.k) split3(b[k + static_cast<int64_t>(j) * rec.k], h, m, l);
      o[x] = h;
      o[x + plane] = m;
      o[x + 2 * plane] = l;
    }
  }
}

__global__ __launch_bounds__(256) void k_scale(double* c, int64_t n, double beta, int bz)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    c[i] = bz ? 0.0 : c[i] * beta;
}

}  // namespace

void launch_as(AsClass k, const AsArgs& a, hipStream_t s)
{
  if (a.nunits <= 0) return;
  const dim3 g(static_cast<unsigned>(a.nunits)), b(64);
  static_assert(kAsUnitRows == 32 && kAsUnitRowsW == 64, "unit rows: 2 and 4 MFMA tiles");
  switch (k) {
    case AsClass::Mfma64: hipLaunchKernelGGL((k_as<2, 8, 0>), g, b, 0, s, a); break;
    case AsClass::Mfma64W: hipLaunchKernelGGL((k_as<4, 4, 0>), g, b, 0, s, a); break;
    case AsClass::Mfma32: hipLaunchKernelGGL((k_as<2, 8, 1>), g, b, 0, s, a); break;
    case AsClass::Mfma32W: hipLaunchKernelGGL((k_as<4, 4, 1>), g, b, 0, s, a); break;
    case AsClass::Bf16x3: hipLaunchKernelGGL((k_as<2, 4, 2>), g, b, 0, s, a); break;
    case AsClass::Bf16x3W: hipLaunchKernelGGL((k_as<4, 4, 2>), g, b, 0, s, a); break;
    case AsClass::Exact: hipLaunchKernelGGL(k_as_exact, g, b, 0, s, a); break;
    default: throw Error(SHT_ERR_INVALID, "unknown A-stationary class");
  }
  SHT_HIP(hipGetLastError());
}

void launch_bt(BtFormat f, const BtRec* recs, int64_t nrec, const double* B, void* BT, int64_t bt_total,
               hipStream_t s)
{
  if (nrec <= 0) return;
  const int g = static_cast<int>(nrec < 65536 ? nrec : 65536);
  switch (f) {
    case BtFormat::F64T:
      hipLaunchKernelGGL(k_bt<double>, dim3(g), dim3(256), 0, s, recs, nrec, B, static_cast<double*>(BT));
      break;
    case BtFormat::F32T:
      hipLaunchKernelGGL(k_bt<float>, dim3(g), dim3(256), 0, s, recs, nrec, B, static_cast<float*>(BT));
      break;
    case BtFormat::BF16X3:
      hipLaunchKernelGGL(k_bt_bf3, dim3(g), dim3(256), 0, s, recs, nrec, B, static_cast<unsigned short*>(BT),
                         bt_total);
      break;
  }
  SHT_HIP(hipGetLastError());
}

void launch_scale(double* c, int64_t nelem, double beta, int beta_zero, hipStream_t s)
{
  if (nelem <= 0) return;
  const int64_t blocks = (nelem + 255) / 256;
  const int g = static_cast<int>(blocks < 16384 ? blocks : 16384);
  hipLaunchKernelGGL(k_scale, dim3(g), dim3(256), 0, s, c, nelem, beta, beta_zero);
  SHT_HIP(hipGetLastError());
}

}  // namespace sht
