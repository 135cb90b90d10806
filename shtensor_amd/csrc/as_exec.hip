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

#include <type_traits>

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
// FP64 / FP32 MFMA sweeps: R row tiles x NS slots, four k-steps (16 k) per
// loop iteration (panel rows are zero-padded to a multiple of 16).
//
// The A operand of a unit is one stream across all its sweeps: a ring of
// four register sets holds the A fragments of the next steps, filled three
// steps ahead, and the last iteration of a sweep already loads the first
// three steps of the next sweep (same pass, next slot group; or the next
// pass), so the A stream from HBM does not restart at sweep boundaries.  B
// is loaded one step ahead into two register sets.  Loads fill raw
// registers; the k < K mask and alpha are applied by the consumer after the
// scheduling barrier, so nothing waits for a load before the MFMAs that
// precede its use.  Every load is in bounds and unconditional: step indexes
// are clamped, A rows past M read on into the next column (tensor buffers
// carry kReadSlack bytes of slack), A columns past K re-read column K-1
// (masked to zero), and panel rows past K are zero.
// ---------------------------------------------------------------------------
struct AStream {
  const double* a;  // A block (wave-uniform)
  int K, M;
};

__device__ inline AStream a_stream(const AsArgs& p, const PassU& ps)
{
  AStream s;
  s.a = static_cast<const double*>(p.A) + ps.a_off;
  s.K = ps.k;
  s.M = ps.m;
  return s;
}

// A fragments of step st of the stream (rows r0 + 16t + li, column 4st + lk)
// A is read once per unit: non-temporal loads keep the stream from evicting
// the panels and C tiles that the other units of the block row reuse in L2.
template <int R>
__device__ inline void load_a(const AStream& s, int st, int row, int lk, double (&a)[R])
{
  const double* col = s.a + static_cast<int64_t>(min(4 * st + lk, s.K - 1)) * s.M + row;
#pragma unroll
  for (int t = 0; t < R; ++t) a[t] = __builtin_nontemporal_load(col + 16 * t);
}

// Step st of the current sweep's stream when st < nsteps, else step
// st - nsteps of the next sweep's (the selects are wave-uniform).
template <int R>
__device__ inline void load_a2(const AStream& cu, const AStream& nx, int nsteps, int st, int row, int lk,
                               double (&a)[R])
{
  const bool n = st >= nsteps;
  AStream s;
  s.a = n ? nx.a : cu.a;
  s.K = n ? nx.K : cu.K;
  s.M = n ? nx.M : cu.M;
  load_a(s, n ? st - nsteps : st, row, lk, a);
}

template <int F> struct BElem { typedef double type; };
template <> struct BElem<1> { typedef float type; };

// One sweep.  F = 0: FP64 MFMA, accumulators start from C (each C element is
// one fma chain in k order); F = 1: FP32 MFMA from zero, then added into C in
// FP64.  `raw_next` is the record of the pass after ps (loaded by the caller
// before this sweep); `next_pass` says whether the next sweep belongs to it.
// Returns the pass of the next sweep.
//
// The MFMA computes the transposed tile, D = Panel^T-fragment x A^T-fragment
// (products commute, so every element is the same fma chain): lane l then
// holds rows r0 + 16t + l%16 of four panel columns, so each C load and store
// instruction covers four columns x 16 consecutive rows (128 contiguous bytes
// per column).  The C offsets of the sweep's columns are staged in LDS.
template <int R, int NS, int F>
__device__ inline PassU sweep_pipe(const AsArgs& p, const PassU& ps, int g, int r0, int li, int lk,
                                   double (&ar)[4][R], const AsPass& raw_next, bool next_pass, int64_t* cbs)
{
  typedef typename BElem<F>::type TB;
  typedef typename std::conditional<F == 0, v4d, v4f>::type Acc;
  const TB* Bl = static_cast<const TB*>(p.B) + ps.bt_off + 16 * g + li + static_cast<int64_t>(lk) * ps.pitch;
  const int64_t bstep = 4 * static_cast<int64_t>(ps.pitch);
  const int K = ps.k;
  double* C = static_cast<double*>(p.C);
  // column c of the sweep held by this lane for accumulator element v
  auto colof = [&](int q, int v) { return F == 0 ? 16 * q + 4 * v + lk : 16 * q + 4 * lk + v; };
  {
    const int lane = threadIdx.x;
    if (lane < 16 * NS) cbs[lane] = col_base(p, ps, 16 * g + lane);
    if (lane + 64 < 16 * NS) cbs[lane + 64] = col_base(p, ps, 16 * g + lane + 64);
    __syncthreads();
  }
  const int rl = r0 + li;  // this lane's row in tile 0
  Acc acc[R][NS];
  if constexpr (F == 0) {
    // Accumulators start from C, all loads in flight at once (straight into
    // accumulation registers).  Pad columns read column 0 of C, rows past M
    // read on into the buffer slack; never stored.
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t cb = cbs[colof(q, v)];
        const double* cq = C + (cb >= 0 ? cb : 0) + rl;
#pragma unroll
        for (int t = 0; t < R; ++t) acc[t][q][v] = cq[16 * t];
      }
  } else {
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int q = 0; q < NS; ++q) acc[t][q] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  TB br[4][NS];
  auto loadB = [&](int st, TB(&b)[NS]) {
    const TB* bs = Bl + st * bstep;
#pragma unroll
    for (int q = 0; q < NS; ++q) b[q] = bs[16 * q];
  };
  const int nsteps = ((K + 15) >> 4) * 4, last = nsteps - 1;
  loadB(0, br[0]);
  loadB(1, br[1]);
  loadB(2, br[2]);
  __builtin_amdgcn_sched_barrier(0);
  // the next sweep's A stream (its record has arrived by now)
  PassU nps = ps;
  if (next_pass) {
    nps.a_off = uni(raw_next.a_off);
    nps.bt_off = uni(raw_next.bt_off);
    nps.k = uni(raw_next.k);
    nps.m = uni(raw_next.m);
    nps.ncol = uni(raw_next.ncol);
    nps.pitch = uni(raw_next.pitch);
    nps.col0 = uni(raw_next.col0);
    nps.cof0 = uni(raw_next.cof0);
  }
  const AStream cu = a_stream(p, ps), nx = a_stream(p, nps);
  auto mma = [&](int st, const double(&araw)[R], const TB(&b)[NS]) {
    const bool kin = 4 * st + lk < K;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if constexpr (F == 0) {
        const double y = kin ? araw[t] : 0.0;
        const double a = p.alpha_one ? y : y * p.alpha;
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[q], a, acc[t][q], 0, 0, 0);
      } else {
        const float a = kin ? static_cast<float>(araw[t]) : 0.f;
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[q], a, acc[t][q], 0, 0, 0);
      }
    }
  };
  // Both operands are loaded three steps ahead into rings of four register
  // sets.  B indexes are clamped to the last step; A indexes past it continue
  // into the next sweep's stream (the A ring then holds its steps 0..2).  In
  // each phase the B loads (mostly L2 hits) are issued before the A loads
  // (from HBM): the vector memory counter retires in issue order.
  const int row = r0 + li;
  for (int st = 0; st < nsteps; st += 4) {
    loadB(min(st + 3, last), br[3]);
    load_a2(cu, nx, nsteps, st + 3, row, lk, ar[3]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st, ar[0], br[0]);
    __builtin_amdgcn_sched_barrier(0);
    loadB(min(st + 4, last), br[0]);
    load_a2(cu, nx, nsteps, st + 4, row, lk, ar[0]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 1, ar[1], br[1]);
    __builtin_amdgcn_sched_barrier(0);
    loadB(min(st + 5, last), br[1]);
    load_a2(cu, nx, nsteps, st + 5, row, lk, ar[1]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 2, ar[2], br[2]);
    __builtin_amdgcn_sched_barrier(0);
    loadB(min(st + 6, last), br[2]);
    load_a2(cu, nx, nsteps, st + 6, row, lk, ar[2]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 3, ar[3], br[3]);
    __builtin_amdgcn_sched_barrier(0);
  }
  // write back
  const int M = ps.m;
  if constexpr (F == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t cb = cbs[colof(q, v)];
#pragma unroll
        for (int t = 0; t < R; ++t)
          if (cb >= 0 && rl + 16 * t < M) C[cb + rl + 16 * t] = acc[t][q][v];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    // C += alpha * (FP32 block product) in FP64, 4/R slots at a time (16
    // doubles per lane in flight).
    constexpr int G = NS < 4 / R ? NS : (4 / R > 0 ? 4 / R : 1);
#pragma unroll
    for (int q0 = 0; q0 < NS; q0 += G) {
      int64_t cb[G][4];
      double cv[G][4][R];
#pragma unroll
      for (int x = 0; x < G; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          cb[x][v] = q0 + x < NS ? cbs[colof(q0 + x, v)] : -1;
          const double* cq = C + (cb[x][v] >= 0 ? cb[x][v] : 0) + rl;
#pragma unroll
          for (int t = 0; t < R; ++t) cv[x][v][t] = cq[16 * t];
        }
#pragma unroll
      for (int x = 0; x < G; ++x) {
        if (q0 + x >= NS) break;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double* cq = C + (cb[x][v] >= 0 ? cb[x][v] : 0) + rl;
#pragma unroll
          for (int t = 0; t < R; ++t)
            if (cb[x][v] >= 0 && rl + 16 * t < M)
              cq[16 * t] = cv[x][v][t] + p.alpha * static_cast<double>(acc[t][q0 + x][v]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  __syncthreads();  // the column offsets are rewritten by the next sweep
  return nps;
}

// F: 0 = FP64 MFMA, 1 = FP32 MFMA.  One wave per unit of 16*R rows; sweeps
// of up to NSMAX slots.
template <int R, int NSMAX, int F>
__global__ __launch_bounds__(64) void k_asp(AsArgs p)
{
  if (static_cast<int64_t>(blockIdx.x) >= p.nunits) return;
  const AsUnit u = p.units[blockIdx.x];
  const int npass = uni(u.npass), pass0 = uni(u.pass0), r0 = uni(u.r0);
  if (npass == 0) return;
  const int li = threadIdx.x & 15, lk = threadIdx.x >> 4;
  const int pend = pass0 + npass;
  __shared__ int64_t cbs[16 * NSMAX];
  PassU ps = load_pass(p.passes, pass0);
  double ar[4][R];
  {
    const AStream s0 = a_stream(p, ps);
    load_a(s0, 0, r0 + li, lk, ar[0]);
    load_a(s0, 1, r0 + li, lk, ar[1]);
    load_a(s0, 2, r0 + li, lk, ar[2]);
  }
  int pi = pass0, g = 0;
  while (true) {
    const int nslots = (ps.ncol + 15) >> 4;
    const bool more_groups = g + NSMAX < nslots;
    const bool next_pass = !more_groups && pi + 1 < pend;
    const AsPass raw_next = p.passes[min(pi + 1, pend - 1)];
    PassU nps;
    switch (min(NSMAX, nslots - g)) {
      case 1: nps = sweep_pipe<R, 1, F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
      case 2: nps = sweep_pipe<R, cmin(2, NSMAX), F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
      case 3: nps = sweep_pipe<R, cmin(3, NSMAX), F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
      case 4: nps = sweep_pipe<R, cmin(4, NSMAX), F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
      case 5: nps = sweep_pipe<R, cmin(5, NSMAX), F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
      case 6: nps = sweep_pipe<R, cmin(6, NSMAX), F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
      case 7: nps = sweep_pipe<R, cmin(7, NSMAX), F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
      default: nps = sweep_pipe<R, cmin(8, NSMAX), F>(p, ps, g, r0, li, lk, ar, raw_next, next_pass, cbs); break;
    }
    if (more_groups) {
      g += NSMAX;
    } else if (next_pass) {
      ps = nps;
      ++pi;
      g = 0;
    } else {
      break;
    }
  }
}

// C += alpha * (FP32 block product) in FP64 for the split-BF16 sweep, 8/R
// slots at a time (32 doubles per lane in flight).  The f32 accumulator
// layout puts rows 4*(l/16) + v of column l%16 in lane l.
template <int R, int NS>
__device__ inline void as_lowp_epilogue(const AsArgs& p, const PassU& ps, int g, int r0, int lk, int li,
                                        const v4f (&acc)[R][NS])
{
  constexpr int G = NS < 8 / R ? NS : 8 / R;
  double* C = static_cast<double*>(p.C);
  const int M = ps.m;
#pragma unroll
  for (int q0 = 0; q0 < NS; q0 += G) {
    int64_t cb[G];
    double cv[G][R][4];
#pragma unroll
    for (int x = 0; x < G; ++x) {
      cb[x] = q0 + x < NS ? col_base(p, ps, 16 * (g + q0 + x) + li) : -1;
      const double* cq = C + (cb[x] >= 0 ? cb[x] : 0) + r0 + 4 * lk;
#pragma unroll
      for (int t = 0; t < R; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) cv[x][t][v] = cq[16 * t + v];
    }
#pragma unroll
    for (int x = 0; x < G; ++x) {
      if (q0 + x >= NS) break;
      double* cq = C + (cb[x] >= 0 ? cb[x] : 0) + r0 + 4 * lk;
#pragma unroll
      for (int t = 0; t < R; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (cb[x] >= 0 && r0 + 16 * t + 4 * lk + v < M)
            cq[16 * t + v] = cv[x][t][v] + p.alpha * static_cast<double>(acc[t][q0 + x][v]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
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
  static_assert(F == 2, "the MFMA forms run k_asp");
  sweep_bf3<R, NS>(p, ps, g, r0, nt, li, lk);
}

// Split-BF16: one wave per unit of 16*R rows; sweeps of up to NSMAX slots.
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
    const int krows = (rec.k + 15) / 16 * 16;  // zero rows up to a multiple of 16
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

int as_unit_rows(AsClass k)
{
  switch (k) {
    case AsClass::Exact:
    case AsClass::Bf16x3W: return kAsUnitRowsW;
    default: return kAsUnitRows;
  }
}

void launch_as(AsClass k, const AsArgs& a, hipStream_t s)
{
  if (a.nunits <= 0) return;
  const dim3 g(static_cast<unsigned>(a.nunits)), b(64);
  static_assert(kAsUnitRows == 32 && kAsUnitRowsW == 64, "unit rows: 2 and 4 MFMA tiles");
  switch (k) {
    case AsClass::Mfma64: hipLaunchKernelGGL((k_asp<2, 8, 0>), g, b, 0, s, a); break;
    case AsClass::Mfma64W: hipLaunchKernelGGL((k_asp<2, 4, 0>), g, b, 0, s, a); break;
    case AsClass::Mfma32: hipLaunchKernelGGL((k_asp<2, 8, 1>), g, b, 0, s, a); break;
    case AsClass::Mfma32W: hipLaunchKernelGGL((k_asp<2, 4, 1>), g, b, 0, s, a); break;
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
