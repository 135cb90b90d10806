// Execution kernels of a contraction plan (C-stationary grouped launches).
//
// Every kernel owns one tile of one C block and walks that block's tasks in
// ascending contracted-block key, so C needs no atomics and every element
// sees its contributions in one fixed order.  Operands are the canonical
// column-major matrices of each block: A (m x k), B (k x n), C (m x n).
//
//  * k_exact<NC>  — the reference JIT's per-element fma chain
//    (src/Kernel/KernelImpl.cpp:594-710) continued across tasks: bitwise equal
//    to the CPU chain.  One wave, 2 rows per lane (128 rows), NC columns held
//    in registers; B chunks staged in LDS and read as wave-wide broadcasts.
//  * k_mfma64<CT> — v_mfma_f64_16x16x4f64, 64 rows x 16*CT columns per wave,
//    accumulators persist across tasks (FP64 throughout).
//  * k_mfma32<CT> — v_mfma_f32_16x16x4f32 per task, task partial added into
//    an FP64 accumulator (FP64 across blocks).
//  * k_bf16x3<CT> — three-term split-BF16 (x = hi + mid + lo), six products
//    of order <= 2 with v_mfma_f32_16x16x32_bf16, FP32 per task, FP64 across.

#include "plan.hpp"

namespace sht {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));

namespace {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// bit-exact fma chain
// ---------------------------------------------------------------------------
template <int NC>
__global__ __launch_bounds__(64) void k_exact(ExecArgs p)
{
  constexpr int R = 2;
  constexpr int KC = 32;
  __shared__ double Bs[KC * NC];
  const TileRec tile = p.tiles[blockIdx.x];
  const CBlkRec cb = p.cblk[tile.cb];
  const int M = cb.m, N = cb.n, n0 = tile.n0;
  const int nc = min(NC, N - n0);
  const int lane = threadIdx.x;
  const double* A = static_cast<const double*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  double* Cb = static_cast<double*>(p.C) + cb.c_off;
  int row[R];
  bool ok[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    row[r] = tile.m0 + lane + r * kWave;
    ok[r] = row[r] < M;
  }
  double acc[R][NC];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int j = 0; j < NC; ++j)
      acc[r][j] = (!p.beta_zero && ok[r] && j < nc) ? Cb[row[r] + static_cast<int64_t>(n0 + j) * M] * p.beta : 0.0;
  if (!p.alpha_zero) {
    for (int t = cb.t0; t < cb.t1; ++t) {
      const TaskRec tr = p.tasks[t];
      const double* Ab = A + tr.a_off;
      const double* Bb = B + tr.b_off;
      const int K = tr.k;
      for (int k0 = 0; k0 < K; k0 += KC) {
        const int kc = min(KC, K - k0);
        __syncthreads();
        for (int e = lane; e < KC * NC; e += kWave) {
          const int kk = e % KC, j = e / KC;
          Bs[kk * NC + j] = (kk < kc && j < nc) ? Bb[(k0 + kk) + static_cast<int64_t>(n0 + j) * K] : 0.0;
        }
        __syncthreads();
#pragma unroll 2
        for (int kk = 0; kk < kc; ++kk) {
          double a[R];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            a[r] = ok[r] ? Ab[row[r] + static_cast<int64_t>(k0 + kk) * M] : 0.0;
            if (!p.alpha_one) a[r] = a[r] * p.alpha;
          }
#pragma unroll
          for (int j = 0; j < NC; ++j) {
            const double b = Bs[kk * NC + j];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r][j] = __fma_rn(a[r], b, acc[r][j]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int j = 0; j < NC; ++j)
      if (ok[r] && j < nc) Cb[row[r] + static_cast<int64_t>(n0 + j) * M] = acc[r][j];
}

// ---------------------------------------------------------------------------
// FP64 MFMA.  Operand layout of v_mfma_f64_16x16x4f64 (wave64, measured):
//   A frag: lane l holds A[l%16][l/16]; B frag: B[l/16][l%16];
//   D: lane l, item v holds D[4*v + l/16][l%16]  (the f32 form below differs:
//   D[4*(l/16) + v][l%16]).
// ---------------------------------------------------------------------------
template <int CT>
__global__ __launch_bounds__(64) void k_mfma64(ExecArgs p)
{
  constexpr int RT = 4;
  const TileRec tile = p.tiles[blockIdx.x];
  const CBlkRec cb = p.cblk[tile.cb];
  const int M = cb.m, N = cb.n;
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = lane >> 4;
  const double* A = static_cast<const double*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  double* Cb = static_cast<double*>(p.C) + cb.c_off;
  v4d acc[RT][CT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + 16 * rt + 4 * v + lk;
        const int j = tile.n0 + 16 * ct + li;
        acc[rt][ct][v] = (!p.beta_zero && i < M && j < N) ? Cb[i + static_cast<int64_t>(j) * M] * p.beta : 0.0;
      }
  if (!p.alpha_zero) {
    for (int t = cb.t0; t < cb.t1; ++t) {
      const TaskRec tr = p.tasks[t];
      const double* Ab = A + tr.a_off;
      const double* Bb = B + tr.b_off;
      const int K = tr.k;
      for (int k = 0; k < K; k += 4) {
        const int kk = k + lk;
        const bool kin = kk < K;
        double a[RT], b[CT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          const int i = tile.m0 + 16 * rt + li;
          a[rt] = (kin && i < M) ? Ab[i + static_cast<int64_t>(kk) * M] : 0.0;
          if (!p.alpha_one) a[rt] *= p.alpha;
        }
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const int j = tile.n0 + 16 * ct + li;
          b[ct] = (kin && j < N) ? Bb[kk + static_cast<int64_t>(j) * K] : 0.0;
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int ct = 0; ct < CT; ++ct)
            acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + 16 * rt + 4 * v + lk;
        const int j = tile.n0 + 16 * ct + li;
        if (i < M && j < N) Cb[i + static_cast<int64_t>(j) * M] = acc[rt][ct][v];
      }
}

// ---------------------------------------------------------------------------
// FP32 MFMA per task, FP64 across tasks.  v_mfma_f32_16x16x4f32 has the same
// operand/result layout as the f64 variant above.
// ---------------------------------------------------------------------------
template <int CT>
__global__ __launch_bounds__(64) void k_mfma32(ExecArgs p)
{
  constexpr int RT = 4;
  const TileRec tile = p.tiles[blockIdx.x];
  const CBlkRec cb = p.cblk[tile.cb];
  const int M = cb.m, N = cb.n;
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = lane >> 4;
  const double* A = static_cast<const double*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  double* Cb = static_cast<double*>(p.C) + cb.c_off;
  v4d acc[RT][CT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + 16 * rt + 4 * lk + v;
        const int j = tile.n0 + 16 * ct + li;
        acc[rt][ct][v] = (!p.beta_zero && i < M && j < N) ? Cb[i + static_cast<int64_t>(j) * M] * p.beta : 0.0;
      }
  if (!p.alpha_zero) {
    for (int t = cb.t0; t < cb.t1; ++t) {
      const TaskRec tr = p.tasks[t];
      const double* Ab = A + tr.a_off;
      const double* Bb = B + tr.b_off;
      const int K = tr.k;
      v4f part[RT][CT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) part[rt][ct] = v4f{0.f, 0.f, 0.f, 0.f};
      for (int k = 0; k < K; k += 4) {
        const int kk = k + lk;
        const bool kin = kk < K;
        float a[RT], b[CT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          const int i = tile.m0 + 16 * rt + li;
          a[rt] = (kin && i < M) ? static_cast<float>(Ab[i + static_cast<int64_t>(kk) * M]) : 0.f;
        }
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const int j = tile.n0 + 16 * ct + li;
          b[ct] = (kin && j < N) ? static_cast<float>(Bb[kk + static_cast<int64_t>(j) * K]) : 0.f;
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int ct = 0; ct < CT; ++ct)
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt], b[ct], part[rt][ct], 0, 0, 0);
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[rt][ct][v] += p.alpha * static_cast<double>(part[rt][ct][v]);
    }
  }
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + 16 * rt + 4 * lk + v;
        const int j = tile.n0 + 16 * ct + li;
        if (i < M && j < N) Cb[i + static_cast<int64_t>(j) * M] = acc[rt][ct][v];
      }
}

// ---------------------------------------------------------------------------
// Three-term split-BF16 (C-stationary).  v_mfma_f32_16x16x32_bf16 (wave64):
// A frag lane l holds A[l%16][8*(l/16) .. 8*(l/16)+7]; B frag holds
// B[8*(l/16)..+7][l%16]; D as the f32 form above.  x = hi + mid + lo (bf16
// each, split3 below); the six products of order <= 2 accumulate in FP32
// per task (smallest first), FP64 across tasks.
// ---------------------------------------------------------------------------
__device__ inline unsigned short bf16_bits_cs(float x)
{
  const unsigned int u = __float_as_uint(x);
  return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ inline void split3_cs(double xd, short& h, short& m, short& l)
{
  const float x = static_cast<float>(xd);
  const unsigned short hb = bf16_bits_cs(x);
  const float r1 = x - __uint_as_float(static_cast<unsigned int>(hb) << 16);
  const unsigned short mb = bf16_bits_cs(r1);
  const float r2 = r1 - __uint_as_float(static_cast<unsigned int>(mb) << 16);
  h = static_cast<short>(hb);
  m = static_cast<short>(mb);
  l = static_cast<short>(bf16_bits_cs(r2));
}

template <int CT>
__global__ __launch_bounds__(64) void k_bf16x3(ExecArgs p)
{
  constexpr int RT = 4;
  const TileRec tile = p.tiles[blockIdx.x];
  const CBlkRec cb = p.cblk[tile.cb];
  const int M = cb.m, N = cb.n;
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = lane >> 4;
  const double* A = static_cast<const double*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  double* Cb = static_cast<double*>(p.C) + cb.c_off;
  v4d acc[RT][CT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + 16 * rt + 4 * lk + v;
        const int j = tile.n0 + 16 * ct + li;
        acc[rt][ct][v] = (!p.beta_zero && i < M && j < N) ? Cb[i + static_cast<int64_t>(j) * M] * p.beta : 0.0;
      }
  if (!p.alpha_zero) {
    for (int t = cb.t0; t < cb.t1; ++t) {
      const TaskRec tr = p.tasks[t];
      const double* Ab = A + tr.a_off;
      const double* Bb = B + tr.b_off;
      const int K = tr.k;
      v4f part[RT][CT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) part[rt][ct] = v4f{0.f, 0.f, 0.f, 0.f};
      for (int k = 0; k < K; k += 32) {
        v8s bh[CT], bm[CT], bl[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const int j = tile.n0 + 16 * ct + li;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int kk = k + 8 * lk + e;
            short h = 0, m = 0, l = 0;
            if (kk < K && j < N) split3_cs(Bb[kk + static_cast<int64_t>(j) * K], h, m, l);
            bh[ct][e] = h;
            bm[ct][e] = m;
            bl[ct][e] = l;
          }
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          const int i = tile.m0 + 16 * rt + li;
          v8s ah, am, al;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int kk = k + 8 * lk + e;
            short h = 0, m = 0, l = 0;
            if (kk < K && i < M) split3_cs(Ab[i + static_cast<int64_t>(kk) * M], h, m, l);
            ah[e] = h;
            am[e] = m;
            al[e] = l;
          }
#pragma unroll
          for (int ct = 0; ct < CT; ++ct) {
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[ct], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm[ct], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[ct], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh[ct], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm[ct], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[ct], part[rt][ct], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[rt][ct][v] += p.alpha * static_cast<double>(part[rt][ct][v]);
    }
  }
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + 16 * rt + 4 * lk + v;
        const int j = tile.n0 + 16 * ct + li;
        if (i < M && j < N) Cb[i + static_cast<int64_t>(j) * M] = acc[rt][ct][v];
      }
}

// ---------------------------------------------------------------------------
// A-stationary FP64 MFMA.  Workgroup = 4 waves = one 64-row chunk of a block
// row; wave w owns rows 16w..16w+15 and walks every pass (A block) of the
// row in ascending K.  Per pass it sweeps up to kAsSlots 16-column slots at
// once: the A fragment (16 rows x 4 k) is loaded once per k-step and feeds
// every slot's MFMA, so each A element leaves HBM once.  The four waves read
// the same B fragments close together in time (L1 reuse).  C tiles are read,
// accumulated over the pass and written back (C pre-scaled by beta).
// ---------------------------------------------------------------------------
template <int S>
__device__ inline void as_sweep(const AsArgs& p, const AsPass& ps, int s_begin, int ns, int r, int li, int lk)
{
  // A fragments are streamed from HBM D k-steps ahead of use (one step's
  // MFMAs do not cover HBM latency); B fragments come from the transposed,
  // zero-padded blocks (L2-resident) one step ahead.  Per-slot state is kept
  // to one B pointer and a pitch; C tile metadata is re-read at the end.
  constexpr int D = 4;
  const double* A = static_cast<const double*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  double* C = static_cast<double*>(p.C);
  const int K = ps.k, M = ps.m;
  const int ia = r + li;
  const bool rowok = ia < M;
  const double* Arow = A + ps.a_off + ia;
  const double* bp[S];
  int pitch[S];
  v4d acc[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const AsSlot sl = p.slots[s_begin + (q < ns ? q : 0)];
    pitch[q] = (sl.n + kBtPad - 1) / kBtPad * kBtPad;
    bp[q] = B + sl.bt_off + sl.n0 + li;
    const int j = sl.n0 + li;
    const bool colok = q < ns && j < sl.n;
    const double* cq = C + sl.c_off + (r + lk) + static_cast<int64_t>(colok ? j : 0) * M;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[q][v] = (colok && r + 4 * v + lk < M) ? cq[4 * v] : 0.0;
  }
  double a_cur[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int kk = 4 * d + lk;
    a_cur[d] = (kk < K && rowok) ? Arow[static_cast<int64_t>(kk) * M] : 0.0;
  }
  double b_nx[S];
#pragma unroll
  for (int q = 0; q < S; ++q) b_nx[q] = (q < ns && lk < K) ? bp[q][static_cast<int64_t>(lk) * pitch[q]] : 0.0;
  const int nsteps = (K + 3) >> 2;
  for (int c = 0; c < nsteps; c += D) {
    double a_nxt[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int kk = 4 * (c + D + d) + lk;
      a_nxt[d] = (kk < K && rowok) ? Arow[static_cast<int64_t>(kk) * M] : 0.0;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int step = c + d;
      if (step < nsteps) {
        double a = a_cur[d];
        if (!p.alpha_one) a *= p.alpha;
        double b[S];
#pragma unroll
        for (int q = 0; q < S; ++q) b[q] = b_nx[q];
        const int kn = 4 * (step + 1) + lk;
#pragma unroll
        for (int q = 0; q < S; ++q)
          if (q < ns) b_nx[q] = kn < K ? bp[q][static_cast<int64_t>(kn) * pitch[q]] : 0.0;
#pragma unroll
        for (int q = 0; q < S; ++q)
          if (q < ns) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[q], acc[q], 0, 0, 0);
      }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) a_cur[d] = a_nxt[d];
  }
#pragma unroll
  for (int q = 0; q < S; ++q) {
    if (q < ns) {
      const AsSlot sl = p.slots[s_begin + q];
      const int j = sl.n0 + li;
      if (j < sl.n) {
        double* cq = C + sl.c_off + (r + lk) + static_cast<int64_t>(j) * M;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (r + 4 * v + lk < M) cq[4 * v] = acc[q][v];
      }
    }
  }
}

// Lean sweep for the one-wave units: B fragments are loaded right before use
// (residency, not prefetch, hides their L2 latency), A streams D steps ahead.
// Sweep with a compile-time slot count NS (dispatched by a switch).  The k
// loop is unrolled by two with two register sets, so the loads of step s+1
// are in flight while the MFMAs of step s issue, and every branch in it is
// wave-uniform: B^T rows are zero-padded to a multiple of 4 in k, A rows past
// M read row M-1 (their outputs are never stored) and A columns past K read
// column K-1 masked to zero.
__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline int64_t uni(int64_t x)
{
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(x) >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

template <int NS>
__device__ inline void as_sweep_lean(const AsArgs& p, const AsPass& ps, int s_begin, int r, int li, int lk)
{
  const double* A = static_cast<const double*>(p.A);
  const double* B = static_cast<const double*>(p.B);
  double* C = static_cast<double*>(p.C);
  const int K = ps.k, M = ps.m;
  const double* abase = A + ps.a_off;
  const int arow = min(r + li, M - 1);
  const double* bbase[NS];  // wave-uniform
  int64_t bstep[NS];        // wave-uniform
  int boff[NS];             // per lane
  v4d acc[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const AsSlot sl = p.slots[s_begin + q];
    const int n = uni(sl.n), n0 = uni(sl.n0);
    const int pitch = (n + kBtPad - 1) / kBtPad * kBtPad;
    bbase[q] = B + uni(sl.bt_off) + n0;
    bstep[q] = 4 * static_cast<int64_t>(pitch);
    boff[q] = li + lk * pitch;
    const int j = n0 + li;
    const bool colok = j < n;
    const double* cq = C + sl.c_off + (r + lk) + static_cast<int64_t>(colok ? j : 0) * M;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[q][v] = (colok && r + 4 * v + lk < M) ? cq[4 * v] : 0.0;
  }
  // Step indexes are clamped to the last step, so every load is in bounds and
  // unconditional; the values of a clamped load are never used.
  const int last = ((K + 3) >> 2) - 1;
  auto load_a = [&](int st) -> double {
    const int kk = 4 * st + lk;
    const double x = abase[arow + static_cast<int64_t>(min(kk, K - 1)) * M];
    const double y = kk < K ? x : 0.0;
    return p.alpha_one ? y : y * p.alpha;
  };
  double a0, a1, b0[NS], b1[NS];
  a0 = load_a(0);
#pragma unroll
  for (int q = 0; q < NS; ++q) b0[q] = bbase[q][boff[q]];
  for (int st = 0; st <= last; st += 2) {
    const int s1 = min(st + 1, last);
    a1 = load_a(s1);
#pragma unroll
    for (int q = 0; q < NS; ++q) b1[q] = (bbase[q] + s1 * bstep[q])[boff[q]];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0[q], acc[q], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    if (st + 1 > last) break;
    const int s2 = min(st + 2, last);
    a0 = load_a(s2);
#pragma unroll
    for (int q = 0; q < NS; ++q) b0[q] = (bbase[q] + s2 * bstep[q])[boff[q]];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1[q], acc[q], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const AsSlot sl = p.slots[s_begin + q];
    const int j = sl.n0 + li;
    if (j < sl.n) {
      double* cq = C + sl.c_off + (r + lk) + static_cast<int64_t>(j) * M;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (r + 4 * v + lk < M) cq[4 * v] = acc[q][v];
    }
  }
}

__device__ inline void as_sweep_switch(const AsArgs& p, const AsPass& ps, int s_begin, int ns, int r, int li, int lk)
{
  switch (ns) {
    case 1: as_sweep_lean<1>(p, ps, s_begin, r, li, lk); break;
    case 2: as_sweep_lean<2>(p, ps, s_begin, r, li, lk); break;
    case 3: as_sweep_lean<3>(p, ps, s_begin, r, li, lk); break;
    case 4: as_sweep_lean<4>(p, ps, s_begin, r, li, lk); break;
    case 5: as_sweep_lean<5>(p, ps, s_begin, r, li, lk); break;
    case 6: as_sweep_lean<6>(p, ps, s_begin, r, li, lk); break;
    case 7: as_sweep_lean<7>(p, ps, s_begin, r, li, lk); break;
    default: as_sweep_lean<8>(p, ps, s_begin, r, li, lk); break;
  }
}

// WPG waves per workgroup; wave w of workgroup g owns unit WPG*g + w (a
// 16-row chunk of one block row); empty units (npass == 0) are padding.
template <int S, int WPG, bool LEAN>
__global__ __launch_bounds__(64 * WPG) void k_as_mfma64(AsArgs p)
{
  const int wave = WPG == 1 ? 0 : uni(static_cast<int>(threadIdx.x >> 6));
  const int64_t ui = static_cast<int64_t>(blockIdx.x) * WPG + wave;
  if (ui >= p.nunits) return;
  const AsUnit u = p.units[ui];
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int npass = uni(u.npass), pass0 = uni(u.pass0), r0 = uni(u.r0);
  for (int pi = pass0; pi < pass0 + npass; ++pi) {
    const AsPass pv = p.passes[pi];
    AsPass ps;
    ps.a_off = uni(pv.a_off);
    ps.k = uni(pv.k);
    ps.m = uni(pv.m);
    ps.s0 = uni(pv.s0);
    ps.ns = uni(pv.ns);
    for (int g = 0; g < ps.ns; g += S) {
      if (LEAN) as_sweep_switch(p, ps, ps.s0 + g, min(S, ps.ns - g), r0, li, lk);
      else as_sweep<S>(p, ps, ps.s0 + g, min(S, ps.ns - g), r0, li, lk);
    }
  }
}

// ---------------------------------------------------------------------------
// A-stationary fma chain (bitwise equal to the oracle).  One wave per unit,
// lane = one C row of the 64-row chunk; per pass, up to kAsxSlots slots of
// kAsxCols columns keep their C values in registers while k runs ascending:
// c = fma(alpha*A(i,k), B(k,j), c).  B(k, n0..n0+7) is wave-uniform and
// contiguous in the transposed block, so it arrives by scalar loads.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_as_exact(AsArgs p)
{
  constexpr int S = kAsxSlots, W = kAsxCols;
  const AsUnit u = p.units[blockIdx.x];
  const int lane = threadIdx.x;
  if (u.npass == 0) return;
  const double* A = static_cast<const double*>(p.A);
  const double* BT = static_cast<const double*>(p.B);
  double* C = static_cast<double*>(p.C);
  const int row = u.r0 + lane;
  for (int pi = u.pass0; pi < u.pass0 + u.npass; ++pi) {
    const AsPass ps = p.passes[pi];
    const int K = ps.k, M = ps.m;
    const bool rowok = row < M;
    const double* Ab = A + ps.a_off + row;
    for (int g = 0; g < ps.ns; g += S) {
      const int ns = min(S, ps.ns - g);
      const double* bt[S];
      int pitch[S];
      double acc[S][W];
#pragma unroll
      for (int q = 0; q < S; ++q) {
        const AsSlot sl = p.slots[ps.s0 + g + (q < ns ? q : 0)];
        pitch[q] = (sl.n + kBtPad - 1) / kBtPad * kBtPad;
        bt[q] = BT + sl.bt_off + sl.n0;
#pragma unroll
        for (int j = 0; j < W; ++j)
          acc[q][j] = (q < ns && rowok && sl.n0 + j < sl.n) ? C[sl.c_off + row + static_cast<int64_t>(sl.n0 + j) * M]
                                                           : 0.0;
      }
      for (int k = 0; k < K; ++k) {
        double a = rowok ? Ab[static_cast<int64_t>(k) * M] : 0.0;
        if (!p.alpha_one) a = a * p.alpha;
#pragma unroll
        for (int q = 0; q < S; ++q) {
          if (q < ns) {
            const double* b = bt[q] + static_cast<int64_t>(k) * pitch[q];
#pragma unroll
            for (int j = 0; j < W; ++j) acc[q][j] = __fma_rn(a, b[j], acc[q][j]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < S; ++q) {
        if (q < ns && rowok) {
          const AsSlot sl = p.slots[ps.s0 + g + q];
#pragma unroll
          for (int j = 0; j < W; ++j)
            if (sl.n0 + j < sl.n) C[sl.c_off + row + static_cast<int64_t>(sl.n0 + j) * M] = acc[q][j];
        }
      }
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_bt(const BtRec* recs, int64_t nrec, const double* B, T* BT)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const BtRec rec = recs[r];
    const int npad = (rec.n + kBtPad - 1) / kBtPad * kBtPad;
    const int krows = (rec.k + 3) / 4 * 4;  // zero rows up to a multiple of 4
    const int64_t tot = static_cast<int64_t>(krows) * npad;
    const double* b = B + rec.b_off;
    T* o = BT + rec.bt_off;
    for (int64_t x = threadIdx.x; x < tot; x += blockDim.x) {
      const int k = static_cast<int>(x / npad), j = static_cast<int>(x - static_cast<int64_t>(k) * npad);
      o[x] = (j < rec.n && k < rec.k) ? static_cast<T>(b[k + static_cast<int64_t>(j) * rec.k]) : T(0);
    }
  }
}

__device__ inline unsigned short bf16_bits(float x)
{
  const unsigned int u = __float_as_uint(x);
  return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// x (as float) = hi + mid + lo, each a bf16 (round to nearest even)
__device__ inline void split3(double xd, unsigned short& h, unsigned short& m, unsigned short& l)
{
  const float x = static_cast<float>(xd);
  h = bf16_bits(x);
  const float r1 = x - __uint_as_float(static_cast<unsigned int>(h) << 16);
  m = bf16_bits(r1);
  const float r2 = r1 - __uint_as_float(static_cast<unsigned int>(m) << 16);
  l = bf16_bits(r2);
}

__global__ __launch_bounds__(256) void k_bt_bf3(const BtRec* recs, int64_t nrec, const double* B, unsigned short* BT,
                                                int64_t plane)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const BtRec rec = recs[r];
    const int npad = (rec.n + kBtPad - 1) / kBtPad * kBtPad;
    const int kpad = (rec.k + kBfK - 1) / kBfK * kBfK;
    const int64_t tot = static_cast<int64_t>(kpad) * npad;
    const double* b = B + rec.b_off;
    unsigned short* o = BT + rec.bt_off;
    for (int64_t x = threadIdx.x; x < tot; x += blockDim.x) {
      const int j = static_cast<int>(x / kpad), k = static_cast<int>(x - static_cast<int64_t>(j) * kpad);
      unsigned short h = 0, m = 0, l = 0;
      if (j < rec.n && k < rec.k) split3(b[k + static_cast<int64_t>(j) * rec.k], h, m, l);
      o[x] = h;
      o[x + plane] = m;
      o[x + 2 * plane] = l;
    }
  }
}

// ---------------------------------------------------------------------------
// A-stationary FP32 MFMA: per (pass, slot) the block product A_blk * B_blk is
// accumulated in FP32 (v_mfma_f32_16x16x4f32, D[4*(l/16)+v][l%16]) and then
// added to the FP64 C tile: FP32 within a block product, FP64 across blocks.
// ---------------------------------------------------------------------------
template <int S>
__device__ inline void as_sweep_f32(const AsArgs& p, const AsPass& ps, int s_begin, int ns, int r, int li, int lk)
{
  constexpr int D = 4;
  const double* A = static_cast<const double*>(p.A);
  const float* B = static_cast<const float*>(p.B);
  double* C = static_cast<double*>(p.C);
  const int K = ps.k, M = ps.m;
  const int ia = r + li;
  const bool rowok = ia < M;
  const double* Arow = A + ps.a_off + ia;
  const float* bp[S];
  int pitch[S];
  v4f acc[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const AsSlot sl = p.slots[s_begin + (q < ns ? q : 0)];
    pitch[q] = (sl.n + kBtPad - 1) / kBtPad * kBtPad;
    bp[q] = B + sl.bt_off + sl.n0 + li;
    acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  float a_cur[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int kk = 4 * d + lk;
    a_cur[d] = (kk < K && rowok) ? static_cast<float>(Arow[static_cast<int64_t>(kk) * M]) : 0.f;
  }
  float b_nx[S];
#pragma unroll
  for (int q = 0; q < S; ++q) b_nx[q] = (q < ns && lk < K) ? bp[q][static_cast<int64_t>(lk) * pitch[q]] : 0.f;
  const int nsteps = (K + 3) >> 2;
  for (int c = 0; c < nsteps; c += D) {
    float a_nxt[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int kk = 4 * (c + D + d) + lk;
      a_nxt[d] = (kk < K && rowok) ? static_cast<float>(Arow[static_cast<int64_t>(kk) * M]) : 0.f;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int step = c + d;
      if (step < nsteps) {
        float b[S];
#pragma unroll
        for (int q = 0; q < S; ++q) b[q] = b_nx[q];
        const int kn = 4 * (step + 1) + lk;
#pragma unroll
        for (int q = 0; q < S; ++q)
          if (q < ns) b_nx[q] = kn < K ? bp[q][static_cast<int64_t>(kn) * pitch[q]] : 0.f;
#pragma unroll
        for (int q = 0; q < S; ++q)
          if (q < ns) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_cur[d], b[q], acc[q], 0, 0, 0);
      }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) a_cur[d] = a_nxt[d];
  }
#pragma unroll
  for (int q = 0; q < S; ++q) {
    if (q < ns) {
      const AsSlot sl = p.slots[s_begin + q];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = r + 4 * lk + v;
        const int j = sl.n0 + li;
        if (i < M && j < sl.n) {
          double* c = C + sl.c_off + i + static_cast<int64_t>(j) * M;
          *c = *c + p.alpha * static_cast<double>(acc[q][v]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// A-stationary three-term split-BF16: x = hi + mid + lo (bf16 each, about
// 24 significant bits); per 32-k chunk six v_mfma_f32_16x16x32_bf16 products
// of order <= 2 (lo*hi, mid*mid, hi*lo, mid*hi, hi*mid, hi*hi, smallest
// first) accumulate in FP32 per block product, then add into the FP64 C tile.
// A is split on the fly; B was split once per execution (BtFormat::BF16X3).
// ---------------------------------------------------------------------------
template <int S>
__device__ inline void as_sweep_bf3(const AsArgs& p, const AsPass& ps, int s_begin, int ns, int r, int li, int lk)
{
  const double* A = static_cast<const double*>(p.A);
  const unsigned short* B = static_cast<const unsigned short*>(p.B);
  const int64_t plane = p.bt_total;
  double* C = static_cast<double*>(p.C);
  const int K = ps.k, M = ps.m;
  const int ia = r + li;
  const bool rowok = ia < M;
  const double* Arow = A + ps.a_off + ia;
  const int kpad = (K + kBfK - 1) / kBfK * kBfK;
  const unsigned short* bp[S];
  v4f acc[S];
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const AsSlot sl = p.slots[s_begin + (q < ns ? q : 0)];
    bp[q] = B + sl.bt_off + static_cast<int64_t>(sl.n0 + li) * kpad + 8 * lk;
    acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  for (int k0 = 0; k0 < K; k0 += kBfK) {
    v8s ah, am, al;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int kk = k0 + 8 * lk + e;
      unsigned short h = 0, m = 0, l = 0;
      if (kk < K && rowok) split3(Arow[static_cast<int64_t>(kk) * M], h, m, l);
      ah[e] = static_cast<short>(h);
      am[e] = static_cast<short>(m);
      al[e] = static_cast<short>(l);
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      if (q < ns) {
        const v8s bh = *reinterpret_cast<const v8s*>(bp[q] + k0);
        const v8s bm = *reinterpret_cast<const v8s*>(bp[q] + k0 + plane);
        const v8s bl = *reinterpret_cast<const v8s*>(bp[q] + k0 + 2 * plane);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[q], 0, 0, 0);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc[q], 0, 0, 0);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[q], 0, 0, 0);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc[q], 0, 0, 0);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc[q], 0, 0, 0);
        acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[q], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < S; ++q) {
    if (q < ns) {
      const AsSlot sl = p.slots[s_begin + q];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = r + 4 * lk + v;
        const int j = sl.n0 + li;
        if (i < M && j < sl.n) {
          double* c = C + sl.c_off + i + static_cast<int64_t>(j) * M;
          *c = *c + p.alpha * static_cast<double>(acc[q][v]);
        }
      }
    }
  }
}

template <int S, int F>
__global__ __launch_bounds__(256) void k_as_lowp(AsArgs p)
{
  const int wave = threadIdx.x >> 6;
  const int64_t ui = static_cast<int64_t>(blockIdx.x) * 4 + wave;
  if (ui >= p.nunits) return;
  const AsUnit u = p.units[ui];
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int r = u.r0;
  if (u.npass == 0) return;
  for (int pi = u.pass0; pi < u.pass0 + u.npass; ++pi) {
    const AsPass ps = p.passes[pi];
    for (int g = 0; g < ps.ns; g += S) {
      if (F == 0) as_sweep_f32<S>(p, ps, ps.s0 + g, min(S, ps.ns - g), r, li, lk);
      else as_sweep_bf3<S>(p, ps, ps.s0 + g, min(S, ps.ns - g), r, li, lk);
    }
  }
}

__global__ __launch_bounds__(256) void k_scale(double* c, int64_t n, double beta, int bz)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    c[i] = bz ? 0.0 : c[i] * beta;
}

// ---------------------------------------------------------------------------
// block permutation (Utils::reshape per block)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_permute(PermArgs p)
{
  const TensorDesc& d = p.d;
  for (int64_t q = blockIdx.x; q < d.nnz; q += gridDim.x) {
    int64_t id = d.indices[q];
    int dims[kMaxModes];
    for (int m = d.N - 1; m >= 0; --m) {
      const int c = static_cast<int>(id / d.bstride[m]);
      id -= static_cast<int64_t>(c) * d.bstride[m];
      dims[m] = d.bsizes[d.boff[m] + c];
    }
    int64_t sstride[kMaxModes];
    int64_t vol = 1;
    for (int m = 0; m < d.N; ++m) {
      sstride[m] = vol;
      vol *= dims[m];
    }
    const T* in = static_cast<const T*>(p.in) + d.offsets[q];
    T* out = static_cast<T*>(p.out) + d.offsets[q];
    for (int64_t x = threadIdx.x; x < vol; x += blockDim.x) {
      // x indexes the canonical layout (dims reordered by `order`)
      int64_t rem = x, src = 0;
      for (int i = 0; i < d.N; ++i) {
        const int ext = dims[p.order[i]];
        const int64_t ci = rem % ext;
        rem /= ext;
        src += ci * sstride[p.order[i]];
      }
      if (p.inverse) out[src] = in[x];
      else out[x] = in[src];
    }
  }
}

}  // namespace

int kclass_rows(KClass k)
{
  switch (k) {
    case KClass::Exact8:
    case KClass::Exact16:
    case KClass::Exact32: return 128;
    default: return 64;
  }
}

int kclass_cols(KClass k)
{
  switch (k) {
    case KClass::Exact8: return 8;
    case KClass::Exact16: return 16;
    case KClass::Exact32: return 32;
    case KClass::Mfma16:
    case KClass::F32x16:
    case KClass::Bf16x16: return 16;
    default: return 32;
  }
}

void launch_exec(KClass k, const ExecArgs& a, hipStream_t s)
{
  if (a.ntiles <= 0) return;
  const dim3 g(static_cast<unsigned>(a.ntiles)), b(kWave);
  switch (k) {
    case KClass::Exact8: hipLaunchKernelGGL(k_exact<8>, g, b, 0, s, a); break;
    case KClass::Exact16: hipLaunchKernelGGL(k_exact<16>, g, b, 0, s, a); break;
    case KClass::Exact32: hipLaunchKernelGGL(k_exact<32>, g, b, 0, s, a); break;
    case KClass::Mfma16: hipLaunchKernelGGL(k_mfma64<1>, g, b, 0, s, a); break;
    case KClass::Mfma32: hipLaunchKernelGGL(k_mfma64<2>, g, b, 0, s, a); break;
    case KClass::F32x16: hipLaunchKernelGGL(k_mfma32<1>, g, b, 0, s, a); break;
    case KClass::F32x32: hipLaunchKernelGGL(k_mfma32<2>, g, b, 0, s, a); break;
    case KClass::Bf16x16: hipLaunchKernelGGL(k_bf16x3<1>, g, b, 0, s, a); break;
    case KClass::Bf16x32: hipLaunchKernelGGL(k_bf16x3<2>, g, b, 0, s, a); break;
    default: throw Error(SHT_ERR_INVALID, "unknown kernel class");
  }
  SHT_HIP(hipGetLastError());
}

void launch_as(AsClass k, const AsArgs& a, hipStream_t s)
{
  if (a.nunits <= 0) return;
  const dim3 g4(static_cast<unsigned>((a.nunits + 3) / 4)), g1(static_cast<unsigned>(a.nunits));
  static_assert(kAsRows == 16, "units are 16-row chunks");
  switch (k) {
    case AsClass::Mfma64: hipLaunchKernelGGL((k_as_mfma64<kAsSlots, 4, false>), g4, dim3(4 * kWave), 0, s, a); break;
    case AsClass::Mfma64s4: hipLaunchKernelGGL((k_as_mfma64<kAsSlots, 1, true>), g1, dim3(kWave), 0, s, a); break;
    case AsClass::Exact: hipLaunchKernelGGL(k_as_exact, g1, dim3(kWave), 0, s, a); break;
    case AsClass::Mfma32: hipLaunchKernelGGL((k_as_lowp<kAsSlots, 0>), g4, dim3(4 * kWave), 0, s, a); break;
    case AsClass::Bf16x3: hipLaunchKernelGGL((k_as_lowp<kAsSlots, 1>), g4, dim3(4 * kWave), 0, s, a); break;
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
    case BtFormat::F64T: hipLaunchKernelGGL(k_bt<double>, dim3(g), dim3(256), 0, s, recs, nrec, B,
                                            static_cast<double*>(BT)); break;
    case BtFormat::F32T: hipLaunchKernelGGL(k_bt<float>, dim3(g), dim3(256), 0, s, recs, nrec, B,
                                            static_cast<float*>(BT)); break;
    case BtFormat::BF16X3: hipLaunchKernelGGL(k_bt_bf3, dim3(g), dim3(256), 0, s, recs, nrec, B,
                                              static_cast<unsigned short*>(BT), bt_total); break;
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

void launch_permute(const PermArgs& p, bool f64, hipStream_t s)
{
  if (p.d.nnz <= 0) return;
  const int g = static_cast<int>(p.d.nnz < 65536 ? p.d.nnz : 65536);
  if (f64) hipLaunchKernelGGL(k_permute<double>, dim3(g), dim3(256), 0, s, p);
  else hipLaunchKernelGGL(k_permute<float>, dim3(g), dim3(256), 0, s, p);
  SHT_HIP(hipGetLastError());
}

}  // namespace sht
