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
//  * k_mfma64<CT,RT> — v_mfma_f64_16x16x4f64, 16*RT rows x 16*CT columns per
//    wave, accumulators persist across tasks (FP64 throughout).
//  * k_mfma32<CT,RT> — v_mfma_f32_16x16x4f32 per task, task partial added
//    into an FP64 accumulator (FP64 across blocks).
//  * k_mfmaw<F,TR,TC> — FP64 (F = 0) or FP32-per-task (F = 1) MFMA, four
//    waves per TR x TC tile sharing every A and B chunk through LDS (64 x 64:
//    C blocks more than 32 columns wide; 32 x 32: a block's last row tile of
//    at most 32 rows and 17 to 32 columns; 64 x 16 and 64 x 32: at most 16
//    and 17 to 32 columns; FP32 128 x 16 and 128 x 32 while more than 96 rows
//    remain in large launches).
//  * k_bf3w<TR,TC> — three-term split-BF16 (x = hi + mid + lo), six
//    products of order <= 2 with v_mfma_f32_16x16x32_bf16, FP32 per task,
//    FP64 across; four waves per TR x TC tile (64 x 64: C blocks more than
//    32 columns wide; 32 x 32: 17 to 32 columns; 64 x 16: at most 16
//    columns), A and B split once per chunk into bf16 planes in LDS.
//  RT = 4 (64-row tiles) in general, RT = 8 while more than 96 rows remain
//  in large FP64 launches, RT = 2 (32-row tiles) for a block's last row
//  tile when at most 32 rows remain (no MFMAs on padding rows).

#include <type_traits>

#include "plan.hpp"

namespace sht {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef double v2d8 __attribute__((ext_vector_type(2), aligned(8)));  // 16 bytes, 8-byte aligned
typedef float v2f4 __attribute__((ext_vector_type(2), aligned(4)));   // 8 bytes, 4-byte aligned

// Two consecutive operand elements, widened to double (FP64 tensors, or the
// FP32 tensors the FP32 and split-BF16 forms may read; the widening is
// exact, so the kernels' later rounding to FP32 or bf16 sees the stored value).
__device__ inline void ld2(const double* p, double& x, double& y)
{
  const v2d8 v = *reinterpret_cast<const v2d8*>(p);
  x = v.x;
  y = v.y;
}
__device__ inline void ld2(const float* p, double& x, double& y)
{
  const v2f4 v = *reinterpret_cast<const v2f4*>(p);
  x = v.x;
  y = v.y;
}

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
  if (tile.pad) return;  // a padding tile of the XCD-interleaved order
  const CBlkRec cb{tile.c_off, tile.m, tile.n, tile.t0, tile.t1};  // (carried by the tile record)
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
    // the first task's record comes with the tile; each next one is loaded
    // while the current task runs
    TaskRec trn{tile.a0, tile.b0, tile.k0, 0};
    for (int t = cb.t0; t < cb.t1; ++t) {
      const TaskRec tr = trn;
      trn = p.tasks[min(t + 1, cb.t1 - 1)];
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
// The kernels compute the transposed tile, D = (B fragment)^T x (A
// fragment)^T (products commute: every element is the same chain), so lane
// l holds row l%16 of four C columns and each C load and store instruction
// covers 16 consecutive rows of four columns.
// ---------------------------------------------------------------------------
// Orders this wave's LDS accesses (each wave owns its LDS buffer).
__device__ inline void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CT, int RT>
__global__ __launch_bounds__(64) void k_mfma64(ExecArgs p)
{
  const TileRec tile = p.tiles[blockIdx.x];
  if (tile.pad) return;  // a padding tile of the XCD-interleaved order
  const CBlkRec cb{tile.c_off, tile.m, tile.n, tile.t0, tile.t1};  // (carried by the tile record)
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
        const int i = tile.m0 + 16 * rt + li;
        const int j = tile.n0 + 16 * ct + 4 * v + lk;
        acc[rt][ct][v] = (!p.beta_zero && i < M && j < N) ? Cb[i + static_cast<int64_t>(j) * M] * p.beta : 0.0;
      }
  const double apad = pad_a(p.alpha, p.alpha_one != 0);
  if (!p.alpha_zero) {
    // the first task's record comes with the tile; each next one is loaded
    // while the current task runs
    TaskRec trn{tile.a0, tile.b0, tile.k0, 0};
    for (int t = cb.t0; t < cb.t1; ++t) {
      const TaskRec tr = trn;
      trn = p.tasks[min(t + 1, cb.t1 - 1)];
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
          a[rt] = (kin && i < M) ? Ab[i + static_cast<int64_t>(kk) * M] : apad;
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
            acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ct], a[rt], acc[rt][ct], 0, 0, 0);
      }
    }
  }
  // Write-back four columns at a time through LDS: lane (li, lk) holds rows
  // 16rt + li of column 4v + lk of the group; lane l then stores row
  // m0 + l % R of column l / R of each group of 64 / R columns (R <= 64), or
  // rows m0 + 64h + l of each column (R > 64), so every store instruction
  // writes whole column pieces of 64 x 8 bytes.
  constexpr int R = 16 * RT;
  constexpr int CPS = R >= 64 ? 1 : 64 / R;  // columns per store instruction
  constexpr int RPS = R >= 64 ? R / 64 : 1;  // store instructions per column
  __shared__ double cst[4 * R];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) cst[lk * R + 16 * rt + li] = acc[rt][ct][v];
      wave_lds_sync();
#pragma unroll
      for (int c = 0; c < 4; c += CPS)
#pragma unroll
        for (int h = 0; h < RPS; ++h) {
          const int cc = R >= 64 ? c : c + lane / R;
          const int r = R >= 64 ? 64 * h + lane : lane % R;
          const int i = tile.m0 + r;
          const int j = tile.n0 + 16 * ct + 4 * v + cc;
          if (i < M && j < N) Cb[i + static_cast<int64_t>(j) * M] = cst[cc * R + r];
        }
      wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------
// MFMA on 64 x 64 C tiles, one workgroup of four waves per tile (C blocks
// more than 32 columns wide); F = 0: FP64 throughout, F = 1: FP32 per task
// (products and sums of one block product), FP64 across tasks.  The
// workgroup walks the block's tasks in 16-k chunks: each chunk of A (64
// rows x 16) and B (16 x 64 columns) is loaded once, by all 256 lanes in
// 32-byte pieces, staged in LDS (two stages, one barrier per chunk) and read
// by the waves as MFMA fragments; wave w owns rows 32(w & 1) .. +31 and
// columns 32(w >> 1) .. +31 of the tile (2 x 2 transposed MFMA tiles, as
// k_mfma64 / k_mfma32).  The next chunk's global loads are issued before
// the current chunk's MFMAs.  Loads are unconditional: rows past M, columns
// past N and contracted indices past K read clamped in-bounds addresses
// (tensor buffers carry slack); k >= K is zeroed when the chunk is written
// to LDS, rows and columns past the block are never stored.  Chunks run task
// after task in task order and k ascending, so every C element sees the
// same chain as in the one-wave kernels of the same precision.
// ---------------------------------------------------------------------------
// The LDS stages of one TR x TC tile.  Row padding against bank conflicts of
// the loaders' stores and the waves' fragment reads, per tile shape
// (measured on C3 b=32, 32 x 32 tiles: FP64 0.315-0.33 -> 0.26-0.31 ms, FP32
// 0.221 -> 0.194 ms on one box and unchanged on another; C0's 64 x 16 FP64
// tiles 0.0426 -> 0.0407 ms; no gain measured for the other shapes, which
// keep none).
template <int F, int TR, int TC>
struct MfmawLds {
  typedef typename std::conditional<F == 0, double, float>::type TL;
  static constexpr bool sq32 = TR == 32 && TC == 32;
  static constexpr int PA = F == 0 ? (sq32 ? 1 : 0) : (sq32 ? 16 : 0);
  static constexpr int PB = F == 0 ? (sq32 ? 1 : (TC == 16 ? 1 : 0)) : (sq32 ? 18 : 0);
  TL sA[2][16][TR + PA];  // [stage][k][row of the tile]
  TL sB[2][16][TC + PB];  // [stage][k][column of the tile]
};

// One tile (called by the whole workgroup; the tile is not a padding tile).
template <int F, int TR, int TC, class TE>
__device__ __attribute__((always_inline)) inline void mfmaw_tile(const ExecArgs& p, const TileRec& tile,
                                                                 MfmawLds<F, TR, TC>& L)
{
  static_assert(F != 0 || std::is_same<TE, double>::value, "FP64 MFMA reads FP64 operands");
  // waves: WR along the rows x WC along the columns of the tile
  constexpr int WC = TC >= 32 ? 2 : 1, WR = 4 / WC;
  constexpr int TRt = TR / WR / 16, TCt = TC / WC / 16;  // 16 x 16 MFMA tiles per wave
  constexpr int EA = TR / 16, EB = TC / 16;  // elements of A / of B per loader lane and chunk
  static_assert(TRt >= 1 && TCt >= 1 && EA % 2 == 0 && (EB == 1 || EB % 2 == 0), "tile shape");
  typedef typename MfmawLds<F, TR, TC>::TL TL;
  const CBlkRec cb{tile.c_off, tile.m, tile.n, tile.t0, tile.t1};  // (carried by the tile record)
  const int M = cb.m, N = cb.n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = (TR / WR) * (w % WR), wc = (TC / WC) * (w / WR);
  const TE* A = static_cast<const TE*>(p.A);
  const TE* B = static_cast<const TE*>(p.B);
  double* Cb = static_cast<double*>(p.C) + cb.c_off;
  auto& sA = L.sA;
  auto& sB = L.sB;
  // column of accumulator element v (the f64 and f32 MFMA results differ)
  auto colof = [&](int ct, int v) { return tile.n0 + wc + 16 * ct + (F == 0 ? 4 * v + lk : 4 * lk + v); };
  v4d acc[TRt][TCt];
#pragma unroll
  for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
    for (int ct = 0; ct < TCt; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + wr + 16 * rt + li;
        const int j = colof(ct, v);
        acc[rt][ct][v] = (!p.beta_zero && i < M && j < N) ? Cb[i + static_cast<int64_t>(j) * M] * p.beta : 0.0;
      }
  if (!p.alpha_zero && cb.t1 > cb.t0) {
    // loader lanes: A column kA (0..15) rows rA..rA+EA-1; B column jB
    // (0..TC-1) contracted rows kB..kB+EB-1
    const int kA = tid >> 4, rA = EA * (tid & 15);
    const int jB = tid / (16 / EB), kB = EB * (tid % (16 / EB));
    const int ia = min(tile.m0 + rA, M - 1);
    const int jb = min(tile.n0 + jB, N - 1);
    int tl = cb.t0, kl = 0;  // load cursor (uniform): task, first contracted index
    TaskRec tr{tile.a0, tile.b0, tile.k0, 0};  // (task tl = t0, carried by the tile record)
    double ra[EA], rb[EB];
    bool ka;       // this lane's A column is inside the task
    int kbb, Kb;   // contracted index of rb[0] and the task's K
    bool lastl;    // the loaded chunk is the last of its task (uniform)
    const double alpha = p.alpha;
    const bool a1 = p.alpha_one != 0;
    const double apad = pad_a(alpha, a1);
    v4f part[TRt][TCt];
#pragma unroll
    for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
      for (int ct = 0; ct < TCt; ++ct) part[rt][ct] = v4f{0.f, 0.f, 0.f, 0.f};
#define SHT_W_LOAD()                                                                       \
  {                                                                                        \
    const int K = tr.k;                                                                    \
    const int kk = kl + kA;                                                                \
    ka = kk < K;                                                                           \
    const TE* pa = A + tr.a_off + static_cast<int64_t>(min(kk, K - 1)) * M + ia;           \
    _Pragma("unroll") for (int e = 0; e < EA; e += 2) ld2(pa + e, ra[e], ra[e + 1]);       \
    kbb = kl + kB;                                                                         \
    Kb = K;                                                                                \
    const TE* pb = B + tr.b_off + static_cast<int64_t>(jb) * K + min(kbb, K - 1);          \
    if constexpr (EB == 1) rb[0] = pb[0];                                                  \
    else {                                                                                 \
      _Pragma("unroll") for (int e = 0; e < EB; e += 2) ld2(pb + e, rb[e], rb[e + 1]);     \
    }                                                                                      \
    lastl = kl + 16 >= K;                                                                  \
    if (tl < cb.t1) {                                                                      \
      kl += 16;                                                                            \
      if (kl >= K) {                                                                       \
        kl = 0;                                                                            \
        ++tl;                                                                              \
        if (tl < cb.t1) tr = p.tasks[tl];                                                  \
      }                                                                                    \
    }                                                                                      \
  }
#define SHT_W_STORE(stg)                                                                   \
  {                                                                                        \
    _Pragma("unroll") for (int e = 0; e < EA; ++e) {                                       \
      const double x = ka ? ra[e] : apad;                                                  \
      if constexpr (F == 0) sA[stg][kA][rA + e] = a1 ? x : x * alpha;                      \
      else sA[stg][kA][rA + e] = static_cast<float>(x);                                    \
    }                                                                                      \
    _Pragma("unroll") for (int e = 0; e < EB; ++e) {                                       \
      const double y = kbb + e < Kb ? rb[e] : 0.0;                                         \
      if constexpr (F == 0) sB[stg][kB + e][jB] = y;                                       \
      else sB[stg][kB + e][jB] = static_cast<float>(y);                                    \
    }                                                                                      \
  }
    SHT_W_LOAD()
    SHT_W_STORE(0)
    bool lastc = lastl;  // the chunk in the current stage ends its task
    __syncthreads();
    int stg = 0;
    for (;;) {
      const bool more = tl < cb.t1;  // a next chunk exists (uniform)
      SHT_W_LOAD()                   // (past the end: a clamped re-read, unused)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        TL a[TRt], b[TCt];
#pragma unroll
        for (int rt = 0; rt < TRt; ++rt) a[rt] = sA[stg][4 * s + lk][wr + 16 * rt + li];
#pragma unroll
        for (int ct = 0; ct < TCt; ++ct) b[ct] = sB[stg][4 * s + lk][wc + 16 * ct + li];
#pragma unroll
        for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
          for (int ct = 0; ct < TCt; ++ct) {
            if constexpr (F == 0)
              acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ct], a[rt], acc[rt][ct], 0, 0, 0);
            else
              part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[ct], a[rt], part[rt][ct], 0, 0, 0);
          }
      }
      if constexpr (F == 1) {
        if (lastc) {  // the task's block product is complete: add it in FP64
#pragma unroll
          for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
            for (int ct = 0; ct < TCt; ++ct) {
#pragma unroll
              for (int v = 0; v < 4; ++v) acc[rt][ct][v] += alpha * static_cast<double>(part[rt][ct][v]);
              part[rt][ct] = v4f{0.f, 0.f, 0.f, 0.f};
            }
        }
      }
      if (!more) break;
      SHT_W_STORE(stg ^ 1)
      lastc = lastl;
      __syncthreads();
      stg ^= 1;
    }
#undef SHT_W_LOAD
#undef SHT_W_STORE
  }
#pragma unroll
  for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
    for (int ct = 0; ct < TCt; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + wr + 16 * rt + li;
        const int j = colof(ct, v);
        if (i < M && j < N) Cb[i + static_cast<int64_t>(j) * M] = acc[rt][ct][v];
      }
}

template <int F, int TR, int TC, class TE = double>
__global__ __launch_bounds__(256) void k_mfmaw(ExecArgs p)
{
  __shared__ MfmawLds<F, TR, TC> L;
  const TileRec tile = p.tiles[blockIdx.x];
  if (tile.pad) return;  // (uniform over the workgroup)
  mfmaw_tile<F, TR, TC, TE>(p, tile, L);
}

// ---------------------------------------------------------------------------
// FP32 MFMA per task, FP64 across tasks.  v_mfma_f32_16x16x4f32 has the same
// operand/result layout as the f64 variant above.
// ---------------------------------------------------------------------------
template <int CT, int RT, class TE = double>
__global__ __launch_bounds__(64) void k_mfma32(ExecArgs p)
{
  const TileRec tile = p.tiles[blockIdx.x];
  if (tile.pad) return;  // a padding tile of the XCD-interleaved order
  const CBlkRec cb{tile.c_off, tile.m, tile.n, tile.t0, tile.t1};  // (carried by the tile record)
  const int M = cb.m, N = cb.n;
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = lane >> 4;
  const TE* A = static_cast<const TE*>(p.A);
  const TE* B = static_cast<const TE*>(p.B);
  double* Cb = static_cast<double*>(p.C) + cb.c_off;
  v4d acc[RT][CT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + 16 * rt + li;
        const int j = tile.n0 + 16 * ct + 4 * lk + v;
        acc[rt][ct][v] = (!p.beta_zero && i < M && j < N) ? Cb[i + static_cast<int64_t>(j) * M] * p.beta : 0.0;
      }
  if (!p.alpha_zero) {
    // the first task's record comes with the tile; each next one is loaded
    // while the current task runs
    TaskRec trn{tile.a0, tile.b0, tile.k0, 0};
    for (int t = cb.t0; t < cb.t1; ++t) {
      const TaskRec tr = trn;
      trn = p.tasks[min(t + 1, cb.t1 - 1)];
      const TE* Ab = A + tr.a_off;
      const TE* Bb = B + tr.b_off;
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
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[ct], a[rt], part[rt][ct], 0, 0, 0);
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[rt][ct][v] += p.alpha * static_cast<double>(part[rt][ct][v]);
    }
  }
  // Write-back through LDS as in k_mfma64 (here lane (li, lk) holds column
  // 4lk + v of each 16-column group): every store instruction writes whole
  // column pieces of 64 x 8 bytes.
  constexpr int R = 16 * RT;
  constexpr int CPS = R >= 64 ? 1 : 64 / R;
  constexpr int RPS = R >= 64 ? R / 64 : 1;
  __shared__ double cst[4 * R];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) cst[lk * R + 16 * rt + li] = acc[rt][ct][v];
      wave_lds_sync();
#pragma unroll
      for (int c = 0; c < 4; c += CPS)
#pragma unroll
        for (int h = 0; h < RPS; ++h) {
          const int cc = R >= 64 ? c : c + lane / R;
          const int r = R >= 64 ? 64 * h + lane : lane % R;
          const int i = tile.m0 + r;
          const int j = tile.n0 + 16 * ct + 4 * cc + v;
          if (i < M && j < N) Cb[i + static_cast<int64_t>(j) * M] = cst[cc * R + r];
        }
      wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------
// Three-term split-BF16 (C-stationary).  v_mfma_f32_16x16x32_bf16 (wave64):
// A frag lane l holds A[l%16][8*(l/16) .. 8*(l/16)+7]; B frag holds
// B[8*(l/16)..+7][l%16]; D as the f32 form above.  x = hi + mid + lo (bf16
// each, split3x2_cs below); the six products of order <= 2 accumulate in
// FP32 per task (smallest first: lo*hi, mid*mid, hi*lo, mid*hi, hi*mid,
// hi*hi), FP64 across tasks.
// ---------------------------------------------------------------------------
// x = hi + mid + lo for two values at once with gfx950's packed conversion
// (v_cvt_pk_bf16_f32, round to nearest even): hi = bf16(x), mid =
// bf16(x - hi), lo = bf16(x - hi - mid), each difference exact in FP32; each
// output word holds the parts of x0 (low half) and x1 (high half).
typedef float v2f_cs __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf_cs __attribute__((ext_vector_type(2)));
__device__ inline void split3x2_cs(float x0, float x1, uint32_t& h, uint32_t& m, uint32_t& l)
{
  const v2f_cs x = {x0, x1};
  const v2bf_cs hb = __builtin_convertvector(x, v2bf_cs);
  const v2f_cs r1 = x - __builtin_convertvector(hb, v2f_cs);
  const v2bf_cs mb = __builtin_convertvector(r1, v2bf_cs);
  const v2f_cs r2 = r1 - __builtin_convertvector(mb, v2f_cs);
  const v2bf_cs lb = __builtin_convertvector(r2, v2bf_cs);
  h = __builtin_bit_cast(uint32_t, hb);
  m = __builtin_bit_cast(uint32_t, mb);
  l = __builtin_bit_cast(uint32_t, lb);
}

// Split-BF16 on TR x TC C tiles, one workgroup of four waves per tile (64 x
// 64 for C blocks more than 32 columns wide, 32 x 32 for 17 to 32 columns,
// 64 x 16 for at most 16).  Each 32-k chunk of A (TR rows x 32) and B (32 x
// TC columns) is loaded once by the workgroup (lane: one row or column,
// TR/8 or TC/8 consecutive k), split into three bf16 planes as it is
// written to LDS (two stages, one barrier per chunk; 32-element rows with
// XOR-rotated 16-byte pieces, w3k) and read by the waves as 16-byte MFMA
// fragments; the waves split
// the tile into 2 x 2 (or 4 x 1 for 16 columns) parts of 16 x 16 MFMA
// tiles.  Loads are unconditional (clamped in-bounds addresses); k >= K is
// zeroed before the split.
constexpr int kW3Pitch = 32;
// LDS position of contracted index k (0..31) in a plane row: the row's four
// 16-byte pieces are XOR-rotated by (row / 2) % 4, so the loaders' stores
// (consecutive rows, one piece) and the waves' 16-byte fragment reads (16
// consecutive rows, four pieces) meet no bank conflicts without row padding
// (unpadded rows keep the 64 x 64 tile at 48 KB of LDS: three workgroups
// per CU instead of two).
__device__ inline int w3k(int row, int k) { return (((k >> 3) ^ ((row >> 1) & 3)) << 3) | (k & 7); }
template <int TR, int TC>
struct Bf3Lds {
  short sA[2][3][TR][kW3Pitch];  // [stage][plane][row of the tile][k]
  short sB[2][3][TC][kW3Pitch];  // [stage][plane][column of the tile][k]
};

// One tile (called by the whole workgroup; the tile is not a padding tile).
template <int TR, int TC, class TE>
__device__ __attribute__((always_inline)) inline void bf3w_tile(const ExecArgs& p, const TileRec& tile,
                                                                Bf3Lds<TR, TC>& L)
{
  // waves: WR along the rows x WC along the columns of the tile
  constexpr int WC = TC >= 32 ? 2 : 1, WR = 4 / WC;
  constexpr int TRt = TR / WR / 16, TCt = TC / WC / 16;  // 16 x 16 MFMA tiles per wave
  constexpr int KA = TR / 8, KB = TC / 8;  // contracted indices per loader lane and chunk
  static_assert(TRt >= 1 && TCt >= 1 && KB >= 2, "tile shape");
  const CBlkRec cb{tile.c_off, tile.m, tile.n, tile.t0, tile.t1};  // (carried by the tile record)
  const int M = cb.m, N = cb.n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = (TR / WR) * (w % WR), wc = (TC / WC) * (w / WR);
  const TE* A = static_cast<const TE*>(p.A);
  const TE* B = static_cast<const TE*>(p.B);
  double* Cb = static_cast<double*>(p.C) + cb.c_off;
  auto& sA = L.sA;
  auto& sB = L.sB;
  v4d acc[TRt][TCt];
#pragma unroll
  for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
    for (int ct = 0; ct < TCt; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + wr + 16 * rt + li;
        const int j = tile.n0 + wc + 16 * ct + 4 * lk + v;
        acc[rt][ct][v] = (!p.beta_zero && i < M && j < N) ? Cb[i + static_cast<int64_t>(j) * M] * p.beta : 0.0;
      }
  if (!p.alpha_zero && cb.t1 > cb.t0) {
    // loader lanes: row ra_ of A with contracted indices KA*ga .. +KA-1 of
    // the chunk; column cb_ of B with KB*gb .. +KB-1
    const int ra_ = tid % TR, ga = tid / TR;
    const int cb_ = tid % TC, gb = tid / TC;
    const int ia = min(tile.m0 + ra_, M - 1);
    const int jb = min(tile.n0 + cb_, N - 1);
    int tl = cb.t0, kl = 0;  // load cursor (uniform): task, first contracted index
    TaskRec tr{tile.a0, tile.b0, tile.k0, 0};  // (task tl = t0, carried by the tile record)
    double ra[KA], rb[KB];
    int k0a, k0b, Kl;  // contracted index of ra[0] and of rb[0], the task's K
    bool lastl;   // the loaded chunk is the last of its task (uniform)
#define SHT_W3_LOAD()                                                                      \
  {                                                                                        \
    const int K = tr.k;                                                                    \
    k0a = kl + KA * ga;                                                                    \
    k0b = kl + KB * gb;                                                                    \
    Kl = K;                                                                                \
    const TE* pa = A + tr.a_off + ia;                                                      \
    _Pragma("unroll") for (int e = 0; e < KA; ++e)                                         \
        ra[e] = pa[static_cast<int64_t>(min(k0a + e, K - 1)) * M];                         \
    const TE* pb = B + tr.b_off + static_cast<int64_t>(jb) * K + min(k0b, K - 1);          \
    _Pragma("unroll") for (int e = 0; e < KB; e += 2) ld2(pb + e, rb[e], rb[e + 1]);       \
    lastl = kl + 32 >= K;                                                                  \
    if (tl < cb.t1) {                                                                      \
      kl += 32;                                                                            \
      if (kl >= K) {                                                                       \
        kl = 0;                                                                            \
        ++tl;                                                                              \
        if (tl < cb.t1) tr = p.tasks[tl];                                                  \
      }                                                                                    \
    }                                                                                      \
  }
#define SHT_W3_STORE(stg)                                                                  \
  {                                                                                        \
    typedef uint32_t vas __attribute__((ext_vector_type(KA / 2)));                         \
    typedef uint32_t vbs __attribute__((ext_vector_type(KB / 2)));                         \
    vas h, m, l;                                                                           \
    _Pragma("unroll") for (int e = 0; e < KA; e += 2) {                                    \
      uint32_t x, y, z;                                                                    \
      split3x2_cs(k0a + e < Kl ? static_cast<float>(ra[e]) : 0.f,                          \
                  k0a + e + 1 < Kl ? static_cast<float>(ra[e + 1]) : 0.f, x, y, z);         \
      h[e / 2] = x;                                                                        \
      m[e / 2] = y;                                                                        \
      l[e / 2] = z;                                                                        \
    }                                                                                      \
    *reinterpret_cast<vas*>(&sA[stg][0][ra_][w3k(ra_, KA * ga)]) = h;                                \
    *reinterpret_cast<vas*>(&sA[stg][1][ra_][w3k(ra_, KA * ga)]) = m;                                \
    *reinterpret_cast<vas*>(&sA[stg][2][ra_][w3k(ra_, KA * ga)]) = l;                                \
    vbs hb, mb, lb;                                                                        \
    _Pragma("unroll") for (int e = 0; e < KB; e += 2) {                                    \
      uint32_t x, y, z;                                                                    \
      split3x2_cs(k0b + e < Kl ? static_cast<float>(rb[e]) : 0.f,                          \
                  k0b + e + 1 < Kl ? static_cast<float>(rb[e + 1]) : 0.f, x, y, z);         \
      hb[e / 2] = x;                                                                       \
      mb[e / 2] = y;                                                                       \
      lb[e / 2] = z;                                                                       \
    }                                                                                      \
    *reinterpret_cast<vbs*>(&sB[stg][0][cb_][w3k(cb_, KB * gb)]) = hb;                               \
    *reinterpret_cast<vbs*>(&sB[stg][1][cb_][w3k(cb_, KB * gb)]) = mb;                               \
    *reinterpret_cast<vbs*>(&sB[stg][2][cb_][w3k(cb_, KB * gb)]) = lb;                               \
  }
    v4f part[TRt][TCt];
#pragma unroll
    for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
      for (int ct = 0; ct < TCt; ++ct) part[rt][ct] = v4f{0.f, 0.f, 0.f, 0.f};
    SHT_W3_LOAD()
    SHT_W3_STORE(0)
    bool lastc = lastl;  // the chunk in the current stage ends its task
    __syncthreads();
    int stg = 0;
    for (;;) {
      const bool more = tl < cb.t1;  // a next chunk exists (uniform)
      SHT_W3_LOAD()                  // (past the end: a clamped re-read, unused)
      {
        v8s ah[TRt], am[TRt], al[TRt], bh[TCt], bm[TCt], bl[TCt];
#pragma unroll
        for (int rt = 0; rt < TRt; ++rt) {
          ah[rt] = *reinterpret_cast<const v8s*>(&sA[stg][0][wr + 16 * rt + li][w3k(li, 8 * lk)]);
          am[rt] = *reinterpret_cast<const v8s*>(&sA[stg][1][wr + 16 * rt + li][w3k(li, 8 * lk)]);
          al[rt] = *reinterpret_cast<const v8s*>(&sA[stg][2][wr + 16 * rt + li][w3k(li, 8 * lk)]);
        }
#pragma unroll
        for (int ct = 0; ct < TCt; ++ct) {
          bh[ct] = *reinterpret_cast<const v8s*>(&sB[stg][0][wc + 16 * ct + li][w3k(li, 8 * lk)]);
          bm[ct] = *reinterpret_cast<const v8s*>(&sB[stg][1][wc + 16 * ct + li][w3k(li, 8 * lk)]);
          bl[ct] = *reinterpret_cast<const v8s*>(&sB[stg][2][wc + 16 * ct + li][w3k(li, 8 * lk)]);
        }
#pragma unroll
        for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
          for (int ct = 0; ct < TCt; ++ct) {
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[ct], al[rt], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm[ct], am[rt], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[ct], ah[rt], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[ct], am[rt], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm[ct], ah[rt], part[rt][ct], 0, 0, 0);
            part[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[ct], ah[rt], part[rt][ct], 0, 0, 0);
          }
      }
      if (lastc) {  // the task's block product is complete: add it in FP64
#pragma unroll
        for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
          for (int ct = 0; ct < TCt; ++ct) {
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[rt][ct][v] += p.alpha * static_cast<double>(part[rt][ct][v]);
            part[rt][ct] = v4f{0.f, 0.f, 0.f, 0.f};
          }
      }
      if (!more) break;
      SHT_W3_STORE(stg ^ 1)
      lastc = lastl;
      __syncthreads();
      stg ^= 1;
    }
#undef SHT_W3_LOAD
#undef SHT_W3_STORE
  }
#pragma unroll
  for (int rt = 0; rt < TRt; ++rt)
#pragma unroll
    for (int ct = 0; ct < TCt; ++ct)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tile.m0 + wr + 16 * rt + li;
        const int j = tile.n0 + wc + 16 * ct + 4 * lk + v;
        if (i < M && j < N) Cb[i + static_cast<int64_t>(j) * M] = acc[rt][ct][v];
      }
}

template <int TR, int TC, class TE = double>
__global__ __launch_bounds__(256) void k_bf3w(ExecArgs p)
{
  __shared__ Bf3Lds<TR, TC> L;
  const TileRec tile = p.tiles[blockIdx.x];
  if (tile.pad) return;  // (uniform over the workgroup)
  bf3w_tile<TR, TC, TE>(p, tile, L);
}

// ---------------------------------------------------------------------------
// The persistent C-stationary launch: one grid of the workgroups the device
// keeps resident (a multiple of eight), serving every tile of the plan (its
// FP64, FP32 or split-BF16 four-wave classes mixed, each tile's class in its
// record).  Workgroup b serves XCD queue b % 8 (queue x holds tiles x, x + 8,
// x + 16, ...): with W workgroups per queue, it takes the queue's positions
// b / 8, b / 8 + W, b / 8 + 2W, ... (a fixed deal: claiming positions from an
// atomic counter per queue made C0 4.6x and C2 1.3x slower than the
// per-class launches, every claim a round trip to one contended address).
// Each tile runs the code of its class's own kernel (mfmaw_tile / bf3w_tile)
// in LDS shared by the classes; a barrier separates consecutive tiles.
// ---------------------------------------------------------------------------
__device__ inline bool next_tile(const ExecArgs& p, int64_t& t, TileRec& tile)
{
  const int64_t x = blockIdx.x & 7u, w = gridDim.x >> 3;
  for (; t < p.qlen; t += w) {
    const int64_t ti = 8 * t + x;
    if (ti >= p.ntiles) continue;  // (the last round of the queues is short)
    tile = p.tiles[ti];
    if (!tile.pad) {
      t += w;
      __syncthreads();  // every wave is done with the previous tile's LDS
      return true;
    }
  }
  return false;
}

template <int F, class TE>
__global__ __launch_bounds__(256) void k_csp(ExecArgs p)
{
  __shared__ union {
    MfmawLds<F, 64, 64> w;
    MfmawLds<F, 32, 32> s;
    MfmawLds<F, 64, 16> n;
    MfmawLds<F, 64, 32> m;
  } L;
  TileRec tile;
  int64_t t = blockIdx.x >> 3;
  while (next_tile(p, t, tile)) {
    switch (static_cast<KClass>(tile.kc)) {
      case (F == 0 ? KClass::Mfma64W : KClass::F32W): mfmaw_tile<F, 64, 64, TE>(p, tile, L.w); break;
      case (F == 0 ? KClass::Mfma64S : KClass::F32S): mfmaw_tile<F, 32, 32, TE>(p, tile, L.s); break;
      case (F == 0 ? KClass::Mfma64N : KClass::F32N): mfmaw_tile<F, 64, 16, TE>(p, tile, L.n); break;
      default: mfmaw_tile<F, 64, 32, TE>(p, tile, L.m); break;  // Mfma64M / F32M
    }
  }
}

template <class TE>
__global__ __launch_bounds__(256) void k_bf3q(ExecArgs p)
{
  __shared__ union {
    Bf3Lds<64, 64> w;
    Bf3Lds<32, 32> s;
    Bf3Lds<64, 16> n;
  } L;
  TileRec tile;
  int64_t t = blockIdx.x >> 3;
  while (next_tile(p, t, tile)) {
    switch (static_cast<KClass>(tile.kc)) {
      case KClass::Bf16W: bf3w_tile<64, 64, TE>(p, tile, L.w); break;
      case KClass::Bf16S: bf3w_tile<32, 32, TE>(p, tile, L.s); break;
      default: bf3w_tile<64, 16, TE>(p, tile, L.n); break;  // Bf16N
    }
  }
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
    case KClass::Mfma16h:
    case KClass::F32x16h:
    case KClass::Bf16S:
    case KClass::Mfma64S:
    case KClass::F32S: return 32;
    case KClass::Mfma16q:
    case KClass::Mfma32q:
    case KClass::F32Nq:
    case KClass::F32Mq: return 128;
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
    case KClass::Mfma16h:
    case KClass::F32x16h:
    case KClass::Mfma16q:
    case KClass::Bf16N:
    case KClass::Mfma64N:
    case KClass::F32N:
    case KClass::F32Nq: return 16;
    case KClass::Mfma64W:
    case KClass::F32W:
    case KClass::Bf16W: return 64;
    case KClass::Bf16S:
    case KClass::Mfma64S:
    case KClass::F32S: return 32;
    default: return 32;
  }
}

int kclass_waves(KClass k)
{
  switch (k) {
    case KClass::Exact8:
    case KClass::Exact16:
    case KClass::Exact32:
    case KClass::Mfma16:
    case KClass::Mfma32:
    case KClass::Mfma16h:
    case KClass::F32x16h:
    case KClass::Mfma16q:
    case KClass::Mfma32q: return 1;
    default: return 4;
  }
}

void launch_exec(KClass k, const ExecArgs& a, hipStream_t s)
{
  if (a.ntiles <= 0) return;
  // one tile per workgroup of kclass_waves(k) waves
  const dim3 g(static_cast<unsigned>(a.ntiles)), b(kclass_waves(k) * kWave);
  // the FP32 and split-BF16 classes read FP64 or FP32 operands
#define SHT_EXEC_F32(kern, ...)                                                            \
  {                                                                                        \
    if (a.f32) hipLaunchKernelGGL((kern<__VA_ARGS__, float>), g, b, 0, s, a);              \
    else hipLaunchKernelGGL((kern<__VA_ARGS__, double>), g, b, 0, s, a);                   \
  }
  SHT_REQUIRE(!a.f32 || k == KClass::F32x16h || k == KClass::F32W || k == KClass::F32N || k == KClass::F32M ||
                  k == KClass::F32Nq || k == KClass::F32Mq || k == KClass::F32S || k == KClass::Bf16W ||
                  k == KClass::Bf16S || k == KClass::Bf16N || k == KClass::Persist32 || k == KClass::PersistBf16,
              SHT_ERR_STATE, "FP32 operands reach an FP64 kernel class");
  switch (k) {
    case KClass::Exact8: hipLaunchKernelGGL(k_exact<8>, g, b, 0, s, a); break;
    case KClass::Exact16: hipLaunchKernelGGL(k_exact<16>, g, b, 0, s, a); break;
    case KClass::Exact32: hipLaunchKernelGGL(k_exact<32>, g, b, 0, s, a); break;
    case KClass::Mfma16: hipLaunchKernelGGL((k_mfma64<1, 4>), g, b, 0, s, a); break;
    case KClass::Mfma32: hipLaunchKernelGGL((k_mfma64<2, 4>), g, b, 0, s, a); break;
    case KClass::Mfma16h: hipLaunchKernelGGL((k_mfma64<1, 2>), g, b, 0, s, a); break;
    case KClass::F32x16h: SHT_EXEC_F32(k_mfma32, 1, 2); break;
    case KClass::Mfma16q: hipLaunchKernelGGL((k_mfma64<1, 8>), g, b, 0, s, a); break;
    case KClass::Mfma32q: hipLaunchKernelGGL((k_mfma64<2, 8>), g, b, 0, s, a); break;
    case KClass::Mfma64W: hipLaunchKernelGGL((k_mfmaw<0, 64, 64>), g, b, 0, s, a); break;
    case KClass::F32W: SHT_EXEC_F32(k_mfmaw, 1, 64, 64); break;
    case KClass::Mfma64S: hipLaunchKernelGGL((k_mfmaw<0, 32, 32>), g, b, 0, s, a); break;
    case KClass::Mfma64N: hipLaunchKernelGGL((k_mfmaw<0, 64, 16>), g, b, 0, s, a); break;
    case KClass::F32N: SHT_EXEC_F32(k_mfmaw, 1, 64, 16); break;
    case KClass::Mfma64M: hipLaunchKernelGGL((k_mfmaw<0, 64, 32>), g, b, 0, s, a); break;
    case KClass::F32M: SHT_EXEC_F32(k_mfmaw, 1, 64, 32); break;
    case KClass::F32Nq: SHT_EXEC_F32(k_mfmaw, 1, 128, 16); break;
    case KClass::F32Mq: SHT_EXEC_F32(k_mfmaw, 1, 128, 32); break;
    case KClass::F32S: SHT_EXEC_F32(k_mfmaw, 1, 32, 32); break;
    case KClass::Bf16W: SHT_EXEC_F32(k_bf3w, 64, 64); break;
    case KClass::Bf16S: SHT_EXEC_F32(k_bf3w, 32, 32); break;
    case KClass::Bf16N: SHT_EXEC_F32(k_bf3w, 64, 16); break;
    case KClass::Persist64:
    case KClass::Persist32:
    case KClass::PersistBf16: {
      ExecArgs q = a;
      q.qlen = (a.ntiles + 7) / 8;
      const void* f = k == KClass::Persist64
                          ? reinterpret_cast<const void*>(k_csp<0, double>)
                          : (k == KClass::Persist32
                                 ? (a.f32 ? reinterpret_cast<const void*>(k_csp<1, float>)
                                          : reinterpret_cast<const void*>(k_csp<1, double>))
                                 : (a.f32 ? reinterpret_cast<const void*>(k_bf3q<float>)
                                          : reinterpret_cast<const void*>(k_bf3q<double>)));
      const dim3 gp(resident_grid(f, 256, a.ntiles));
      if (k == KClass::Persist64) hipLaunchKernelGGL((k_csp<0, double>), gp, b, 0, s, q);
      else if (k == KClass::Persist32 && a.f32) hipLaunchKernelGGL((k_csp<1, float>), gp, b, 0, s, q);
      else if (k == KClass::Persist32) hipLaunchKernelGGL((k_csp<1, double>), gp, b, 0, s, q);
      else if (a.f32) hipLaunchKernelGGL(k_bf3q<float>, gp, b, 0, s, q);
      else hipLaunchKernelGGL(k_bf3q<double>, gp, b, 0, s, q);
      break;
    }
    default: throw Error(SHT_ERR_INVALID, "unknown kernel class");
  }
#undef SHT_EXEC_F32
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
