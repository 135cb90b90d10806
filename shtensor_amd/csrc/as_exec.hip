// A-stationary row-panel kernels of a contraction plan.
//
// A unit (one wave) owns a chunk of rows of block row I and walks the sweep
// list of I: for every A block (I, K), in ascending contracted key K, one or
// more sweeps over the packed panel of K (the partners (K, J) side by side),
// eight 16-column slots at most per sweep.  A sweep loads its slots' C tiles
// into accumulators, runs the A block's k loop with one panel fragment per
// slot and step shared by every 16-row tile of the chunk, and writes the
// tiles back.  A C element belongs to exactly one unit, and its
// contributions arrive in the task order (C id, K): no atomics, one fixed
// order.
//
// Kernels: k_asp (one wave per unit; schedule 1, 64-row units for schedule
// 3, and the split-BF16 groups of schedule 2 sharing panels through L1),
// k_asl (schedule 2 for FP64/FP32: the four units of a block row in one
// workgroup, each panel fragment loaded once and shared through LDS),
// k_as_exact (the fma chain, one row per lane).
//
// MFMA operand layouts (lane l of 64; measured on gfx950):
//   v_mfma_f64_16x16x4f64   A[l%16][l/16], B[l/16][l%16], D[4v + l/16][l%16]
//   v_mfma_f32_16x16x4f32   A[l%16][l/16], B[l/16][l%16], D[4(l/16) + v][l%16]
//   v_mfma_f32_16x16x32bf16 A[l%16][8(l/16)+e], B[8(l/16)+e][l%16], D as f32
// The kernels compute the transposed tile, D = (panel fragment) x (A
// fragment)^T: products commute, so every element is the same chain, and
// lane l then holds rows r0 + 16t + l%16 of four panel columns, so each C
// load and store instruction covers four columns x 16 consecutive rows.
//
//  * F = 0, FP64 MFMA: each C element accumulates as one fused multiply-add
//    chain in k order, c = fma(alpha*a, b, c), continued across sweeps —
//    bitwise equal to the reference JIT's chain (src/Kernel/KernelImpl.cpp:
//    594-710) continued over the task list.
//  * F = 1, FP32 MFMA per block product, added into C in FP64.
//  * F = 2, three-term split-BF16 (x = hi + mid + lo, about 24 significant
//    bits), six bf16 products of order <= 2 per 32-k step, FP32 per block
//    product, FP64 across blocks.
//  * k_as_exact: the fma chain with one row per lane (bitwise as well).

#include <map>
#include <mutex>
#include <type_traits>

#include "plan.hpp"

namespace sht {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));

namespace {

// wave-uniform copies (sweep and unit records are the same for every lane)
__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline int64_t uni(int64_t x)
{
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(x) >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Orders this wave's LDS writes before its later LDS reads of other lanes'
// entries (each wave owns its own LDS slice; no workgroup barrier).
__device__ inline void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct SweepU {
  int64_t a_off, b_off, c0;
  int k, m, pitch, ns;
};

__device__ inline SweepU to_uni(const AsSweep& v)
{
  SweepU u;
  u.a_off = uni(v.a_off);
  u.b_off = uni(v.b_off);
  u.c0 = uni(v.c0);
  u.k = uni(v.k);
  u.m = uni(v.m);
  u.pitch = uni(v.pitch);
  u.ns = uni(v.ns);
  return u;
}

// C offsets of this lane's two sweep columns (lane, lane + 64); -1 past the
// sweep's 16*ns columns.
__device__ inline void load_cols(const AsArgs& p, const SweepU& sw, int lane, int64_t& c0, int64_t& c1)
{
  const int n = 16 * sw.ns;
  const int64_t x0 = p.ccol[sw.c0 + min(lane, n - 1)];
  const int64_t x1 = p.ccol[sw.c0 + min(lane + 64, n - 1)];
  c0 = lane < n ? x0 : -1;
  c1 = lane + 64 < n ? x1 : -1;
}

constexpr int cmin(int a, int b) { return a < b ? a : b; }

// A sweep column entry: -1 (pad column or dropped block), an element offset
// v of C(0, column), or -(v + 2) for the first sweep of the block row that
// touches the column.
__device__ inline int64_t ccol_off(int64_t e) { return e < -1 ? -e - 2 : e; }

// Where the values a sweep column starts from are read, for rows from
// `row`: the C column; for a column's first sweep with p.zero set (beta ==
// 0), a line of zeros (C is not read); for a pad column, column 0 of C (its
// values are never stored).  Branch-free (address selects only).
__device__ inline const double* c_src(const AsArgs& p, int64_t e, int row)
{
  const double* C = static_cast<const double*>(p.C);
  const double* zs = p.zero ? p.zero + (row & 63) : C + (-e - 2) + row;
  return e >= 0 ? C + e + row : (e == -1 ? C + row : zs);
}

// ---------------------------------------------------------------------------
// The A operand of a unit is one stream across all its sweeps: a ring of
// four register sets holds the A fragments of the next steps, filled three
// steps ahead, and the last iteration of a sweep already loads the first
// three steps of the next sweep, so the A stream from HBM does not restart at
// sweep boundaries.  Every load is in bounds and unconditional: step indexes
// are clamped, A rows past M read on into the next column (tensor buffers
// carry kReadSlack bytes of slack), A columns past K re-read column K-1
// (masked to zero by the consumer), and panel rows past K are zero.  A is read
// once per unit: non-temporal loads keep the stream from evicting the panels
// and C tiles that the other units of the block row reuse in L2.
// ---------------------------------------------------------------------------
template <class TE>
struct AStream {
  const TE* a;  // A block (wave-uniform)
  int K, M;
};

// A fragments of step st of the stream (rows row + 16t, column 4st + lk).  A
// blocks of the schedule hold fewer than 2^28 elements, so the element
// offset fits 32 bits: the load is a uniform base plus a lane offset.
template <int R, class TE>
__device__ inline void load_a(const AStream<TE>& s, int st, int row, int lk, TE (&a)[R])
{
  const uint32_t off = static_cast<uint32_t>(min(4 * st + lk, s.K - 1) * s.M + row);
#pragma unroll
  for (int t = 0; t < R; ++t) a[t] = __builtin_nontemporal_load(s.a + (off + 16u * t));
}

// Step st of the current sweep's stream when st < nsteps, else step
// st - nsteps of the next sweep's (the selects are wave-uniform).
template <int R, class TE>
__device__ inline void load_a2(const AStream<TE>& cu, const AStream<TE>& nx, int nsteps, int st, int row, int lk,
                               TE (&a)[R])
{
  const bool n = st >= nsteps;
  AStream<TE> s;
  s.a = n ? nx.a : cu.a;
  s.K = n ? nx.K : cu.K;
  s.M = n ? nx.M : cu.M;
  load_a(s, n ? st - nsteps : st, row, lk, a);
}

// Two consecutive rows of one A column (16 bytes, 8-byte aligned), streamed.
typedef double v2d8 __attribute__((ext_vector_type(2), aligned(8)));
__device__ inline __attribute__((unused)) void load_a_pair(const double* a, uint32_t off, double (&x)[2])
{
  const v2d8 v = __builtin_nontemporal_load(reinterpret_cast<const v2d8*>(a + off));
  x[0] = v.x;
  x[1] = v.y;
}
// FP32 operands (FP32 and split-BF16 forms): the same two rows, 8 bytes.
typedef float v2f4 __attribute__((ext_vector_type(2), aligned(4)));
__device__ inline __attribute__((unused)) void load_a_pair(const float* a, uint32_t off, float (&x)[2])
{
  const v2f4 v = __builtin_nontemporal_load(reinterpret_cast<const v2f4*>(a + off));
  x[0] = v.x;
  x[1] = v.y;
}
// Two consecutive elements of an operand (plain load), widened to double.
__device__ inline __attribute__((unused)) void load_pair(const double* a, double& x, double& y)
{
  const v2d8 v = *reinterpret_cast<const v2d8*>(a);
  x = v.x;
  y = v.y;
}
__device__ inline __attribute__((unused)) void load_pair(const float* a, float& x, float& y)
{
  const v2f4 v = *reinterpret_cast<const v2f4*>(a);
  x = v.x;
  y = v.y;
}

template <int F> struct BElem { typedef double type; };
template <> struct BElem<1> { typedef float type; };

// ---------------------------------------------------------------------------
// FP64 / FP32 MFMA sweep: R row tiles x NS slots, four k-steps (16 k) per
// loop iteration (panel rows are zero-padded to a multiple of 16).  A is
// loaded three steps ahead (ring of four register sets), the panel one step
// ahead (two sets); in each phase the panel loads (mostly L2 hits) are
// issued before the A loads (from HBM): the vector memory counter retires in
// issue order, so waiting for the panel does not wait for the younger A
// loads.  Loads fill raw registers; the k < K mask and alpha are applied by
// the consumer after the scheduling barrier.  `cbs` (LDS) holds the C offset
// of each sweep column.
// ---------------------------------------------------------------------------
template <int R, int NS, int F, bool A1, class TE>
__device__ inline void sweep_mfma(const AsArgs& p, const SweepU& sw, const SweepU& nxs, int r0, int li, int lk,
                                  TE (&ar)[4][R], const int64_t* cbs)
{
  static_assert(F != 0 || std::is_same<TE, double>::value, "FP64 MFMA reads FP64 operands");
  typedef typename BElem<F>::type TB;
  typedef typename std::conditional<F == 0, v4d, v4f>::type Acc;
  const TB* Bl = static_cast<const TB*>(p.B) + sw.b_off + li + static_cast<int64_t>(lk) * sw.pitch;
  const int64_t bstep = 4 * static_cast<int64_t>(sw.pitch);
  const int K = sw.k, M = sw.m;
  double* C = static_cast<double*>(p.C);
  // column of the sweep held by this lane for accumulator element v
  auto colof = [&](int q, int v) { return F == 0 ? 16 * q + 4 * v + lk : 16 * q + 4 * lk + v; };
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
        const double* cq = c_src(p, cbs[colof(q, v)], rl);
#pragma unroll
        for (int t = 0; t < R; ++t) acc[t][q][v] = cq[16 * t];
      }
  } else {
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int q = 0; q < NS; ++q) acc[t][q] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  TB b0[NS], b1[NS];
  auto loadB = [&](int st, TB(&b)[NS]) {
    const TB* bs = Bl + st * bstep;
#pragma unroll
    for (int q = 0; q < NS; ++q) b[q] = bs[16 * q];
  };
  const int nsteps = ((K + 15) >> 4) * 4, last = nsteps - 1;
  loadB(0, b0);
  __builtin_amdgcn_sched_barrier(0);
  AStream<TE> cu, nx;
  cu.a = static_cast<const TE*>(p.A) + sw.a_off;
  cu.K = K;
  cu.M = M;
  nx.a = static_cast<const TE*>(p.A) + nxs.a_off;
  nx.K = nxs.k;
  nx.M = nxs.m;
  // steps past the contracted length (the loop runs whole 16-k iterations)
  // issue no MFMAs: every product they hold is zero
  const int kq = (K + 3) >> 2;
  const double apad = pad_a(p.alpha, A1);
  auto mma = [&](int st, const TE(&araw)[R], const TB(&b)[NS]) {
    if (st >= kq) return;
    const bool kin = 4 * st + lk < K;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if constexpr (F == 0) {
        const double y = kin ? araw[t] : apad;
        const double a = A1 ? y : y * p.alpha;
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[q], a, acc[t][q], 0, 0, 0);
      } else {
        const float a = kin ? static_cast<float>(araw[t]) : 0.f;
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[q], a, acc[t][q], 0, 0, 0);
      }
    }
  };
  for (int st = 0; st < nsteps; st += 4) {
    loadB(st + 1, b1);
    load_a2(cu, nx, nsteps, st + 3, rl, lk, ar[3]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st, ar[0], b0);
    __builtin_amdgcn_sched_barrier(0);
    loadB(st + 2, b0);
    load_a2(cu, nx, nsteps, st + 4, rl, lk, ar[0]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 1, ar[1], b1);
    __builtin_amdgcn_sched_barrier(0);
    loadB(st + 3, b1);
    load_a2(cu, nx, nsteps, st + 5, rl, lk, ar[1]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 2, ar[2], b0);
    __builtin_amdgcn_sched_barrier(0);
    loadB(min(st + 4, last), b0);
    load_a2(cu, nx, nsteps, st + 6, rl, lk, ar[2]);
    __builtin_amdgcn_sched_barrier(0);
    mma(st + 3, ar[3], b1);
    __builtin_amdgcn_sched_barrier(0);
  }
  // write back
  if constexpr (F == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t cb = ccol_off(cbs[colof(q, v)]);
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
          const int64_t e = q0 + x < NS ? cbs[colof(q0 + x, v)] : -1;
          cb[x][v] = ccol_off(e);
          const double* cq = c_src(p, e, rl);
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
}

// ---------------------------------------------------------------------------
// The same sweep for a workgroup of W waves that run W row chunks of one
// block row through the same sweep list: each k-step's panel fragments are
// fetched once per workgroup (wave w loads the slots q with q % W == w) and
// shared through LDS, so every wave reads all NS fragments from LDS instead
// of from L1/L2.  Per step s: issue this wave's panel loads for step s+4 and
// A loads for step s+AD-1 (running offsets), read step s's fragments from
// stage s % 4, run the MFMAs; after every odd step write this wave's share
// of steps s+1 and s+2 and pass a workgroup barrier.  Stage s % 4 is written
// before the barrier that precedes step s, and rewritten (with step s+4)
// only after the barrier that follows step s.  A ghost wave (a padding unit
// of the workgroup) only takes part in the panel loads and barriers.
// ---------------------------------------------------------------------------
constexpr int kAslStages = 4;

template <int NS, int F, bool A1, int W, int AD, class TE>
__device__ inline void sweep_lds(const AsArgs& p, const SweepU& sw, const SweepU& nxs, int r0, int li, int lk, int wave,
                                 bool ghost, TE (&ar)[AD][2], const int64_t* cbs,
                                 typename BElem<F>::type (*stage)[kAsSweepSlots][64])
{
  constexpr int R = 2;
  constexpr int NSW = (NS + W - 1) / W;  // slots this wave loads (at most)
  typedef typename BElem<F>::type TB;
  typedef typename std::conditional<F == 0, v4d, v4f>::type Acc;
  const int lane = li + 16 * lk;
  const TB* Bl = static_cast<const TB*>(p.B) + sw.b_off + li + static_cast<int64_t>(lk) * sw.pitch;
  const int64_t bstep = 4 * static_cast<int64_t>(sw.pitch);
  const int K = sw.k, M = sw.m;
  double* C = static_cast<double*>(p.C);
  auto colof = [&](int q, int v) { return F == 0 ? 16 * q + 4 * v + lk : 16 * q + 4 * lk + v; };
  // Lane li holds rows r0 + 2li (tile 0) and r0 + 2li + 1 (tile 1): one
  // 16-byte load covers both tiles' rows of an A or C column.
  const int rl2 = r0 + 2 * li;
  // AD steps per loop iteration (the A ring's size); steps past the panel's
  // rows are masked (k >= K) and re-read its last row
  const int nsteps = ((K + 15) >> 4) * 4, last = nsteps - 1;
  const int nstepsU = (nsteps + AD - 1) / AD * AD;
  const int kq = (K + 3) >> 2;  // steps holding contracted indices < K
  // Every global load below is unconditional: a load under a branch makes
  // the compiler drain all outstanding loads where the paths merge, which
  // would discard the A and panel prefetch.  Ghost waves load their share of
  // the panel like the others, but their A and C loads all read one fixed
  // line (rows 0..31 of column 0: L1 hits); they skip MFMAs and stores.
  Acc acc[R][NS];
  if constexpr (F == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const v2d8 c2 = *reinterpret_cast<const v2d8*>(ghost ? C + rl2 : c_src(p, cbs[colof(q, v)], rl2));
        acc[0][q][v] = c2.x;
        acc[1][q][v] = c2.y;
      }
  } else {
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int q = 0; q < NS; ++q) acc[t][q] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  // this wave's share of the panel: slots wave, wave + W, ...; a share past
  // the sweep's slots re-reads slot NS-1 into an unused stage slot
  static_assert(W * ((kAsSweepSlots + W - 1) / W) <= kAsSweepSlots, "stage slots");
  int soff[NSW];
#pragma unroll
  for (int x = 0; x < NSW; ++x) soff[x] = 16 * min(wave + W * x, NS - 1);
  auto loadS = [&](int64_t boff, TB(&g)[NSW]) {
    const TB* bs = Bl + boff;
#pragma unroll
    for (int x = 0; x < NSW; ++x) g[x] = bs[soff[x]];
  };
  auto writeS = [&](int stg, const TB(&g)[NSW]) {
#pragma unroll
    for (int x = 0; x < NSW; ++x) stage[stg][wave + W * x][lane] = g[x];
  };
  // this wave's share of steps s+1 .. s+BD of the panel, in flight in a
  // ring of BD register sets (set s % BD)
  constexpr int BD = 4;
  TB g[BD][NSW];
#pragma unroll
  for (int i = 0; i < BD; ++i) loadS(min(i, last) * bstep, g[i]);
  writeS(0, g[0]);
  writeS(1, g[1]);
  __syncthreads();
  // panel offset of the next step to load (step s + BD in phase s; stays at
  // the last row once past it)
  int64_t bnext = min(BD, last) * bstep;
  // The A stream, one running offset per lane: the next step to load is
  // s + AD - 1 in phase s; its column 4*step + lk is clamped to K - 1
  // (offsets grow with the column, so clamping the offset is the same), and
  // at step nstepsU the stream moves on to the next sweep's A block.  Ghost
  // waves use K = 1 (one fixed line).
  const TE* abase = static_cast<const TE*>(p.A) + sw.a_off;
  // (a lane whose two rows are both past the block's last row re-reads rows
  // 0 and 1 instead of running on into the next columns: its products only
  // reach padding rows of the C tiles; a lane holding the last real row and
  // one padding row keeps its pair, whose second row is the next column's
  // first)
  const int ra2 = rl2 < M ? rl2 : 0;
  uint32_t arun = static_cast<uint32_t>((4 * (AD - 1) + lk) * M + ra2);
  uint32_t alast = static_cast<uint32_t>((ghost ? 0 : K - 1) * M + ra2);
  uint32_t a4m = static_cast<uint32_t>(4 * M);
  const TE* nbase = static_cast<const TE*>(p.A) + nxs.a_off;
  const uint32_t nrun = static_cast<uint32_t>(lk * nxs.m + ra2);
  const uint32_t nlast = static_cast<uint32_t>((ghost ? 0 : nxs.k - 1) * nxs.m + ra2);
  const uint32_t n4m = static_cast<uint32_t>(4 * nxs.m);
  TB b[NS];
  // one step's MFMAs (the A values already masked where the step passes K)
  const int kfull = K >> 2;  // steps whose four contracted indices are < K
  const double apad = pad_a(p.alpha, A1);
  auto mma_full = [&](const TE(&araw)[R]) {
#pragma unroll
    for (int t = 0; t < R; ++t) {
      if constexpr (F == 0) {
        const double a = A1 ? araw[t] : araw[t] * p.alpha;
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[q], a, acc[t][q], 0, 0, 0);
      } else {
        const float a = static_cast<float>(araw[t]);
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[q], a, acc[t][q], 0, 0, 0);
      }
    }
  };
  // one step s = st + j: j = position in the 4-step iteration, so ring and
  // stage indexes are compile-time
  auto phase = [&](int st, auto jc) {
    constexpr int j = decltype(jc)::value;
    loadS(bnext, g[j % BD]);  // the set held step st+j, already in LDS
    bnext += st + j + BD < last ? bstep : 0;
    if (st + j + AD - 1 == nstepsU) {  // (wave-uniform; no load inside)
      abase = nbase;
      arun = nrun;
      alast = nlast;
      a4m = n4m;
    }
    {
      load_a_pair(abase, min(arun, alast), ar[(j + AD - 1) % AD]);
      arun += a4m;
    }
    __builtin_amdgcn_sched_barrier(0);
    // (steps past the contracted length issue no MFMAs: their products are
    // all zero; the loads, LDS writes and barriers above and below still run)
    if (!ghost && st + j < kq) {
#pragma unroll
      for (int q = 0; q < NS; ++q) b[q] = stage[j][q][lane];
      // Only the last step of a contracted length that is not a multiple of
      // 4 needs the k < K mask: a wave-uniform branch masks the A values of
      // that step (the empty asm statement keeps the branch from being
      // turned back into per-step selects), and every step runs the same
      // MFMA sequence.
      TE am[R];
#pragma unroll
      for (int t = 0; t < R; ++t) am[t] = ar[j][t];
      if (st + j >= kfull) {
        const bool kin = 4 * (st + j) + lk < K;
#pragma unroll
        for (int t = 0; t < R; ++t) {
          asm volatile("" : "+v"(am[t]));
          am[t] = kin ? am[t] : static_cast<TE>(F == 0 ? apad : 0.0);
        }
      }
      mma_full(am);
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (j % 2 == 1) {  // steps s+1, s+2 into the stages of steps s-3, s-2
      writeS((j + 1) % 4, g[(j + 1) % BD]);
      writeS((j + 2) % 4, g[(j + 2) % BD]);
      __syncthreads();
    }
  };
  static_assert(AD == 4 && BD == 4 && kAslStages == 4, "stage s % 4 of step s");
  for (int st = 0; st < nstepsU; st += AD) {
    phase(st, std::integral_constant<int, 0>());
    phase(st, std::integral_constant<int, 1>());
    phase(st, std::integral_constant<int, 2>());
    phase(st, std::integral_constant<int, 3>());
    if constexpr (AD == 8) {
      phase(st, std::integral_constant<int, 4>());
      phase(st, std::integral_constant<int, 5>());
      phase(st, std::integral_constant<int, 6>());
      phase(st, std::integral_constant<int, 7>());
    }
  }
  // write back (as in sweep_mfma; ghosts store nothing)
  if constexpr (F == 0) {
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t cb = ccol_off(cbs[colof(q, v)]);
        if (!ghost && cb >= 0 && rl2 + 1 < M) {
          v2d8 c2;
          c2.x = acc[0][q][v];
          c2.y = acc[1][q][v];
          *reinterpret_cast<v2d8*>(C + cb + rl2) = c2;
        } else if (!ghost && cb >= 0 && rl2 < M) {
          C[cb + rl2] = acc[0][q][v];
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    constexpr int G = NS < 2 ? NS : 2;
#pragma unroll
    for (int q0 = 0; q0 < NS; q0 += G) {
      int64_t cb[G][4];
      double cv[G][4][R];
#pragma unroll
      for (int x = 0; x < G; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t e = q0 + x < NS && !ghost ? cbs[colof(q0 + x, v)] : -1;
          cb[x][v] = ccol_off(e);
          const v2d8 c2 = *reinterpret_cast<const v2d8*>(c_src(p, e, rl2));
          cv[x][v][0] = c2.x;
          cv[x][v][1] = c2.y;
        }
#pragma unroll
      for (int x = 0; x < G; ++x) {
        if (q0 + x >= NS) break;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double* cq = C + (cb[x][v] >= 0 ? cb[x][v] : 0) + rl2;
          v2d8 c2;
          c2.x = cv[x][v][0] + p.alpha * static_cast<double>(acc[0][q0 + x][v]);
          c2.y = cv[x][v][1] + p.alpha * static_cast<double>(acc[1][q0 + x][v]);
          if (!ghost && cb[x][v] >= 0 && rl2 + 1 < M) *reinterpret_cast<v2d8*>(cq) = c2;
          else if (!ghost && cb[x][v] >= 0 && rl2 < M) cq[0] = c2.x;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// One workgroup task of the grouped kernel: the W units (row chunks) of one
// block row walk the row's sweep list.
template <int F, int W, bool A1, int AD, class TE>
__device__ __attribute__((always_inline)) inline void asl_task(const AsArgs& p, int wave, int lane, int li, int lk,
                                                               int nsw, int s0, int r0, bool ghost, int64_t* cbs,
                                                               typename BElem<F>::type (*stage)[kAsSweepSlots][64])
{
  SweepU cur = to_uni(p.sweeps[s0]);
  SweepU nxt = to_uni(p.sweeps[s0 + min(1, nsw - 1)]);
  AsSweep raw2 = p.sweeps[s0 + min(2, nsw - 1)];
  int64_t cc0, cc1;
  load_cols(p, cur, lane, cc0, cc1);
  TE ar[AD][2];
  {
    // first steps of the A stream (rows r0 + 2li, +1; ghosts: one fixed line)
    const TE* a = static_cast<const TE*>(p.A) + cur.a_off;
    const int kmax = ghost ? 0 : cur.k - 1;
#pragma unroll
    for (int i = 0; i < AD - 1; ++i)
      load_a_pair(a, static_cast<uint32_t>(min(4 * i + lk, kmax) * cur.m + (r0 + 2 * li < cur.m ? r0 + 2 * li : 0)),
                  ar[i]);
  }
  for (int j = 0; j < nsw; ++j) {
    cbs[lane] = cc0;
    cbs[lane + 64] = cc1;
    wave_sync();
    load_cols(p, nxt, lane, cc0, cc1);
    switch (cur.ns) {
      case 1: sweep_lds<1, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
      case 2: sweep_lds<2, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
      case 3: sweep_lds<3, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
      case 4: sweep_lds<4, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
      case 5: sweep_lds<5, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
      case 6: sweep_lds<6, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
      case 7: sweep_lds<7, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
      default: sweep_lds<8, F, A1, W, AD, TE>(p, cur, nxt, r0, li, lk, wave, ghost, ar, cbs, stage); break;
    }
    wave_sync();
    cur = nxt;
    nxt = to_uni(raw2);
    raw2 = p.sweeps[s0 + min(j + 3, nsw - 1)];
  }
}

// Workgroups of W waves over W row chunks of one block row (the LDS-shared
// panel path; F: 0 = FP64 MFMA, 1 = FP32 MFMA).  Every wave of a workgroup
// runs the same sweep list, ghosts included, so the workgroup barriers match.
// Persistent: the grid holds as many workgroups as the device keeps resident
// (a multiple of eight); workgroup b serves XCD queue b % 8 of the unit list
// (workgroup task t of queue x is the list's workgroup 8t + x) and takes its
// next task from that queue's counter until the queue is empty.  The task
// index is read by every wave after a barrier, so every exit and skip is
// taken by the whole workgroup.
template <int F, int W, bool A1, int AD, class TE>
__global__ __launch_bounds__(64 * W) void k_asl(AsArgs p)
{
  static_assert(AD == 4 || AD == 8, "A ring of 4 or 8 steps");
  typedef typename BElem<F>::type TB;
  const int wave = uni(static_cast<int>(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  __shared__ int64_t cbs_all[W][128];
  __shared__ TB stage[kAslStages][kAsSweepSlots][64];
  __shared__ unsigned long long next_task;
  int64_t* cbs = cbs_all[wave];
  const unsigned x = blockIdx.x & 7u;
  const int64_t ntask = (p.nunits / W + 7) / 8;  // workgroup tasks per queue
  for (;;) {
    __syncthreads();  // every wave is done with the previous task's LDS
    if (threadIdx.x == 0) next_task = atomicAdd(p.qctr + x, 1ull);
    __syncthreads();
    const int64_t t = uni(static_cast<int64_t>(next_task));
    if (t >= ntask) break;
    const int64_t ui = (8 * t + x) * W;
    if (ui >= p.nunits) continue;  // (the last round of the queues is short)
    const AsUnit u = p.units[ui + wave];
    const int nsw = uni(u.nsweep), s0 = uni(u.sweep0), r0 = uni(u.r0);
    const bool ghost = uni(u.pad) != 0;
    if (nsw == 0) continue;  // a whole-workgroup filler (every wave of it)
    asl_task<F, W, A1, AD, TE>(p, wave, lane, li, lk, nsw, s0, r0, ghost, cbs, stage);
  }
}

// ---------------------------------------------------------------------------
// Three-term split-BF16 sweep.  x (rounded to FP32) = hi + mid + lo, each a
// bf16 (round to nearest even); per 32-k step six products of order <= 2,
// smallest first: lo*hi, mid*mid, hi*lo, mid*hi, hi*mid, hi*hi.  The raw A
// values of the next step are loaded while the current step's products run.
// ---------------------------------------------------------------------------
__device__ inline unsigned short bf16_bits(float x)
{
  const unsigned int u = __float_as_uint(x);
  return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// The split of eight consecutive contracted indices of one row (an MFMA
// operand) with gfx950's packed conversion (v_cvt_pk_bf16_f32, round to
// nearest even like bf16_bits): word e/2 of each plane holds the parts of
// x[e] (low half) and x[e+1] (high half).
typedef float v2f_s __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf_s __attribute__((ext_vector_type(2)));
typedef unsigned int v4u_s __attribute__((ext_vector_type(4)));
__device__ inline __attribute__((unused)) void split3x8(const float (&x)[8], v8s& h, v8s& m, v8s& l)
{
  v4u_s hw, mw, lw;
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    const v2f_s v = {x[e], x[e + 1]};
    const v2bf_s hb = __builtin_convertvector(v, v2bf_s);
    const v2f_s r1 = v - __builtin_convertvector(hb, v2f_s);
    const v2bf_s mb = __builtin_convertvector(r1, v2bf_s);
    const v2f_s r2 = r1 - __builtin_convertvector(mb, v2f_s);
    const v2bf_s lb = __builtin_convertvector(r2, v2bf_s);
    hw[e / 2] = __builtin_bit_cast(unsigned int, hb);
    mw[e / 2] = __builtin_bit_cast(unsigned int, mb);
    lw[e / 2] = __builtin_bit_cast(unsigned int, lb);
  }
  h = __builtin_bit_cast(v8s, hw);
  m = __builtin_bit_cast(v8s, mw);
  l = __builtin_bit_cast(v8s, lw);
}

__device__ inline __attribute__((unused)) void split3(float x, unsigned short& h, unsigned short& m, unsigned short& l)
{
  h = bf16_bits(x);
  const float r1 = x - __uint_as_float(static_cast<unsigned int>(h) << 16);
  m = bf16_bits(r1);
  const float r2 = r1 - __uint_as_float(static_cast<unsigned int>(m) << 16);
  l = bf16_bits(r2);
}

template <int R, int NS, class TE>
__device__ inline void sweep_bf3(const AsArgs& p, const SweepU& sw, int r0, int li, int lk, const int64_t* cbs)
{
  const TE* A = static_cast<const TE*>(p.A) + sw.a_off;
  const int kp = sw.pitch;
  const unsigned short* Bl =
      static_cast<const unsigned short*>(p.B) + sw.b_off + static_cast<int64_t>(li) * kp + 8 * lk;
  const int64_t plane = p.bt_total;
  const int64_t qstep = 16 * static_cast<int64_t>(kp);
  double* C = static_cast<double*>(p.C);
  const int K = sw.k, M = sw.m;
  const int rl = r0 + li;
  v4f acc[R][NS];
#pragma unroll
  for (int t = 0; t < R; ++t)
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[t][q] = v4f{0.f, 0.f, 0.f, 0.f};
  const int nsteps = (K + kBfK - 1) / kBfK;
  TE x_nx[R][8];
  auto load_a8 = [&](int st) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int kk = kBfK * st + 8 * lk + e;
      const uint32_t off = static_cast<uint32_t>(min(kk, K - 1) * M + rl);
#pragma unroll
      for (int t = 0; t < R; ++t) {
        const TE x = A[off + 16u * t];
        x_nx[t][e] = kk < K ? x : static_cast<TE>(0);
      }
    }
  };
  load_a8(0);
  for (int st = 0; st < nsteps; ++st) {
    v8s ah[R], am[R], al[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      float xf[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) xf[e] = static_cast<float>(x_nx[t][e]);
      split3x8(xf, ah[t], am[t], al[t]);
    }
    load_a8(min(st + 1, nsteps - 1));
    const int k0 = kBfK * st;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const unsigned short* b = Bl + q * qstep + k0;
      const v8s bh = *reinterpret_cast<const v8s*>(b);
      const v8s bm = *reinterpret_cast<const v8s*>(b + plane);
      const v8s bl = *reinterpret_cast<const v8s*>(b + 2 * plane);
#pragma unroll
      for (int t = 0; t < R; ++t) {
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm, am[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, am[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm, ah[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah[t], acc[t][q], 0, 0, 0);
      }
    }
  }
  // C += alpha * (FP32 block product) in FP64; lane l holds rows r0 + 16t +
  // l%16 of columns 16q + 4(l/16) + v.
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    double cv[4][R];
    int64_t cb[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t e = cbs[16 * q + 4 * lk + v];
      cb[v] = ccol_off(e);
      const double* cq = c_src(p, e, rl);
#pragma unroll
      for (int t = 0; t < R; ++t) cv[v][t] = cq[16 * t];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double* cq = C + (cb[v] >= 0 ? cb[v] : 0) + rl;
#pragma unroll
      for (int t = 0; t < R; ++t)
        if (cb[v] >= 0 && rl + 16 * t < M) cq[16 * t] = cv[v][t] + p.alpha * static_cast<double>(acc[t][q][v]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int R, int NS, int F, bool A1, class TE>
__device__ inline void run_sweep(const AsArgs& p, const SweepU& sw, const SweepU& nx, int r0, int li, int lk,
                                 TE (&ar)[4][R], const int64_t* cbs)
{
  if constexpr (F == 2) sweep_bf3<R, NS, TE>(p, sw, r0, li, lk, cbs);
  else sweep_mfma<R, NS, F, A1, TE>(p, sw, nx, r0, li, lk, ar, cbs);
}

// One wave per unit of 16*R rows, WPG waves (units of one block row) per
// workgroup; F: 0 = FP64 MFMA, 1 = FP32 MFMA, 2 = split-BF16; A1: alpha == 1.
// Waves are independent (no workgroup barrier): an empty unit just returns.
// The records of the next two sweeps and the next sweep's column offsets are
// loaded one sweep ahead, so a sweep starts without a dependent load chain.
// The split-BF16 instances (second translation unit) hold 64 AGPR
// accumulators next to about 190 VGPRs: they are held to two waves per
// SIMD (256 registers in all), which their speed depends on.
#ifdef SHT_AS_BF3_TU
#define SHT_ASP_OCC __attribute__((amdgpu_waves_per_eu(2)))
#else
#define SHT_ASP_OCC
#endif
template <int R, int NSMAX, int F, int WPG, bool A1, class TE>
__global__ __launch_bounds__(64 * WPG) SHT_ASP_OCC void k_asp(AsArgs p)
{
  static_assert(16 * NSMAX <= 128, "two column offsets per lane");
  const int wave = WPG == 1 ? 0 : uni(static_cast<int>(threadIdx.x >> 6));
  const int64_t ui = static_cast<int64_t>(blockIdx.x) * WPG + wave;
  if (ui >= p.nunits) return;
  const AsUnit u = p.units[ui];
  const int nsw = uni(u.nsweep), s0 = uni(u.sweep0), r0 = uni(u.r0);
  if (nsw == 0 || uni(u.pad) != 0) return;  // empty or ghost unit
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  __shared__ int64_t cbs_all[WPG][128];
  int64_t* cbs = cbs_all[wave];
  SweepU cur = to_uni(p.sweeps[s0]);
  SweepU nxt = to_uni(p.sweeps[s0 + min(1, nsw - 1)]);
  AsSweep raw2 = p.sweeps[s0 + min(2, nsw - 1)];
  int64_t cc0, cc1;
  load_cols(p, cur, lane, cc0, cc1);
  TE ar[4][R];
  if constexpr (F != 2) {
    AStream<TE> s;
    s.a = static_cast<const TE*>(p.A) + cur.a_off;
    s.K = cur.k;
    s.M = cur.m;
    load_a(s, 0, r0 + li, lk, ar[0]);
    load_a(s, 1, r0 + li, lk, ar[1]);
    load_a(s, 2, r0 + li, lk, ar[2]);
  }
  for (int j = 0; j < nsw; ++j) {
    cbs[lane] = cc0;
    cbs[lane + 64] = cc1;
    wave_sync();
    load_cols(p, nxt, lane, cc0, cc1);
    switch (min(cur.ns, NSMAX)) {
      case 1: run_sweep<R, 1, F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
      case 2: run_sweep<R, cmin(2, NSMAX), F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
      case 3: run_sweep<R, cmin(3, NSMAX), F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
      case 4: run_sweep<R, cmin(4, NSMAX), F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
      case 5: run_sweep<R, cmin(5, NSMAX), F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
      case 6: run_sweep<R, cmin(6, NSMAX), F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
      case 7: run_sweep<R, cmin(7, NSMAX), F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
      default: run_sweep<R, cmin(8, NSMAX), F, A1, TE>(p, cur, nxt, r0, li, lk, ar, cbs); break;
    }
    wave_sync();  // the column offsets are rewritten by the next sweep
    cur = nxt;
    nxt = to_uni(raw2);
    raw2 = p.sweeps[s0 + min(j + 3, nsw - 1)];
  }
}

#ifdef SHT_AS_BF3_TU
// ---------------------------------------------------------------------------
// Split-BF16 for four waves over four 32-row chunks of one block row (ghost
// units pad the workgroup).  Each 32-k step's panel fragments are read from
// an FP32 panel (BtFormat::F32K; 32 bytes per lane and slot, two thirds of
// the bytes of pre-split bf16 planes) once per workgroup (wave w: slots w and
// w + 4) one step ahead, split into the three bf16 planes as they are staged
// through two LDS stages, one barrier per step.  The raw A values of the next
// step are loaded (unconditionally; ghosts read one fixed line) while the
// current step's products run, and are masked (k >= K) and split when
// consumed.  Products, their order and the FP32-per-sweep / FP64-C update are
// those of sweep_bf3, so the results are the same.
// ---------------------------------------------------------------------------
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

template <int NS, class TE>
__device__ inline void sweep_bf3_lds(const AsArgs& p, const SweepU& sw, int r0, int li, int lk, int wave, bool ghost,
                                     const int64_t* cbs, v4u (*stage)[3][kAsSweepSlots][64])
{
  constexpr int R = 2;
  constexpr int NSW = (NS + 3) / 4;
  const int lane = li + 16 * lk;
  const TE* A = static_cast<const TE*>(p.A) + sw.a_off;
  const int kp = sw.pitch;
  // FP32 panel (BtFormat::F32K): column-major, contracted indices contiguous
  const float* Bl = static_cast<const float*>(p.B) + sw.b_off + static_cast<int64_t>(li) * kp + 8 * lk;
  const int64_t qstep = 16 * static_cast<int64_t>(kp);
  double* C = static_cast<double*>(p.C);
  const int K = sw.k, M = sw.m;
  // lane li holds rows r0 + 2li (tile 0) and r0 + 2li + 1 (tile 1): one
  // 16-byte load covers both tiles' rows of an A or C column
  const int rl2 = r0 + 2 * li;
  const int ra2 = rl2 < M ? rl2 : 0;  // A rows (a pair of padding rows re-reads rows 0, 1)
  const int nsteps = (K + kBfK - 1) / kBfK;
  const int nst2 = (nsteps + 1) & ~1;  // the loop runs whole pairs of steps
  v4f acc[R][NS];
#pragma unroll
  for (int t = 0; t < R; ++t)
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[t][q] = v4f{0.f, 0.f, 0.f, 0.f};
  // this wave's panel share (a share past the sweep's slots re-reads slot
  // NS - 1 into an unused stage slot)
  int soff[NSW];  // (slot * 16 * pitch: below 2^31)
#pragma unroll
  for (int x = 0; x < NSW; ++x) soff[x] = static_cast<int>(min(wave + 4 * x, NS - 1) * qstep);
  // this wave's share of one step's panel: eight FP32 values (contracted
  // indices 32st + 8lk .. +7 of column li) per slot, split into the three
  // bf16 planes as they are written to LDS
  v4f g[NSW][2];
  auto loadP = [&](int st) {
    const float* b = Bl + kBfK * min(st, nsteps - 1);
#pragma unroll
    for (int x = 0; x < NSW; ++x) {
      g[x][0] = *reinterpret_cast<const v4f*>(b + soff[x]);
      g[x][1] = *reinterpret_cast<const v4f*>(b + soff[x] + 4);
    }
  };
  auto writeP = [&](int stg) {
#pragma unroll
    for (int x = 0; x < NSW; ++x) {
      const float xf[8] = {g[x][0][0], g[x][0][1], g[x][0][2], g[x][0][3],
                           g[x][1][0], g[x][1][1], g[x][1][2], g[x][1][3]};
      v8s h, m, l;
      split3x8(xf, h, m, l);
      stage[stg][0][wave + 4 * x][lane] = __builtin_bit_cast(v4u, h);
      stage[stg][1][wave + 4 * x][lane] = __builtin_bit_cast(v4u, m);
      stage[stg][2][wave + 4 * x][lane] = __builtin_bit_cast(v4u, l);
    }
  };
  // raw A of step st: rows rl2, rl2 + 1, contracted indices 32st + 8lk + e
  auto loadA = [&](int st, TE (&x)[R][8]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int kk = kBfK * min(st, nsteps - 1) + 8 * lk + e;
      const uint32_t off = ghost ? static_cast<uint32_t>(ra2) : static_cast<uint32_t>(min(kk, K - 1) * M + ra2);
      load_pair(A + off, x[0][e], x[1][e]);
    }
  };
  auto step = [&](int st, const TE (&x)[R][8], int stg) {
    v8s ah[R], am[R], al[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      float xf[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) xf[e] = kBfK * st + 8 * lk + e < K ? static_cast<float>(x[t][e]) : 0.f;
      split3x8(xf, ah[t], am[t], al[t]);
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const v8s bh = __builtin_bit_cast(v8s, stage[stg][0][q][lane]);
      const v8s bm = __builtin_bit_cast(v8s, stage[stg][1][q][lane]);
      const v8s bl = __builtin_bit_cast(v8s, stage[stg][2][q][lane]);
#pragma unroll
      for (int t = 0; t < R; ++t) {
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm, am[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, am[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm, ah[t], acc[t][q], 0, 0, 0);
        acc[t][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah[t], acc[t][q], 0, 0, 0);
      }
    }
  };
  TE x0[R][8], x1[R][8];
  loadP(0);
  writeP(0);
  loadA(0, x0);
  __syncthreads();
  for (int st = 0; st < nst2; st += 2) {
    // step st from stage 0 (loads of step st + 1 in flight)
    loadP(st + 1);
    loadA(st + 1, x1);
    __builtin_amdgcn_sched_barrier(0);
    if (!ghost && st < nsteps) step(st, x0, 0);
    __builtin_amdgcn_sched_barrier(0);
    writeP(1);
    __syncthreads();
    // step st + 1 from stage 1
    loadP(st + 2);
    loadA(st + 2, x0);
    __builtin_amdgcn_sched_barrier(0);
    if (!ghost && st + 1 < nsteps) step(st + 1, x1, 1);
    __builtin_amdgcn_sched_barrier(0);
    writeP(0);
    __syncthreads();
  }
  // C += alpha * (FP32 block product) in FP64 (ghosts store nothing); lane l
  // holds rows r0 + 2(l%16) + t of columns 16q + 4(l/16) + v
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    double cv[4][R];
    int64_t cb[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t e = ghost ? -1 : cbs[16 * q + 4 * lk + v];
      cb[v] = ccol_off(e);
      const v2d8 c2 = *reinterpret_cast<const v2d8*>(c_src(p, e, rl2));
      cv[v][0] = c2.x;
      cv[v][1] = c2.y;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double* cq = C + (cb[v] >= 0 ? cb[v] : 0) + rl2;
      v2d8 c2;
      c2.x = cv[v][0] + p.alpha * static_cast<double>(acc[0][q][v]);
      c2.y = cv[v][1] + p.alpha * static_cast<double>(acc[1][q][v]);
      if (cb[v] >= 0 && rl2 + 1 < M) *reinterpret_cast<v2d8*>(cq) = c2;
      else if (cb[v] >= 0 && rl2 < M) cq[0] = c2.x;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Workgroups of four waves over four 32-row chunks of one block row (the
// LDS-shared split-BF16 path).  Every wave runs the same sweep list, ghosts
// included, so the workgroup barriers match.
template <class TE>
__device__ __attribute__((always_inline)) inline void asb_body(const AsArgs& p)
{
  constexpr int W = 4;
  const int wave = uni(static_cast<int>(threadIdx.x >> 6));
  const int64_t ui = static_cast<int64_t>(blockIdx.x) * W;
  if (ui >= p.nunits) return;  // (the unit list is padded to whole workgroups)
  const AsUnit u = p.units[ui + wave];
  const int nsw = uni(u.nsweep), s0 = uni(u.sweep0), r0 = uni(u.r0);
  const bool ghost = uni(u.pad) != 0;
  if (nsw == 0) return;  // a whole-workgroup filler (every wave of it)
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  __shared__ int64_t cbs_all[W][128];
  __shared__ v4u stage[2][3][kAsSweepSlots][64];
  int64_t* cbs = cbs_all[wave];
  SweepU cur = to_uni(p.sweeps[s0]);
  SweepU nxt = to_uni(p.sweeps[s0 + min(1, nsw - 1)]);
  AsSweep raw2 = p.sweeps[s0 + min(2, nsw - 1)];
  int64_t cc0, cc1;
  load_cols(p, cur, lane, cc0, cc1);
  for (int j = 0; j < nsw; ++j) {
    cbs[lane] = cc0;
    cbs[lane + 64] = cc1;
    wave_sync();
    load_cols(p, nxt, lane, cc0, cc1);
    switch (cur.ns) {
      case 1: sweep_bf3_lds<1, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
      case 2: sweep_bf3_lds<2, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
      case 3: sweep_bf3_lds<3, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
      case 4: sweep_bf3_lds<4, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
      case 5: sweep_bf3_lds<5, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
      case 6: sweep_bf3_lds<6, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
      case 7: sweep_bf3_lds<7, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
      default: sweep_bf3_lds<8, TE>(p, cur, r0, li, lk, wave, ghost, cbs, stage); break;
    }
    wave_sync();
    cur = nxt;
    nxt = to_uni(raw2);
    raw2 = p.sweeps[s0 + min(j + 3, nsw - 1)];
  }
}

// FP32 operands: held to 168 registers, three waves per SIMD (three
// workgroups per CU with their 52 KB of LDS); the compiler then spills a few
// sweep-level values, none inside the step loop, and C1 runs 33.5 -> 31.9 ms
// against two waves per SIMD.  FP64 operands (more registers for the raw A
// values) keep two waves per SIMD: at three they spill more and run slower
// (40.2 -> 40.6 ms).
template <class TE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_asb3(AsArgs p)
{
  asb_body<TE>(p);
}
template <class TE>
__global__ __launch_bounds__(256) SHT_ASP_OCC void k_asb(AsArgs p)
{
  asb_body<TE>(p);
}
#endif  // SHT_AS_BF3_TU

#ifndef SHT_AS_BF3_TU
// ---------------------------------------------------------------------------
// A-stationary fma chain (bitwise equal to the oracle).  One wave per 64-row
// unit, lane = one C row; per chunk of 64 sweep columns (kAsxSlots slots of
// kAsxCols) the C values stay in registers while k runs ascending:
// c = fma(alpha*A(i,k), B(k,j), c).  B(k, c0..c0+7) is wave-uniform and
// contiguous in the panel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_as_exact(AsArgs p)
{
  constexpr int S = kAsxSlots, W = kAsxCols;
  if (static_cast<int64_t>(blockIdx.x) >= p.nunits) return;
  const AsUnit u = p.units[blockIdx.x];
  const int nsw = uni(u.nsweep), s0 = uni(u.sweep0), r0 = uni(u.r0);
  const double* A = static_cast<const double*>(p.A);
  const double* BT = static_cast<const double*>(p.B);
  double* C = static_cast<double*>(p.C);
  const int row = r0 + static_cast<int>(threadIdx.x);
  for (int j = 0; j < nsw; ++j) {
    const SweepU sw = to_uni(p.sweeps[s0 + j]);
    const int K = sw.k, M = sw.m;
    const bool rowok = row < M;
    const double* Ab = A + sw.a_off + min(row, M - 1);
    const int ncol = 16 * sw.ns;
    for (int c0 = 0; c0 < ncol; c0 += S * W) {
      const int ns = min(S, (ncol - c0) / W);
      double acc[S][W];
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int x = 0; x < W; ++x) {
          const int64_t e = q < ns ? p.ccol[sw.c0 + c0 + q * W + x] : -1;
          acc[q][x] = (e != -1 && rowok) ? *c_src(p, e, row) : 0.0;
        }
      const double* bt = BT + sw.b_off + c0;
      for (int k = 0; k < K; ++k) {
        double a = Ab[static_cast<int64_t>(k) * M];
        if (!p.alpha_one) a = a * p.alpha;
        const double* bk = bt + static_cast<int64_t>(k) * sw.pitch;
#pragma unroll
        for (int q = 0; q < S; ++q) {
          if (q < ns) {
#pragma unroll
            for (int x = 0; x < W; ++x) acc[q][x] = __fma_rn(a, bk[q * W + x], acc[q][x]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < S; ++q)
#pragma unroll
        for (int x = 0; x < W; ++x) {
          const int64_t cb = q < ns ? ccol_off(p.ccol[sw.c0 + c0 + q * W + x]) : -1;
          if (cb >= 0 && rowok) C[cb + row] = acc[q][x];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Panel conversion, once per execution (one workgroup per B block record).
// ---------------------------------------------------------------------------
// FP32 panel in the split-BF16 layout (BtFormat::F32K): column j of a record
// at j * pitch, contracted indices contiguous, zero from k up to the pitch.
template <class TE>
__global__ __launch_bounds__(256) void k_bt_f32k(const BtRec* recs, int64_t nrec, const TE* B, float* BT)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const BtRec rec = recs[r];
    const int kp = rec.pitch;
    const int64_t tot = static_cast<int64_t>(kp) * rec.n;
    const TE* b = B + rec.b_off;
    float* o = BT + rec.bt_off;
    for (int64_t x = threadIdx.x; x < tot; x += blockDim.x) {
      const int j = static_cast<int>(x / kp), k = static_cast<int>(x - static_cast<int64_t>(j) * kp);
      o[x] = k < rec.k ? static_cast<float>(b[k + static_cast<int64_t>(j) * rec.k]) : 0.f;
    }
  }
}

template <class T, class TE>
__global__ __launch_bounds__(256) void k_bt(const BtRec* recs, int64_t nrec, const TE* B, T* BT)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const BtRec rec = recs[r];
    const int krows = (rec.k + 15) / 16 * 16;  // zero rows up to a multiple of 16
    const int64_t tot = static_cast<int64_t>(krows) * rec.n;
    const TE* b = B + rec.b_off;
    T* o = BT + rec.bt_off;
    for (int64_t x = threadIdx.x; x < tot; x += blockDim.x) {
      const int k = static_cast<int>(x / rec.n), j = static_cast<int>(x - static_cast<int64_t>(k) * rec.n);
      o[static_cast<int64_t>(k) * rec.pitch + j] =
          k < rec.k ? static_cast<T>(b[k + static_cast<int64_t>(j) * rec.k]) : T(0);
    }
  }
}

template <class TE>
__global__ __launch_bounds__(256) void k_bt_bf3(const BtRec* recs, int64_t nrec, const TE* B, unsigned short* BT,
                                                int64_t plane)
{
  for (int64_t r = blockIdx.x; r < nrec; r += gridDim.x) {
    const BtRec rec = recs[r];
    const int kp = rec.pitch;
    const int64_t tot = static_cast<int64_t>(kp) * rec.n;
    const TE* b = B + rec.b_off;
    unsigned short* o = BT + rec.bt_off;
    for (int64_t x = threadIdx.x; x < tot; x += blockDim.x) {
      const int j = static_cast<int>(x / kp), k = static_cast<int>(x - static_cast<int64_t>(j) * kp);
      unsigned short h = 0, m = 0, l = 0;
      if (k < rec.k) split3(static_cast<float>(b[k + static_cast<int64_t>(j) * rec.k]), h, m, l);
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
#endif  // !SHT_AS_BF3_TU

}  // namespace

// This file is compiled twice: the FP64 / FP32 / fma-chain kernels with
// VGPR-form MFMAs (accumulators in ordinary VGPRs: with AGPR accumulators
// the compiler copies them between register files inside the k loops), and,
// with SHT_AS_BF3_TU defined, the split-BF16 kernels with AGPR accumulators
// (their VGPR form needs more than 256 VGPRs: one wave per SIMD).
#ifdef SHT_AS_BF3_TU

void launch_as_bf3(AsClass k, const AsArgs& a, hipStream_t s)
{
  constexpr int G = kAsGroupWaves;
  const dim3 g1(static_cast<unsigned>(a.nunits)), b1(64);
  const dim3 gg(static_cast<unsigned>((a.nunits + G - 1) / G)), bg(64 * G);
  switch (k) {
    case AsClass::Bf16x3:
      if (a.f32) hipLaunchKernelGGL((k_asp<2, 8, 2, 1, false, float>), g1, b1, 0, s, a);
      else hipLaunchKernelGGL((k_asp<2, 8, 2, 1, false, double>), g1, b1, 0, s, a);
      break;
    case AsClass::Bf16x3G:
      if (a.f32) hipLaunchKernelGGL(k_asb3<float>, gg, bg, 0, s, a);
      else hipLaunchKernelGGL(k_asb<double>, gg, bg, 0, s, a);
      break;
    default: throw Error(SHT_ERR_INVALID, "not a split-BF16 class");
  }
}

#else

int as_group_waves(AsClass k)
{
  switch (k) {
    case AsClass::Mfma64G:
    case AsClass::Mfma32G:
    case AsClass::Bf16x3G: return kAsGroupWaves;
    default: return 1;
  }
}

namespace {

__global__ void k_clear_qctr(unsigned long long* c)
{
  if (threadIdx.x < 8) c[threadIdx.x] = 0;
}

// Grid of the persistent grouped kernel f (one workgroup task per work item).
template <class Kern>
unsigned persistent_grid(Kern f, int threads, int64_t ntasks)
{
  return resident_grid(reinterpret_cast<const void*>(f), threads, ntasks);
}

}  // namespace

unsigned resident_grid(const void* f, int threads, int64_t nwork)
{
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, int> cache;
  int dev = 0;
  SHT_HIP(hipGetDevice(&dev));
  int resident = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(dev, f);
    auto it = cache.find(key);
    if (it == cache.end()) {
      int per_cu = 0, cus = 0;
      SHT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, threads, 0));
      SHT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      it = cache.emplace(key, std::max(8, per_cu * cus / 8 * 8)).first;
    }
    resident = it->second;
  }
  const int64_t want = (nwork + 7) / 8 * 8;
  return static_cast<unsigned>(std::min<int64_t>(want, resident));
}

void launch_as(AsClass k, const AsArgs& a, hipStream_t s)
{
  if (a.nunits <= 0) return;
  constexpr int G = kAsGroupWaves;
  const dim3 g1(static_cast<unsigned>(a.nunits)), b1(64);
  const dim3 bg(64 * G);
  const int64_t nwg = (a.nunits + G - 1) / G;  // grouped workgroup tasks
  auto persistent = [&](auto kern) {
    SHT_REQUIRE(a.qctr, SHT_ERR_STATE, "grouped kernel without work counters");
    hipLaunchKernelGGL(k_clear_qctr, dim3(1), dim3(64), 0, s, a.qctr);
    hipLaunchKernelGGL(kern, dim3(persistent_grid(kern, 64 * G, nwg)), bg, 0, s, a);
  };
  static_assert(kAsUnitRows == 32 && kAsUnitRowsW == 64, "unit rows: two MFMA tiles; 64 lanes");
  static_assert(kAsSweepSlots == 8, "sweep slots");
  SHT_REQUIRE(!a.f32 || k == AsClass::Mfma32 || k == AsClass::Mfma32G || k == AsClass::Mfma32N ||
                  k == AsClass::Bf16x3 || k == AsClass::Bf16x3G,
              SHT_ERR_STATE, "FP32 operands reach an FP64 kernel class");
  // the grouped classes read W unit records per workgroup
  SHT_REQUIRE(as_group_waves(k) == 1 || a.nunits % G == 0, SHT_ERR_STATE, "grouped unit list not padded");
  switch (k) {
    case AsClass::Mfma64:
      if (a.alpha_one) hipLaunchKernelGGL((k_asp<2, 8, 0, 1, true, double>), g1, b1, 0, s, a);
      else hipLaunchKernelGGL((k_asp<2, 8, 0, 1, false, double>), g1, b1, 0, s, a);
      break;
    case AsClass::Mfma64G:
      if (a.alpha_one) persistent(k_asl<0, G, true, 4, double>);
      else persistent(k_asl<0, G, false, 4, double>);
      break;
    case AsClass::Mfma32:
      if (a.f32) hipLaunchKernelGGL((k_asp<2, 8, 1, 1, false, float>), g1, b1, 0, s, a);
      else hipLaunchKernelGGL((k_asp<2, 8, 1, 1, false, double>), g1, b1, 0, s, a);
      break;
    case AsClass::Mfma32G:
      if (a.f32) persistent(k_asl<1, G, false, 4, float>);
      else persistent(k_asl<1, G, false, 4, double>);
      break;
    case AsClass::Bf16x3:
    case AsClass::Bf16x3G: launch_as_bf3(k, a, s); break;
    case AsClass::Mfma64N:
      if (a.alpha_one) hipLaunchKernelGGL((k_asp<4, kAsNarrowSlots, 0, 1, true, double>), g1, b1, 0, s, a);
      else hipLaunchKernelGGL((k_asp<4, kAsNarrowSlots, 0, 1, false, double>), g1, b1, 0, s, a);
      break;
    case AsClass::Mfma32N:
      if (a.f32) hipLaunchKernelGGL((k_asp<4, kAsNarrowSlots, 1, 1, false, float>), g1, b1, 0, s, a);
      else hipLaunchKernelGGL((k_asp<4, kAsNarrowSlots, 1, 1, false, double>), g1, b1, 0, s, a);
      break;
    case AsClass::Exact: hipLaunchKernelGGL(k_as_exact, g1, b1, 0, s, a); break;
    default: throw Error(SHT_ERR_INVALID, "unknown A-stationary class");
  }
  SHT_HIP(hipGetLastError());
}

void launch_bt(BtFormat f, const BtRec* recs, int64_t nrec, const void* B, bool b_f32, void* BT, int64_t bt_total,
               hipStream_t s)
{
  if (nrec <= 0) return;
  const int g = static_cast<int>(nrec < 65536 ? nrec : 65536);
  const double* b64 = static_cast<const double*>(B);
  const float* b32 = static_cast<const float*>(B);
  SHT_REQUIRE(!b_f32 || f != BtFormat::F64T, SHT_ERR_STATE, "FP32 operand for an FP64 panel");
  switch (f) {
    case BtFormat::F64T:
      hipLaunchKernelGGL((k_bt<double, double>), dim3(g), dim3(256), 0, s, recs, nrec, b64, static_cast<double*>(BT));
      break;
    case BtFormat::F32T:
      if (b_f32)
        hipLaunchKernelGGL((k_bt<float, float>), dim3(g), dim3(256), 0, s, recs, nrec, b32, static_cast<float*>(BT));
      else
        hipLaunchKernelGGL((k_bt<float, double>), dim3(g), dim3(256), 0, s, recs, nrec, b64, static_cast<float*>(BT));
      break;
    case BtFormat::F32K:
      if (b_f32)
        hipLaunchKernelGGL(k_bt_f32k<float>, dim3(g), dim3(256), 0, s, recs, nrec, b32, static_cast<float*>(BT));
      else
        hipLaunchKernelGGL(k_bt_f32k<double>, dim3(g), dim3(256), 0, s, recs, nrec, b64, static_cast<float*>(BT));
      break;
    case BtFormat::BF16X3:
      if (b_f32)
        hipLaunchKernelGGL(k_bt_bf3<float>, dim3(g), dim3(256), 0, s, recs, nrec, b32,
                           static_cast<unsigned short*>(BT), bt_total);
      else
        hipLaunchKernelGGL(k_bt_bf3<double>, dim3(g), dim3(256), 0, s, recs, nrec, b64,
                           static_cast<unsigned short*>(BT), bt_total);
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

#endif  // SHT_AS_BF3_TU

}  // namespace sht
