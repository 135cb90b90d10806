// ORACLE — CPU restatement of shtensor's block-sparse contraction semantics.
// Test infrastructure only: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py, never by the product library.
//
// Every function names the reference source (relative to the reference tree)
// whose behaviour it restates.  Compiled with -ffp-contract=off so that the
// only fused multiply-adds are the explicit ones.

#include "oracle.h"

#include <immintrin.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

void set_err(char* err, int errlen, const std::string& msg)
{
  if (err && errlen > 0) {
    std::snprintf(err, (size_t)errlen, "%s", msg.c_str());
  }
}

// ---------------------------------------------------------------------------
// Expression analysis — src/Basics/ContractInfo.h:68-346
// ---------------------------------------------------------------------------
struct Maps {
  std::vector<int> arow, acol, brow, bcol, crow, ccol;
};

struct Tok {
  int type;  // 0 indices, 1 comma, 2 arrow
  int pos;
  std::string text;
};

bool analyse(const std::string& expr, const std::vector<int>& da, const std::vector<int>& db,
             const std::vector<int>& dc, Maps& m, std::string& err)
{
  if (expr.empty()) { err = "Equation is empty"; return false; }
  // tokenizer (ContractInfo.h:99-147)
  std::vector<Tok> toks;
  const int len = (int)expr.size();
  auto is_idx = [](char ch) { return std::isalpha((unsigned char)ch) || ch == '_'; };
  for (int p = 0; p < len; ++p) {
    char ch = expr[p];
    if (std::isspace((unsigned char)ch)) continue;
    if (is_idx(ch)) {
      int q = p;
      std::string s;
      while (q < len && is_idx(expr[q])) s += expr[q++];
      toks.push_back({0, p, s});
      p = q - 1;
      continue;
    }
    if (ch == '-' && p + 1 < len && expr[p + 1] == '>') { toks.push_back({2, p, "->"}); ++p; continue; }
    if (ch == ',') { toks.push_back({1, p, ","}); continue; }
    err = std::string("Unknown character in equation at position ") + std::to_string(p) + ": " + ch;
    return false;
  }
  // grammar: indices , indices -> indices (ContractInfo.h:155-214)
  const int want[5] = {0, 1, 0, 2, 0};
  const char* names[3] = {"index", "\",\"", "->"};
  for (int t = 0; t < 5; ++t) {
    if (t >= (int)toks.size()) {
      err = std::string("Unexpected end of expression, expected ") + names[want[t]];
      return false;
    }
    if (toks[t].type != want[t]) {
      err = std::string("Expected ") + names[want[t]] + " at position " + std::to_string(toks[t].pos);
      return false;
    }
  }
  if (toks.size() != 5) { err = "Invalid syntax"; return false; }

  // axis ids by first appearance, A then B then C (ContractInfo.h:175-206)
  std::map<char, int> ids;
  auto number = [&ids](const std::string& s) {
    std::vector<int> v;
    for (char ch : s) {
      auto it = ids.find(ch);
      if (it == ids.end()) { int nid = (int)ids.size(); ids[ch] = nid; v.push_back(nid); }
      else v.push_back(it->second);
    }
    return v;
  };
  std::vector<int> xa = number(toks[0].text), xb = number(toks[2].text), xc = number(toks[4].text);
  std::set<int> sa(xa.begin(), xa.end()), sb(xb.begin(), xb.end()), sc(xc.begin(), xc.end());
  if (sa.size() != xa.size() || sb.size() != xb.size() || sc.size() != xc.size()) {
    err = "Duplicate tensor indices"; return false;
  }
  std::set<int> con, n1, n2;
  for (int x : sa) (sb.count(x) ? con : n1).insert(x);
  for (int x : sb) if (!sa.count(x)) n2.insert(x);
  if (n1.empty() || n2.empty()) { err = "Element-wise multiplication not supported"; return false; }
  if (con.empty()) { err = "No indices found over which to contract"; return false; }
  // The reference checks only the count; an output index absent from A and B
  // is rejected here as the same mismatch.
  std::set<int> free_ax(n1); free_ax.insert(n2.begin(), n2.end());
  if (n1.size() + n2.size() != sc.size() || free_ax != sc) { err = "Dimension mismatch"; return false; }
  if (da.size() != xa.size()) { err = "Number of sizes does not match number of indices of tensor A"; return false; }
  if (db.size() != xb.size()) { err = "Number of sizes does not match number of indices of tensor B"; return false; }
  if (dc.size() != xc.size()) { err = "Number of sizes does not match number of indices of tensor C"; return false; }

  auto positions = [](const std::vector<int>& all, const std::set<int>& sel) {
    std::vector<int> out;
    for (int g : sel) {
      for (int i = 0; i < (int)all.size(); ++i) if (all[i] == g) { out.push_back(i); break; }
    }
    return out;
  };
  m.arow = positions(xa, n1); m.acol = positions(xa, con);
  m.brow = positions(xb, con); m.bcol = positions(xb, n2);
  m.crow = positions(xc, n1); m.ccol = positions(xc, n2);

  // size checks (ContractInfo.h:293-328)
  auto check = [&err](const std::vector<int>& m1, const std::vector<int>& m2, const std::vector<int>& d1,
                      const std::vector<int>& d2, const char* nm1, const char* nm2) {
    for (size_t i = 0; i < m1.size(); ++i) {
      if (d1[m1[i]] != d2[m2[i]]) {
        err = std::string("Sizes for dimensions ") + std::to_string(m1[i]) + " in " + nm1 +
              " and for dimension " + std::to_string(m2[i]) + " in " + nm2 + " do not match.";
        return false;
      }
    }
    return true;
  };
  if (!check(m.arow, m.crow, da, dc, "tensor A", "tensor C")) return false;
  if (!check(m.bcol, m.ccol, db, dc, "tensor B", "tensor C")) return false;
  if (!check(m.acol, m.brow, da, db, "tensor A", "tensor B")) return false;
  return true;
}

// create_scatter_vector, ContractInfo.h:38-66 (column-major strides Utils.h:80-94)
std::vector<int64_t> scatter(const std::vector<int>& map, const std::vector<int>& sizes)
{
  std::vector<int64_t> stride(sizes.size(), 1);
  for (size_t d = 1; d < sizes.size(); ++d) stride[d] = stride[d - 1] * sizes[d - 1];
  int64_t n = 1;
  for (int ax : map) n *= sizes[ax];
  std::vector<int64_t> out((size_t)n, 0);
  for (int64_t idx = 0; idx < n; ++idx) {
    int64_t rem = idx, off = 0;
    for (int ax : map) { off += (rem % sizes[ax]) * stride[ax]; rem /= sizes[ax]; }
    out[(size_t)idx] = off;
  }
  return out;
}

int64_t prod(const std::vector<int>& v, const std::vector<int>& map)
{
  int64_t p = 1;
  for (int ax : map) p *= v[ax];
  return p;
}

// ---------------------------------------------------------------------------
// XMM chain — src/Kernel/KernelImpl.cpp:553-735, AvxHelper.cpp:250-285
// ---------------------------------------------------------------------------
template <class T>
bool bytes_zero(T v)
{
  unsigned char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  for (unsigned char x : b) if (x) return false;
  return true;
}

template <class T>
T fused(T a, T b, T c);
template <> double fused<double>(double a, double b, double c) { return std::fma(a, b, c); }
template <> float fused<float>(float a, float b, float c) { return std::fmaf(a, b, c); }

template <class T>
int kernel_xmm(const char* expr, const int* sa, int na, const int* sb, int nb, const int* sc, int nc,
               T alpha, T beta, const T* a, const T* b, T* c, int64_t nb_ops, char* err, int errlen)
{
  std::vector<int> da(sa, sa + na), db(sb, sb + nb), dc(sc, sc + nc);
  Maps m;
  std::string e;
  if (!analyse(expr, da, db, dc, m, e)) { set_err(err, errlen, e); return 0; }
  auto ar = scatter(m.arow, da), ac = scatter(m.acol, da);
  auto br = scatter(m.brow, db), bc = scatter(m.bcol, db);
  auto cr = scatter(m.crow, dc), cc = scatter(m.ccol, dc);
  const int64_t M = (int64_t)ar.size(), K = (int64_t)ac.size(), N = (int64_t)bc.size();
  const int64_t sza = M * K, szb = K * N, szc = M * N;
  const bool az = bytes_zero(alpha), bz = bytes_zero(beta);
  // nb_ops <= 0 is a no-op (the reference's do-while runs one op; fixed).
  for (int64_t op = 0; op < nb_ops; ++op) {
    const T* pa = a + op * sza;
    const T* pb = b + op * szb;
    T* pc = c + op * szc;
    for (int64_t j = 0; j < N; ++j) {
      for (int64_t i = 0; i < M; ++i) {
        const int64_t oc = cr[i] + cc[j];
        T acc = bz ? T(0) : pc[oc] * beta;
        if (!az) {
          for (int64_t k = 0; k < K; ++k) {
            T av = pa[ar[i] + ac[k]] * alpha;
            acc = fused<T>(av, pb[br[k] + bc[j]], acc);
          }
        }
        pc[oc] = acc;
      }
    }
  }
  return 1;
}

// netlib reference BLAS dgemm, TRANSA='N', TRANSB='N' (BLAS 3.9 dgemm.f):
//   for j: if beta==0 C(:,j)=0 elif beta!=1 C(:,j)=beta*C(:,j)
//          for l: temp = alpha*B(l,j); for i: C(i,j) = C(i,j) + temp*A(i,l)
void dgemm_nn(int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
              double beta, double* C, int ldc)
{
  for (int j = 0; j < n; ++j) {
    if (beta == 0.0) {
      for (int i = 0; i < m; ++i) C[i + (int64_t)j * ldc] = 0.0;
    } else if (beta != 1.0) {
      for (int i = 0; i < m; ++i) C[i + (int64_t)j * ldc] = beta * C[i + (int64_t)j * ldc];
    }
    for (int l = 0; l < k; ++l) {
      const double temp = alpha * B[l + (int64_t)j * ldb];
      for (int i = 0; i < m; ++i) {
        const double prodv = temp * A[i + (int64_t)l * lda];
        C[i + (int64_t)j * ldc] = C[i + (int64_t)j * ldc] + prodv;
      }
    }
  }
}

// Utils::reshape, src/Basics/Utils.h:315-359
template <class T>
void reshape(const T* in, const std::vector<int>& sizes, const std::vector<int>& order, T* out)
{
  const int dim = (int)sizes.size();
  std::vector<int> rs(dim);
  for (int i = 0; i < dim; ++i) rs[i] = sizes[order[i]];
  std::vector<int64_t> ss(dim, 1), rst(dim, 1);
  for (int i = 1; i < dim; ++i) { ss[i] = ss[i - 1] * sizes[i - 1]; rst[i] = rst[i - 1] * rs[i - 1]; }
  int64_t n = 1;
  for (int s : sizes) n *= s;
  std::vector<int64_t> sidx(dim);
  for (int64_t s = 0; s < n; ++s) {
    int64_t t = s;
    for (int i = dim - 1; i >= 0; --i) { sidx[i] = t / ss[i]; t = t % ss[i]; }
    int64_t r = 0;
    for (int i = 0; i < dim; ++i) r += sidx[order[i]] * rst[i];
    out[r] = in[s];
  }
}

// ---------------------------------------------------------------------------
// Work-stealing parallel-for — src/Basics/ThreadPool.cpp:57-196 semantics:
// static split by divide_equally (Utils.h:273-287), per-thread atomic cursor,
// then random-victim stealing until every thread has drained its range.
// ---------------------------------------------------------------------------
template <class F>
void pool_run(int64_t start, int64_t end, int nthreads, F&& fn)
{
  if (nthreads <= 1 || end - start <= 1) {
    for (int64_t t = start; t < end; ++t) fn(t);
    return;
  }
  const int64_t n = end - start;
  std::vector<std::atomic<int64_t>> cur(nthreads);
  std::vector<int64_t> lim(nthreads);
  for (int t = 0; t < nthreads; ++t) {
    const int64_t rest = n % nthreads, lower = n / nthreads;
    const int64_t off = (t == 0) ? 0 : rest;
    const int64_t sz = (t == 0) ? lower + rest : lower;
    const int64_t s = t * lower + off;
    cur[t].store(start + s);
    lim[t] = start + s + sz;
  }
  std::atomic<int> done{0};
  auto worker = [&](int id) {
    uint64_t rng = 0x2545F4914F6CDD1DULL * (uint64_t)(id + 1);
    int64_t task = cur[id].fetch_add(1);
    while (task < lim[id]) { fn(task); task = cur[id].fetch_add(1); }
    done.fetch_add(1);
    while (done.load() != nthreads) {
      rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
      const int v = (int)(rng % (uint64_t)nthreads);
      if (cur[v].load() >= lim[v]) {
        // every range drained: stop stealing
        bool any = false;
        for (int t = 0; t < nthreads; ++t) if (cur[t].load() < lim[t]) { any = true; break; }
        if (!any) break;
        continue;
      }
      task = cur[v].fetch_add(1);
      if (task >= lim[v]) continue;
      fn(task);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; ++t) th.emplace_back(worker, t);
  worker(0);
  for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Block-sparse helpers
// ---------------------------------------------------------------------------
struct TView {
  int N;
  std::vector<int> nblk;
  std::vector<std::vector<int>> bs;  // block sizes per mode
  std::vector<int64_t> bstride;      // rolled-index strides (mode 0 fastest)
  explicit TView(const orc_bst* t) : N(t->N), nblk(t->nblk, t->nblk + t->N), bs(t->N), bstride(t->N, 1)
  {
    int off = 0;
    for (int d = 0; d < N; ++d) {
      bs[d].assign(t->bsizes + off, t->bsizes + off + nblk[d]);
      off += nblk[d];
      if (d) bstride[d] = bstride[d - 1] * nblk[d - 1];
    }
  }
  void coords(int64_t id, std::vector<int>& c) const
  {
    c.resize(N);
    for (int d = N - 1; d >= 0; --d) { c[d] = (int)(id / bstride[d]); id %= bstride[d]; }
  }
  std::vector<int> dims(int64_t id) const
  {
    std::vector<int> c, out(N);
    coords(id, c);
    for (int d = 0; d < N; ++d) out[d] = bs[d][c[d]];
    return out;
  }
};

struct Plan {
  Maps m;
  std::vector<int64_t> ta, tb, tc, tk;  // task arrays (positions), sorted by (cid, kkey)
  std::vector<int64_t> cids;            // C pattern
  std::vector<int64_t> cstart;          // task range per C position (size ncid+1)
};

bool make_plan(const char* expr, const orc_bst* A, const orc_bst* B, const orc_bst* C, Plan& P,
               std::string& e)
{
  if (!A || !B || !C) { e = "null tensor"; return false; }
  TView va(A), vb(B), vc(C);
  if (!analyse(expr, va.nblk, vb.nblk, vc.nblk, P.m, e)) return false;
  const Maps& m = P.m;
  // identified modes must share their block-size vectors
  for (size_t r = 0; r < m.arow.size(); ++r)
    if (va.bs[m.arow[r]] != vc.bs[m.crow[r]]) { e = "Block sizes of A and C do not match"; return false; }
  for (size_t r = 0; r < m.bcol.size(); ++r)
    if (vb.bs[m.bcol[r]] != vc.bs[m.ccol[r]]) { e = "Block sizes of B and C do not match"; return false; }
  for (size_t r = 0; r < m.acol.size(); ++r)
    if (va.bs[m.acol[r]] != vb.bs[m.brow[r]]) { e = "Block sizes of A and B do not match"; return false; }

  auto key = [](const std::vector<int>& c, const std::vector<int>& map, const std::vector<int>& nb) {
    int64_t k = 0, s = 1;
    for (int ax : map) { k += c[ax] * s; s *= nb[ax]; }
    return k;
  };
  std::vector<int> c;
  // group B by contracted key
  std::map<int64_t, std::vector<int64_t>> bk;
  for (int64_t p = 0; p < B->nnz; ++p) {
    vb.coords(B->indices[p], c);
    bk[key(c, m.brow, vb.nblk)].push_back(p);
  }
  struct T4 { int64_t cid, kk, a, b; };
  std::vector<T4> tasks;
  std::vector<int> cb, ccoord(vc.N);
  for (int64_t p = 0; p < A->nnz; ++p) {
    va.coords(A->indices[p], c);
    auto it = bk.find(key(c, m.acol, va.nblk));
    if (it == bk.end()) continue;
    for (int64_t q : it->second) {
      vb.coords(B->indices[q], cb);
      for (size_t r = 0; r < m.crow.size(); ++r) ccoord[m.crow[r]] = c[m.arow[r]];
      for (size_t r = 0; r < m.ccol.size(); ++r) ccoord[m.ccol[r]] = cb[m.bcol[r]];
      int64_t cid = 0;
      for (int d = 0; d < vc.N; ++d) cid += ccoord[d] * vc.bstride[d];
      tasks.push_back({cid, it->first, p, q});
    }
  }
  std::sort(tasks.begin(), tasks.end(), [](const T4& x, const T4& y) {
    return x.cid != y.cid ? x.cid < y.cid : x.kk < y.kk;
  });
  // C pattern
  if (C->nnz == 0) {
    for (const auto& t : tasks) if (P.cids.empty() || P.cids.back() != t.cid) P.cids.push_back(t.cid);
  } else {
    P.cids.assign(C->indices, C->indices + C->nnz);
  }
  P.cstart.assign(P.cids.size() + 1, 0);
  size_t ci = 0;
  for (const auto& t : tasks) {
    while (ci < P.cids.size() && P.cids[ci] < t.cid) { ++ci; P.cstart[ci] = (int64_t)P.ta.size(); }
    if (ci >= P.cids.size() || P.cids[ci] != t.cid) continue;  // dropped: outside kept pattern
    P.ta.push_back(t.a); P.tb.push_back(t.b); P.tc.push_back((int64_t)ci); P.tk.push_back(t.kk);
  }
  while (ci < P.cids.size()) { ++ci; P.cstart[ci] = (int64_t)P.ta.size(); }
  return true;
}

bool contiguous_prefix(const std::vector<int>& map, int first)
{
  for (size_t r = 0; r < map.size(); ++r) if (map[r] != first + (int)r) return false;
  return true;
}

// One task's contribution with the XMM chain, C accumulator in-place.
void chain_task(const double* pa, const std::vector<int>& da, const double* pb, const std::vector<int>& db,
                double* pc, const std::vector<int>& dc, const Maps& m, double alpha)
{
  const int64_t M = prod(da, m.arow), K = prod(da, m.acol), N = prod(db, m.bcol);
  const bool nn = contiguous_prefix(m.arow, 0) && contiguous_prefix(m.acol, (int)m.arow.size()) &&
                  contiguous_prefix(m.brow, 0) && contiguous_prefix(m.bcol, (int)m.brow.size()) &&
                  contiguous_prefix(m.crow, 0) && contiguous_prefix(m.ccol, (int)m.crow.size());
  if (nn) {
    // AVX2 restatement of the JIT's i-chunk loop: 4 rows per register,
    // fma chain over k with b broadcast; scalar fma epilogue.
    const __m256d valpha = _mm256_set1_pd(alpha);
    const bool unit = (alpha == 1.0);
    for (int64_t j = 0; j < N; ++j) {
      const double* bj = pb + j * K;
      double* cj = pc + j * M;
      int64_t i = 0;
      for (; i + 16 <= M; i += 16) {
        __m256d c0 = _mm256_loadu_pd(cj + i), c1 = _mm256_loadu_pd(cj + i + 4);
        __m256d c2 = _mm256_loadu_pd(cj + i + 8), c3 = _mm256_loadu_pd(cj + i + 12);
        const double* ak = pa + i;
        for (int64_t k = 0; k < K; ++k, ak += M) {
          const __m256d bv = _mm256_broadcast_sd(bj + k);
          __m256d a0 = _mm256_loadu_pd(ak), a1 = _mm256_loadu_pd(ak + 4);
          __m256d a2 = _mm256_loadu_pd(ak + 8), a3 = _mm256_loadu_pd(ak + 12);
          if (!unit) {
            a0 = _mm256_mul_pd(a0, valpha); a1 = _mm256_mul_pd(a1, valpha);
            a2 = _mm256_mul_pd(a2, valpha); a3 = _mm256_mul_pd(a3, valpha);
          }
          c0 = _mm256_fmadd_pd(a0, bv, c0); c1 = _mm256_fmadd_pd(a1, bv, c1);
          c2 = _mm256_fmadd_pd(a2, bv, c2); c3 = _mm256_fmadd_pd(a3, bv, c3);
        }
        _mm256_storeu_pd(cj + i, c0); _mm256_storeu_pd(cj + i + 4, c1);
        _mm256_storeu_pd(cj + i + 8, c2); _mm256_storeu_pd(cj + i + 12, c3);
      }
      for (; i + 4 <= M; i += 4) {
        __m256d c0 = _mm256_loadu_pd(cj + i);
        const double* ak = pa + i;
        for (int64_t k = 0; k < K; ++k, ak += M) {
          __m256d a0 = _mm256_loadu_pd(ak);
          if (!unit) a0 = _mm256_mul_pd(a0, valpha);
          c0 = _mm256_fmadd_pd(a0, _mm256_broadcast_sd(bj + k), c0);
        }
        _mm256_storeu_pd(cj + i, c0);
      }
      for (; i < M; ++i) {
        double acc = cj[i];
        for (int64_t k = 0; k < K; ++k) acc = std::fma(pa[i + k * M] * alpha, bj[k], acc);
        cj[i] = acc;
      }
    }
    return;
  }
  auto ar = scatter(m.arow, da), ac = scatter(m.acol, da);
  auto br = scatter(m.brow, db), bc = scatter(m.bcol, db);
  auto cr = scatter(m.crow, dc), cc = scatter(m.ccol, dc);
  for (int64_t j = 0; j < N; ++j)
    for (int64_t i = 0; i < M; ++i) {
      double acc = pc[cr[i] + cc[j]];
      for (int64_t k = 0; k < K; ++k) acc = std::fma(pa[ar[i] + ac[k]] * alpha, pb[br[k] + bc[j]], acc);
      pc[cr[i] + cc[j]] = acc;
    }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int orc_contract_info(const char* expr, const int* sa, int na, const int* sb, int nb, const int* sc,
                      int nc, int* maps_out, char* err, int errlen)
{
  Maps m;
  std::string e;
  if (!analyse(expr ? expr : "", std::vector<int>(sa, sa + na), std::vector<int>(sb, sb + nb),
               std::vector<int>(sc, sc + nc), m, e)) {
    set_err(err, errlen, e);
    return 0;
  }
  const std::vector<int>* v[6] = {&m.arow, &m.acol, &m.brow, &m.bcol, &m.crow, &m.ccol};
  for (int r = 0; r < 6; ++r) {
    maps_out[r * 17] = (int)v[r]->size();
    for (size_t i = 0; i < v[r]->size() && i < 16; ++i) maps_out[r * 17 + 1 + i] = (*v[r])[i];
  }
  return 1;
}

int orc_scatter(const int* map, int nmap, const int* sizes, int nsizes, int* out)
{
  auto s = scatter(std::vector<int>(map, map + nmap), std::vector<int>(sizes, sizes + nsizes));
  for (size_t i = 0; i < s.size(); ++i) out[i] = (int)s[i];
  return (int)s.size();
}

int orc_kernel_xmm_f64(const char* expr, const int* sa, int na, const int* sb, int nb, const int* sc,
                       int nc, double alpha, double beta, const double* a, const double* b, double* c,
                       int64_t nb_ops, char* err, int errlen)
{
  return kernel_xmm<double>(expr, sa, na, sb, nb, sc, nc, alpha, beta, a, b, c, nb_ops, err, errlen);
}

int orc_kernel_xmm_f32(const char* expr, const int* sa, int na, const int* sb, int nb, const int* sc,
                       int nc, float alpha, float beta, const float* a, const float* b, float* c,
                       int64_t nb_ops, char* err, int errlen)
{
  return kernel_xmm<float>(expr, sa, na, sb, nb, sc, nc, alpha, beta, a, b, c, nb_ops, err, errlen);
}

int orc_kernel_lapack_f64(const char* expr, const int* sa, int na, const int* sb, int nb, const int* sc,
                          int nc, double alpha, double beta, const double* a, const double* b, double* c,
                          int64_t nb_ops, char* err, int errlen)
{
  std::vector<int> da(sa, sa + na), db(sb, sb + nb), dc(sc, sc + nc);
  Maps m;
  std::string e;
  if (!analyse(expr, da, db, dc, m, e)) { set_err(err, errlen, e); return 0; }
  const int64_t sza = std::accumulate(da.begin(), da.end(), (int64_t)1, std::multiplies<int64_t>());
  const int64_t szb = std::accumulate(db.begin(), db.end(), (int64_t)1, std::multiplies<int64_t>());
  const int64_t szc = std::accumulate(dc.begin(), dc.end(), (int64_t)1, std::multiplies<int64_t>());
  std::vector<double> ba((size_t)sza), bb((size_t)szb), bc((size_t)szc);
  std::vector<int> oa(m.arow), ob(m.brow), oc(m.crow);
  oa.insert(oa.end(), m.acol.begin(), m.acol.end());
  ob.insert(ob.end(), m.bcol.begin(), m.bcol.end());
  oc.insert(oc.end(), m.ccol.begin(), m.ccol.end());
  const int M = (int)prod(da, m.arow), N = (int)prod(db, m.bcol), K = (int)prod(da, m.acol);
  std::vector<int> ocr(oc.size()), scr(oc.size());
  for (size_t i = 0; i < oc.size(); ++i) { ocr[oc[i]] = (int)i; scr[i] = dc[oc[i]]; }
  for (int64_t op = 0; op < nb_ops; ++op) {
    reshape(a + op * sza, da, oa, ba.data());
    reshape(b + op * szb, db, ob, bb.data());
    reshape(c + op * szc, dc, oc, bc.data());
    dgemm_nn(M, N, K, alpha, ba.data(), M, bb.data(), K, beta, bc.data(), M);
    reshape(bc.data(), scr, ocr, c + op * szc);
  }
  return 1;
}

void orc_reshape_f64(const double* in, const int* sizes, int n, const int* order, double* out)
{
  reshape(in, std::vector<int>(sizes, sizes + n), std::vector<int>(order, order + n), out);
}

int64_t orc_reserve(int N, const int* nblk, const int* bsizes, const int* idx, int64_t nb,
                    int64_t* indices_out, int64_t* offsets_out, int64_t* rows_out, char* err, int errlen)
{
  std::vector<int64_t> stride(N, 1);
  for (int d = 1; d < N; ++d) stride[d] = stride[d - 1] * nblk[d - 1];
  std::vector<int> boff(N, 0);
  for (int d = 1; d < N; ++d) boff[d] = boff[d - 1] + nblk[d - 1];
  std::vector<int64_t> rolled((size_t)nb);
  std::vector<int64_t> per_row(nblk[0], 0);
  for (int64_t b = 0; b < nb; ++b) {
    int64_t r = 0;
    for (int d = 0; d < N; ++d) {
      const int x = idx[(int64_t)d * nb + b];
      if (x < 0 || x >= nblk[d]) { set_err(err, errlen, "reserve: block index out of range"); return -1; }
      r += x * stride[d];
    }
    rolled[(size_t)b] = r;
    per_row[idx[b]]++;
  }
  std::vector<int64_t> sorted(rolled);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) {
    set_err(err, errlen, "reserve: duplicate block index");
    return -1;
  }
  int64_t off = 0;
  for (int64_t b = 0; b < nb; ++b) {
    int64_t id = sorted[(size_t)b], vol = 1;
    for (int d = N - 1; d >= 0; --d) { vol *= bsizes[boff[d] + id / stride[d]]; id %= stride[d]; }
    indices_out[b] = sorted[(size_t)b];
    offsets_out[b] = off;
    off += vol;
  }
  if (rows_out) {
    int64_t s = 0;
    for (int r = 0; r < nblk[0]; ++r) { rows_out[r] = s; s += per_row[r]; }
  }
  return off;
}

int64_t orc_tasks(const char* expr, const orc_bst* A, const orc_bst* B, const orc_bst* C,
                  int64_t* tasks_out, int64_t cap, int64_t* c_ids_out, int64_t* nc_out, int64_t* kkey_out,
                  char* err, int errlen)
{
  Plan P;
  std::string e;
  if (!make_plan(expr, A, B, C, P, e)) { set_err(err, errlen, e); return -1; }
  const int64_t nt = (int64_t)P.ta.size();
  if (nc_out) *nc_out = (int64_t)P.cids.size();
  if (cap >= nt) {
    for (int64_t t = 0; t < nt; ++t) {
      tasks_out[3 * t] = P.ta[t]; tasks_out[3 * t + 1] = P.tb[t]; tasks_out[3 * t + 2] = P.tc[t];
      if (kkey_out) kkey_out[t] = P.tk[t];
    }
    if (c_ids_out) std::copy(P.cids.begin(), P.cids.end(), c_ids_out);
  }
  return nt;
}

int64_t orc_contract_f64(const char* expr, double alpha, const orc_bst* A, const orc_bst* B, double beta,
                         orc_bst* C, int mode, const int64_t* subset, int64_t nsubset, int nthreads,
                         char* err, int errlen)
{
  return orc_contract_f64_t(expr, alpha, A, B, beta, C, mode, subset, nsubset, nthreads, err, errlen, nullptr,
                            nullptr);
}

int64_t orc_contract_f64_t(const char* expr, double alpha, const orc_bst* A, const orc_bst* B, double beta,
                           orc_bst* C, int mode, const int64_t* subset, int64_t nsubset, int nthreads,
                           char* err, int errlen, double* plan_s, double* compute_s)
{
  const auto tp0 = std::chrono::steady_clock::now();
  Plan P;
  std::string e;
  if (!make_plan(expr, A, B, C, P, e)) { set_err(err, errlen, e); return -1; }
  const auto tp1 = std::chrono::steady_clock::now();
  if (plan_s) *plan_s = std::chrono::duration<double>(tp1 - tp0).count();
  if ((int64_t)P.cids.size() != C->nnz) { set_err(err, errlen, "C pattern not reserved"); return -1; }
  for (int64_t i = 0; i < C->nnz; ++i)
    if (C->indices[i] != P.cids[i]) { set_err(err, errlen, "C pattern mismatch"); return -1; }
  TView va(A), vb(B), vc(C);
  const double* ad = (const double*)A->data;
  const double* bd = (const double*)B->data;
  double* cd = (double*)C->data;
  const bool az = bytes_zero(alpha), bz = bytes_zero(beta);
  std::vector<int64_t> all;
  if (!subset) { all.resize((size_t)C->nnz); std::iota(all.begin(), all.end(), 0); subset = all.data(); nsubset = C->nnz; }
  std::atomic<int64_t> ntask{0};
  auto body = [&](int64_t s) {
    const int64_t pc = subset[s];
    const std::vector<int> dc = vc.dims(C->indices[pc]);
    int64_t vol = 1;
    for (int x : dc) vol *= x;
    double* c = cd + C->offsets[pc];
    const int64_t t0 = P.cstart[pc], t1 = P.cstart[pc + 1];
    ntask.fetch_add(t1 - t0);
    if (mode == 0) {
      for (int64_t x = 0; x < vol; ++x) c[x] = bz ? 0.0 : c[x] * beta;
      if (az) return;
      for (int64_t t = t0; t < t1; ++t) {
        const int64_t pa = P.ta[t], pb = P.tb[t];
        chain_task(ad + A->offsets[pa], va.dims(A->indices[pa]), bd + B->offsets[pb], vb.dims(B->indices[pb]),
                   c, dc, P.m, alpha);
      }
    } else {
      std::vector<long double> acc((size_t)vol);
      for (int64_t x = 0; x < vol; ++x) acc[x] = bz ? 0.0L : (long double)c[x] * beta;
      auto cr = scatter(P.m.crow, dc), cc = scatter(P.m.ccol, dc);
      for (int64_t t = t0; t < t1 && !az; ++t) {
        const int64_t pa = P.ta[t], pb = P.tb[t];
        const auto da = va.dims(A->indices[pa]), db = vb.dims(B->indices[pb]);
        auto ar = scatter(P.m.arow, da), ac = scatter(P.m.acol, da);
        auto br = scatter(P.m.brow, db), bc = scatter(P.m.bcol, db);
        const double* a = ad + A->offsets[pa];
        const double* b = bd + B->offsets[pb];
        for (size_t j = 0; j < bc.size(); ++j)
          for (size_t i = 0; i < ar.size(); ++i) {
            long double s2 = 0.0L;
            for (size_t k = 0; k < ac.size(); ++k) s2 += (long double)a[ar[i] + ac[k]] * (long double)b[br[k] + bc[j]];
            acc[(size_t)(cr[i] + cc[j])] += (long double)alpha * s2;
          }
      }
      for (int64_t x = 0; x < vol; ++x) c[x] = (double)acc[(size_t)x];
    }
  };
  pool_run(0, nsubset, nthreads, body);
  if (compute_s) *compute_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tp1).count();
  return ntask.load();
}

void orc_block_norms_f64(const orc_bst* T, int norm, double* out)
{
  TView v(T);
  const double* d = (const double*)T->data;
  for (int64_t p = 0; p < T->nnz; ++p) {
    auto dims = v.dims(T->indices[p]);
    int64_t vol = 1;
    for (int x : dims) vol *= x;
    const double* b = d + T->offsets[p];
    double r = 0.0;
    if (norm == 0) {
      for (int64_t i = 0; i < vol; ++i) r += b[i] * b[i];
      r = std::sqrt(r);
    } else {
      for (int64_t i = 0; i < vol; ++i) r = std::max(r, std::fabs(b[i]));
    }
    out[p] = r;
  }
}

// ---- generator --------------------------------------------------------------
uint64_t orc_splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static uint64_t tensor_key(uint64_t seed, int tid)
{
  return orc_splitmix64(seed + 0x632BE59BD9B4E019ULL * (uint64_t)(tid + 1));
}

int orc_gen_block_sizes(uint64_t seed, int axis, int length, const int* choices, int nchoices, int* out, int cap)
{
  const uint64_t key = orc_splitmix64(seed ^ (0xA0761D6478BD642FULL * (uint64_t)(axis + 1)));
  int n = 0, filled = 0;
  uint64_t ctr = 0;
  while (filled < length && n < cap) {
    const int b = choices[orc_splitmix64(key + ctr++) % (uint64_t)nchoices];
    const int take = std::min(b, length - filled);
    out[n++] = take;
    filled += take;
  }
  return n;
}

int64_t orc_gen_presence(uint64_t seed, int tid, int N, const int* nblk, double p, int64_t* out, int64_t cap)
{
  int64_t total = 1;
  for (int d = 0; d < N; ++d) total *= nblk[d];
  const uint64_t key = tensor_key(seed, tid);
  int64_t n = 0;
  for (int64_t id = 0; id < total; ++id) {
    const uint64_t h = orc_splitmix64(key ^ orc_splitmix64((uint64_t)id));
    const double u = (double)(h >> 11) * 0x1.0p-53;
    if (u < p) {
      if (n < cap) out[n] = id;
      ++n;
    }
  }
  return n;
}

void orc_gen_values_f64(uint64_t seed, int tid, const orc_bst* T, double lambda, const int64_t* positions,
                        int64_t npos)
{
  TView v(T);
  const uint64_t key = tensor_key(seed, tid);
  double* d = (double*)T->data;
  const int nth = (int)std::min<int64_t>(16, std::max<int64_t>(1, npos / 64));
  pool_run(0, npos, nth, [&](int64_t s) {
    std::vector<int> c;
    const int64_t p = positions ? positions[s] : s;
    const int64_t id = T->indices[p];
    v.coords(id, c);
    auto dims = v.dims(id);
    int64_t vol = 1;
    for (int x : dims) vol *= x;
    const double f = (lambda != 0.0 && v.N >= 2) ? std::exp(-lambda * (double)std::abs(c[0] - c[1])) : 1.0;
    const uint64_t bkey = orc_splitmix64(key ^ orc_splitmix64((uint64_t)id ^ 0x5851F42D4C957F2DULL));
    double* b = d + T->offsets[p];
    for (int64_t x = 0; x < vol; ++x) {
      const uint64_t h = orc_splitmix64(bkey + (uint64_t)x);
      const double u = (double)(h >> 11) * 0x1.0p-53;
      b[x] = (2.0 * u - 1.0) * f;
    }
  });
}

}  // extern "C"
