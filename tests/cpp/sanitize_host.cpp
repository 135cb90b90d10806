// Host-code check under AddressSanitizer + UndefinedBehaviorSanitizer or
// ThreadSanitizer (built by tests/cpp/Makefile.sanitize with g++, host code
// only): the oracle (thread-pool contraction, task list, kernel chain,
// reserve), the HBM arena's allocator logic over host memory from several
// threads at once, and the expression analysis including its error paths.
// Exit status = failures.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "arena.hpp"
#include "common.hpp"
#include "oracle.h"

namespace {

int g_fail = 0;

void report(bool ok, const char* what)
{
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  std::fflush(stdout);
  if (!ok) ++g_fail;
}

// ik,kj->ij over 12 x 9 x 11 blocks of sizes 3..8, p = 0.6, serial and with
// 8 pool threads: identical bits; task count equal to the brute-force count.
void oracle_threads()
{
  std::mt19937 gen(5);
  std::uniform_int_distribution<int> bsz(3, 8);
  std::bernoulli_distribution pres(0.6);
  std::uniform_real_distribution<double> val(-1.0, 1.0);
  const int ni = 12, nk = 9, nj = 11;
  std::vector<int> bi(ni), bk(nk), bj(nj);
  for (auto& x : bi) x = bsz(gen);
  for (auto& x : bk) x = bsz(gen);
  for (auto& x : bj) x = bsz(gen);
  auto build = [&](const std::vector<int>& b0, const std::vector<int>& b1, std::vector<int>& nblk,
                   std::vector<int>& flat, std::vector<int64_t>& ind, std::vector<int64_t>& off,
                   std::vector<double>& data) {
    nblk = {static_cast<int>(b0.size()), static_cast<int>(b1.size())};
    flat = b0;
    flat.insert(flat.end(), b1.begin(), b1.end());
    int64_t o = 0;
    for (size_t c1 = 0; c1 < b1.size(); ++c1)
      for (size_t c0 = 0; c0 < b0.size(); ++c0)
        if (pres(gen)) {
          ind.push_back(static_cast<int64_t>(c0 + b0.size() * c1));
          off.push_back(o);
          o += static_cast<int64_t>(b0[c0]) * b1[c1];
        }
    data.resize(o);
    for (auto& x : data) x = val(gen);
  };
  std::vector<int> na, fa, nb, fb;
  std::vector<int64_t> ia, oa, ib, ob;
  std::vector<double> da, db;
  build(bi, bk, na, fa, ia, oa, da);
  build(bk, bj, nb, fb, ib, ob, db);
  orc_bst A{2, na.data(), fa.data(), static_cast<int64_t>(ia.size()), ia.data(), oa.data(), da.data()};
  orc_bst B{2, nb.data(), fb.data(), static_cast<int64_t>(ib.size()), ib.data(), ob.data(), db.data()};
  std::vector<int> nc = {ni, nj}, fc = bi;
  fc.insert(fc.end(), bj.begin(), bj.end());
  orc_bst E{2, nc.data(), fc.data(), 0, nullptr, nullptr, nullptr};
  char err[256];
  int64_t ncid = 0;
  const int64_t nt = orc_tasks("ik,kj->ij", &A, &B, &E, nullptr, 0, nullptr, &ncid, nullptr, err, sizeof(err));
  std::vector<int64_t> tasks(3 * nt), cids(ncid);
  orc_tasks("ik,kj->ij", &A, &B, &E, tasks.data(), nt, cids.data(), &ncid, nullptr, err, sizeof(err));
  int64_t brute = 0;
  for (size_t x = 0; x < ia.size(); ++x)
    for (size_t y = 0; y < ib.size(); ++y) brute += (ia[x] / ni) == (ib[y] % nk);
  std::vector<int64_t> oc(ncid);
  int64_t o = 0;
  for (int64_t c = 0; c < ncid; ++c) {
    oc[c] = o;
    o += static_cast<int64_t>(bi[cids[c] % ni]) * bj[cids[c] / ni];
  }
  std::vector<double> c1(o, 0.0), c8(o, 0.0);
  orc_bst C1{2, nc.data(), fc.data(), ncid, cids.data(), oc.data(), c1.data()};
  orc_bst C8{2, nc.data(), fc.data(), ncid, cids.data(), oc.data(), c8.data()};
  orc_contract_f64("ik,kj->ij", 1.0, &A, &B, 0.0, &C1, 0, nullptr, 0, 1, err, sizeof(err));
  orc_contract_f64("ik,kj->ij", 1.0, &A, &B, 0.0, &C8, 0, nullptr, 0, 8, err, sizeof(err));
  report(nt == brute && nt > 0 && std::memcmp(c1.data(), c8.data(), o * sizeof(double)) == 0,
         "oracle: task list = brute force; 8-thread pool = serial, bit for bit");
}

// Arena over host memory: eight threads allocate, grow, shrink and free
// concurrently; every chunk keeps its own bytes; all memory returns.
void arena_threads()
{
  const size_t total = size_t(64) << 20;
  std::vector<char> backing(total + sht::Arena::kAlign);
  void* base = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(backing.data()) + sht::Arena::kAlign - 1) /
                                       sht::Arena::kAlign * sht::Arena::kAlign);
  sht::Arena arena;
  arena.init(base, total, true);
  std::atomic<int> bad{0};
  std::vector<std::thread> th;
  for (int t = 0; t < 8; ++t)
    th.emplace_back([&, t] {
      std::mt19937 gen(100 + t);
      std::uniform_int_distribution<size_t> sz(1, 200000);
      std::vector<std::pair<char*, size_t>> live;
      for (int it = 0; it < 400; ++it) {
        const int op = static_cast<int>(gen() % 3);
        if (op == 0 || live.empty()) {
          const size_t n = sz(gen);
          try {
            char* p = static_cast<char*>(arena.allocate(n));
            std::memset(p, t + 1, n);
            live.push_back({p, n});
          } catch (const sht::Error&) {  // out of arena: fine under contention
          }
        } else if (op == 1) {
          auto& e = live[gen() % live.size()];
          for (size_t x = 0; x < e.second; x += 997)
            if (e.first[x] != static_cast<char>(t + 1)) bad++;
          arena.release(e.first);
          e = live.back();
          live.pop_back();
        } else {
          auto& e = live[gen() % live.size()];
          const size_t n = sz(gen);
          try {
            char* q = static_cast<char*>(arena.reallocate(e.first, n));
            for (size_t x = 0; x < std::min(n, e.second); x += 997)
              if (q[x] != static_cast<char>(t + 1)) bad++;
            std::memset(q, t + 1, n);
            e = {q, n};
          } catch (const sht::Error&) {
          }
        }
      }
      for (auto& e : live) arena.release(e.first);
    });
  for (auto& x : th) x.join();
  report(bad.load() == 0 && arena.free_bytes() == arena.total() && arena.largest_free() == arena.total(),
         "arena: concurrent allocate/reallocate/release keep contents; all memory coalesces back");
}

// compute_contract_info semantics: valid maps and the reference's errors.
void contract_info()
{
  const sht::ContractMaps m = sht::analyse_expression("jik, kmj -> im", {6, 5, 8}, {8, 4, 6}, {5, 4});
  bool ok = m.a_row == std::vector<int>{1} && m.a_col == std::vector<int>{0, 2} && m.b_row == std::vector<int>{2, 0};
  const char* bad[] = {"ii,ik->ik", "ij,jk", "ij,jk->ikk", "ij,kl->ijkl", ""};
  for (const char* e : bad) {
    bool threw = false;
    try {
      sht::analyse_expression(e, {2, 2}, {2, 2}, {2, 2});
    } catch (const sht::Error&) {
      threw = true;
    }
    ok = ok && threw;
  }
  report(ok, "contract_info: maps of jik,kmj->im and five rejected expressions");
}

}  // namespace

namespace sht {
// (the library's thread-local error slot, defined in tensor.cpp, which this
// host-only build does not link)
void set_last_error(const std::string&) {}
}  // namespace sht

int main()
{
  oracle_threads();
  arena_threads();
  contract_info();
  std::printf("%d failure(s)\n", g_fail);
  return g_fail;
}
