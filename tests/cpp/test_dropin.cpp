// Drop-in check of include/shtensor_amd.hpp: the reference's own test
// scenarios (test/test_kernel.cpp:119-189 test_contract, :241-285
// test_kernel, :300-325 main; test/test_tensor.cpp:39-120) written against
// Shtensor::Kernel<T> / Tensor<T,N> with std::vector host data, plus a
// reserve -> contract flow.  Every result is checked bitwise against the
// oracle (liboracle.so, test infrastructure), not only by the reference's
// RMS comparison.  Prints one line per case; exit status = failures.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "oracle.h"
#include "shtensor_amd.hpp"

namespace {

int g_fail = 0;

void report(bool ok, const std::string& what)
{
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok) ++g_fail;
}

template <class T>
void fill_random(std::vector<T>& v, unsigned seed)
{
  std::mt19937 gen(seed);
  std::uniform_real_distribution<> dis(-1.0, 1.0);
  for (auto& x : v) x = static_cast<T>(dis(gen));
}

template <class T>
T rms(const std::vector<T>& v)
{
  double s = 0.0;
  for (T x : v) s += static_cast<double>(x) * x;
  return static_cast<T>(std::sqrt(s / static_cast<double>(v.size())));
}

template <class T>
bool same_bits(const std::vector<T>& a, const std::vector<T>& b)
{
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0;
}

template <class T>
void oracle_xmm(const std::string& expr, const std::vector<int>& sa, const std::vector<int>& sb,
                const std::vector<int>& sc, T alpha, T beta, const std::vector<T>& a, const std::vector<T>& b,
                std::vector<T>& c, int64_t nb)
{
  char err[256];
  int ok;
  if constexpr (std::is_same<T, double>::value)
    ok = orc_kernel_xmm_f64(expr.c_str(), sa.data(), (int)sa.size(), sb.data(), (int)sb.size(), sc.data(),
                            (int)sc.size(), alpha, beta, a.data(), b.data(), c.data(), nb, err, sizeof(err));
  else
    ok = orc_kernel_xmm_f32(expr.c_str(), sa.data(), (int)sa.size(), sb.data(), (int)sb.size(), sc.data(),
                            (int)sc.size(), alpha, beta, a.data(), b.data(), c.data(), nb, err, sizeof(err));
  if (!ok) std::printf("oracle: %s\n", err);
}

// test_kernel.cpp:119-189: LAPACK and XMM kernels on nb_cons packed ops,
// RMS within 1e-6 of each other (the reference's check), and here also
// bitwise equal to the oracle's XMM chain.
template <class T, size_t N, size_t M, size_t K>
void test_contract(const std::string& expr, const std::array<int, N>& sa, const std::array<int, M>& sb,
                   const std::array<int, K>& sc, int nb_cons)
{
  const int na = std::accumulate(sa.begin(), sa.end(), 1, std::multiplies<int>());
  const int nbe = std::accumulate(sb.begin(), sb.end(), 1, std::multiplies<int>());
  const int nc = std::accumulate(sc.begin(), sc.end(), 1, std::multiplies<int>());
  std::vector<T> a(nb_cons * na), b(nb_cons * nbe), c(nb_cons * nc);
  fill_random(a, 1);
  fill_random(b, 2);
  fill_random(c, 3);
  Shtensor::Kernel<T> lapack(expr, sa, sb, sc, T(1), T(0), Shtensor::KernelType::LAPACK);
  std::vector<T> rc = c;
  lapack.call(a.data(), b.data(), rc.data(), nb_cons);
  const T lapack_rms = rms(rc);
  Shtensor::Kernel<T> xmm(expr, sa, sb, sc, T(1), T(0), Shtensor::KernelType::XMM);
  std::vector<T> rx = c;
  xmm.call(a.data(), b.data(), rx.data(), nb_cons);
  const T xmm_rms = rms(rx);
  std::vector<T> want = c;
  oracle_xmm<T>(expr, {sa.begin(), sa.end()}, {sb.begin(), sb.end()}, {sc.begin(), sc.end()}, T(1), T(0), a, b,
                want, nb_cons);
  const bool ok = std::fabs(lapack_rms - xmm_rms) <= 1e-6 && same_bits(rx, want) && same_bits(rc, want) &&
                  xmm.get_info().find("Expression: " + expr) != std::string::npos;
  report(ok, "test_contract<" + std::string(sizeof(T) == 8 ? "double" : "float") + "> " + expr);
}

// test_kernel.cpp:241-285: alpha = 0, beta = 2 scales C (sum identity), here
// also elementwise.
template <class T, size_t N, size_t M, size_t K>
void test_kernel(const std::string& expr, const std::array<int, N>& sa, const std::array<int, M>& sb,
                 const std::array<int, K>& sc)
{
  const int na = std::accumulate(sa.begin(), sa.end(), 1, std::multiplies<int>());
  const int nbe = std::accumulate(sb.begin(), sb.end(), 1, std::multiplies<int>());
  const int nc = std::accumulate(sc.begin(), sc.end(), 1, std::multiplies<int>());
  std::vector<T> a(na), b(nbe), c(nc);
  fill_random(a, 4);
  fill_random(b, 5);
  fill_random(c, 6);
  const T sum_c = std::accumulate(c.begin(), c.end(), T());
  const T beta = 2.0;
  Shtensor::Kernel<T> kernel(expr, sa, sb, sc, T(), beta, Shtensor::KernelType::XMM);
  std::vector<T> out = c;
  kernel.call(a.data(), b.data(), out.data(), 1);
  const T sum_after = std::accumulate(out.begin(), out.end(), T());
  bool ok = std::fabs(beta * sum_c - sum_after) <= 1e-6 * std::max<T>(1, std::fabs(sum_after));
  for (int i = 0; i < nc; ++i) ok = ok && out[i] == beta * c[i];
  report(ok, "test_kernel<" + std::string(sizeof(T) == 8 ? "double" : "float") + "> " + expr);
}

// Kernel::call on device pointers (the arena) with the asynchronous form.
void test_device_pointers()
{
  const std::array<int, 3> sa{16, 16, 16}, sc{16, 16, 16};
  const std::array<int, 2> sb{16, 16};
  const int nb = 9;
  std::vector<double> a(nb * 4096), b(nb * 256), c(nb * 4096);
  fill_random(a, 7);
  fill_random(b, 8);
  fill_random(c, 9);
  Shtensor::Context ctx(64 << 20);
  void *da, *db, *dc;
  Shtensor::check(sht_arena_alloc(ctx.handle(), a.size() * 8, &da));
  Shtensor::check(sht_arena_alloc(ctx.handle(), b.size() * 8, &db));
  Shtensor::check(sht_arena_alloc(ctx.handle(), c.size() * 8, &dc));
  Shtensor::check(sht_copy_h2d(da, a.data(), a.size() * 8));
  Shtensor::check(sht_copy_h2d(db, b.data(), b.size() * 8));
  Shtensor::check(sht_copy_h2d(dc, c.data(), c.size() * 8));
  Shtensor::Kernel<double> k("ijk,kl->ijl", sa, sb, sc, 0.5, 1.25, Shtensor::KernelType::XMM);
  void* s = nullptr;
  Shtensor::check(sht_stream_create(&s));
  k.call_async(static_cast<double*>(da), static_cast<double*>(db), static_cast<double*>(dc), nb, s);
  Shtensor::check(sht_stream_synchronize(s));
  Shtensor::check(sht_stream_destroy(s));
  std::vector<double> out(c.size());
  Shtensor::check(sht_copy_d2h(out.data(), dc, out.size() * 8));
  std::vector<double> want = c;
  oracle_xmm<double>("ijk,kl->ijl", {16, 16, 16}, {16, 16}, {16, 16, 16}, 0.5, 1.25, a, b, want, nb);
  report(same_bits(out, want), "Kernel<double>::call_async on arena pointers");
  Shtensor::check(sht_arena_free(ctx.handle(), da));
  Shtensor::check(sht_arena_free(ctx.handle(), db));
  Shtensor::check(sht_arena_free(ctx.handle(), dc));
}

// test_tensor.cpp:39-120: reserve_all of a 4-mode tensor (360 blocks), block
// dims per position, filter of alternating 0/1 blocks.
void test_tensor()
{
  Shtensor::Context ctx(64 << 20);
  const std::vector<int> dim0 = {4, 5, 3, 4}, dim1 = {5, 5, 8, 9, 2}, dim2 = {5, 6, 3, 7, 7, 9}, dim3 = {4, 3, 2};
  Shtensor::VArray<4> bs4 = {dim0, dim1, dim2, dim3};
  Shtensor::Tensor<float, 4> t4(ctx, bs4);
  t4.reserve_all();
  bool ok = t4.get_nb_nzblocks_global() == 360 && t4.get_occupation() == 1.0;
  const auto info = t4.get_sparse_info();
  std::vector<float> data(t4.nelem());
  int64_t kept = 0;
  for (int64_t p = 0; p < info.nnz; ++p) {
    const auto blk = t4.get_local_block(p);
    const auto idx = t4.get_indices(blk.block_index);
    for (int d = 0; d < 4; ++d) ok = ok && blk.dim(d) == bs4[d][idx[d]];
    const float v = blk.block_index % 2 == 0 ? 0.f : 1.f;
    kept += blk.block_index % 2;
    std::fill(data.begin() + info.offsets[p], data.begin() + info.offsets[p] + blk.size(), v);
  }
  t4.upload(data);
  t4.filter(1e-6f, Shtensor::BlockNorm::FROBENIUS, true);
  ok = ok && t4.get_nb_nzblocks_local() == kept;
  const std::vector<float> after = t4.download();
  ok = ok && std::all_of(after.begin(), after.end(), [](float x) { return x == 1.f; });
  // BaseTensor::reserve from per-mode coordinate vectors (test_tensor.cpp:51-57)
  Shtensor::VArray<3> bs3 = {dim0, dim1, dim2};
  Shtensor::Tensor<double, 3> t3(ctx, bs3);
  Shtensor::VArray<3> idx = {std::vector<int>{0, 0, 0, 1, 2, 2, 3, 3, 3}, std::vector<int>{0, 0, 1, 0, 2, 3, 2, 3, 4},
                             std::vector<int>{0, 1, 1, 0, 1, 2, 0, 1, 2}};
  t3.reserve(idx);
  ok = ok && t3.get_nb_nzblocks_local() == 9;
  bool threw = false;
  try {
    Shtensor::VArray<3> bad = {std::vector<int>{0, 1}, std::vector<int>{0}, std::vector<int>{0, 1}};
    t3.reserve(bad);
  } catch (const std::runtime_error& e) {
    threw = std::string(e.what()).find("same size") != std::string::npos;
  }
  report(ok && threw, "Tensor<float,4> reserve_all/get_local_block/filter, Tensor<double,3> reserve");
}

orc_bst view(const int* nblk, const std::vector<int>& flat, int N, const int64_t* ind, const int64_t* off,
             int64_t nnz, double* data)
{
  orc_bst v{};
  v.N = N;
  v.nblk = nblk;
  v.bsizes = flat.data();
  v.nnz = nnz;
  v.indices = ind;
  v.offsets = off;
  v.data = data;
  return v;
}

template <int N>
std::vector<int> flat_sizes(const Shtensor::VArray<N>& bs, std::vector<int>& nblk)
{
  std::vector<int> f;
  nblk.clear();
  for (const auto& v : bs) {
    nblk.push_back(static_cast<int>(v.size()));
    f.insert(f.end(), v.begin(), v.end());
  }
  return f;
}

// reserve -> contract -> download, against the oracle's contraction (the fma
// chain continued over the task list): bitwise for FP64 and FP64_EXACT, C
// pattern identical.
void test_contract_flow(Shtensor::Precision prec, const char* name)
{
  Shtensor::Context ctx(256 << 20);
  const std::vector<int> bi = {16, 5, 21, 9}, bj = {7, 13, 28}, bk = {11, 16, 5, 9, 15}, bl = {28, 7, 21};
  Shtensor::VArray<3> bsa = {bi, bj, bk};
  Shtensor::VArray<2> bsb = {bk, bl};
  Shtensor::VArray<3> bsc = {bi, bj, bl};
  Shtensor::Tensor<double, 3> A(ctx, bsa);
  Shtensor::Tensor<double, 2> B(ctx, bsb);
  Shtensor::Tensor<double, 3> C(ctx, bsc);
  std::mt19937 gen(11);
  std::bernoulli_distribution pres(0.5);
  Shtensor::VArray<3> ia;
  for (int k = 0; k < 5; ++k)
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 4; ++i)
        if (pres(gen)) {
          ia[0].push_back(i);
          ia[1].push_back(j);
          ia[2].push_back(k);
        }
  Shtensor::VArray<2> ib;
  for (int l = 0; l < 3; ++l)
    for (int k = 0; k < 5; ++k)
      if (pres(gen)) {
        ib[0].push_back(k);
        ib[1].push_back(l);
      }
  A.reserve(ia);
  B.reserve(ib);
  std::vector<double> ha(A.nelem()), hb(B.nelem());
  fill_random(ha, 12);
  fill_random(hb, 13);
  A.upload(ha);
  B.upload(hb);
  Shtensor::contract(ctx, "ijk,kl->ijl", 1.0, A, B, 0.0, C, prec);
  const std::vector<double> got = C.download();
  // oracle: same pattern rule and the chain over tasks
  std::vector<int> na, nb, nc;
  const std::vector<int> fa = flat_sizes<3>(bsa, na), fb = flat_sizes<2>(bsb, nb), fc = flat_sizes<3>(bsc, nc);
  const auto sa = A.get_sparse_info();
  const auto sb = B.get_sparse_info();
  const auto sc = C.get_sparse_info();
  orc_bst oa = view(na.data(), fa, 3, sa.indices, sa.offsets, sa.nnz, ha.data());
  orc_bst ob = view(nb.data(), fb, 2, sb.indices, sb.offsets, sb.nnz, hb.data());
  orc_bst empty = view(nc.data(), fc, 3, nullptr, nullptr, 0, nullptr);
  char err[256];
  int64_t ncids = 0;
  const int64_t nt = orc_tasks("ijk,kl->ijl", &oa, &ob, &empty, nullptr, 0, nullptr, &ncids, nullptr, err, sizeof(err));
  std::vector<int64_t> tasks(3 * std::max<int64_t>(nt, 1)), cids(std::max<int64_t>(ncids, 1));
  orc_tasks("ijk,kl->ijl", &oa, &ob, &empty, tasks.data(), nt, cids.data(), &ncids, nullptr, err, sizeof(err));
  bool ok = ncids == sc.nnz && std::equal(cids.begin(), cids.begin() + ncids, sc.indices);
  std::vector<double> want(got.size(), 0.0);
  orc_bst occ = view(nc.data(), fc, 3, sc.indices, sc.offsets, sc.nnz, want.data());
  orc_contract_f64("ijk,kl->ijl", 1.0, &oa, &ob, 0.0, &occ, 0, nullptr, 0, 1, err, sizeof(err));
  ok = ok && same_bits(got, want) && nt > 0;
  report(ok, std::string("reserve -> contract -> download (") + name + "), " + std::to_string(nt) + " tasks");
}

// FP32 operands (Tensor<float,N>, the reference's T = float) for the FP32
// and split-BF16 forms: convert -> contract gives the same C bit for bit as
// the same form reading the double tensors the floats were rounded from.
void test_fp32_operands(Shtensor::Precision prec, const char* name)
{
  Shtensor::Context ctx(256 << 20);
  const std::vector<int> bi = {16, 5, 21, 9}, bj = {7, 13, 28}, bk = {11, 16, 5, 9, 15}, bl = {28, 7, 21};
  Shtensor::VArray<3> bsa = {bi, bj, bk};
  Shtensor::VArray<2> bsb = {bk, bl};
  Shtensor::VArray<3> bsc = {bi, bj, bl};
  Shtensor::Tensor<double, 3> A(ctx, bsa);
  Shtensor::Tensor<double, 2> B(ctx, bsb);
  A.reserve_all();
  B.reserve_all();
  std::vector<double> ha(A.nelem()), hb(B.nelem());
  fill_random(ha, 21);
  fill_random(hb, 22);
  A.upload(ha);
  B.upload(hb);
  Shtensor::Tensor<float, 3> A32(ctx, bsa);
  Shtensor::Tensor<float, 2> B32(ctx, bsb);
  Shtensor::convert(A, A32);
  Shtensor::convert(B, B32);
  const std::vector<float> a32 = A32.download();
  bool ok = a32.size() == ha.size();
  for (size_t i = 0; ok && i < ha.size(); ++i) ok = a32[i] == static_cast<float>(ha[i]);
  Shtensor::Tensor<double, 3> C64(ctx, bsc), C32(ctx, bsc);
  Shtensor::contract(ctx, "ijk,kl->ijl", 1.0, A, B, 0.0, C64, prec);
  Shtensor::contract(ctx, "ijk,kl->ijl", 1.0, A32, B32, 0.0, C32, prec);
  ok = ok && same_bits(C32.download(), C64.download());
  report(ok, std::string("convert -> FP32 operands -> contract (") + name + ") == double operands");
}

}  // namespace

int main()
{
  try {
    // test/test_kernel.cpp:300-311
    test_contract<float>("ik, kj -> ij", std::array<int, 2>{8, 5}, std::array<int, 2>{5, 6}, std::array<int, 2>{8, 6}, 20);
    test_contract<double>("ik, kj -> ij", std::array<int, 2>{8, 5}, std::array<int, 2>{5, 6}, std::array<int, 2>{8, 6}, 20);
    test_contract<float>("jik, kmj -> im", std::array<int, 3>{6, 5, 8}, std::array<int, 3>{8, 4, 6},
                         std::array<int, 2>{5, 4}, 20);
    test_contract<double>("jik, kmj -> im", std::array<int, 3>{6, 5, 8}, std::array<int, 3>{8, 4, 6},
                          std::array<int, 2>{5, 4}, 20);
    test_contract<float>("ikl, lkmj -> mij", std::array<int, 3>{6, 5, 7}, std::array<int, 4>{7, 5, 6, 6},
                         std::array<int, 3>{6, 6, 6}, 20);
    test_contract<double>("ikl, lkmj -> mij", std::array<int, 3>{6, 5, 7}, std::array<int, 4>{7, 5, 6, 6},
                          std::array<int, 3>{6, 6, 6}, 20);
    // (an MFMA-sized shape through the same host-data path)
    test_contract<double>("ijk,kl->ijl", std::array<int, 3>{16, 16, 16}, std::array<int, 2>{16, 16},
                          std::array<int, 3>{16, 16, 16}, 20);
    // test/test_kernel.cpp:313-334
    test_kernel<float>("ijml, lmk -> jki", std::array<int, 4>{5, 6, 8, 9}, std::array<int, 3>{9, 8, 7},
                       std::array<int, 3>{6, 7, 5});
    test_kernel<float>("ij, jk -> ik", std::array<int, 2>{2, 2}, std::array<int, 2>{2, 2}, std::array<int, 2>{2, 2});
    test_kernel<float>("ij, jk -> ik", std::array<int, 2>{5, 5}, std::array<int, 2>{5, 5}, std::array<int, 2>{5, 5});
    test_kernel<float>("ij, jk -> ik", std::array<int, 2>{8, 8}, std::array<int, 2>{8, 8}, std::array<int, 2>{8, 8});
    test_kernel<float>("ij, jk -> ik", std::array<int, 2>{10, 10}, std::array<int, 2>{10, 10},
                       std::array<int, 2>{10, 10});
    test_kernel<double>("ij, jk -> ik", std::array<int, 2>{3, 3}, std::array<int, 2>{3, 3}, std::array<int, 2>{3, 3});
    test_kernel<double>("ij, jk -> ik", std::array<int, 2>{4, 4}, std::array<int, 2>{4, 4}, std::array<int, 2>{4, 4});
    test_kernel<double>("ij, jk -> ki", std::array<int, 2>{4, 4}, std::array<int, 2>{4, 4}, std::array<int, 2>{4, 4});
    test_device_pointers();
    test_tensor();
    test_contract_flow(Shtensor::Precision::FP64, "FP64");
    test_contract_flow(Shtensor::Precision::FP64_EXACT, "FP64_EXACT");
    test_fp32_operands(Shtensor::Precision::FP32, "FP32");
    test_fp32_operands(Shtensor::Precision::BF16X3, "BF16X3");
    // an expression error surfaces as std::runtime_error with the reference's text
    bool threw = false;
    try {
      Shtensor::Kernel<double> bad("ii,ik->ik", std::array<int, 2>{2, 2}, std::array<int, 2>{2, 2},
                                   std::array<int, 2>{2, 2}, 1.0, 0.0, Shtensor::KernelType::XMM);
    } catch (const std::runtime_error& e) {
      threw = std::string(e.what()).find("Duplicate tensor indices") != std::string::npos;
    }
    report(threw, "expression error as std::runtime_error");
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1 + g_fail;
  }
  std::printf("%d failure(s)\n", g_fail);
  return g_fail;
}
