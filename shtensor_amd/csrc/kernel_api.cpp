// Kernel<T> replacement: expression analysis + hipRTC-specialised block
// product (one code object per (expr maps, M, N, K, dtype, alpha/beta class)).
//
// The reference JIT emits one x86 function per Kernel instance with asmjit
// (src/Kernel/KernelImpl.cpp:210-770) and keeps no cross-instance cache.  Here
// the per-shape HIP source is compiled with hipRTC for gfx950 and cached in
// process (by key) and on disk (code objects under $SHT_JIT_CACHE, default
// ~/.cache/shtensor_amd).  If hipRTC is unavailable or SHT_JIT=0, the generic
// table-driven kernel (blockprod.hip) runs instead; both compute the identical
// fma chain.

#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <unordered_map>

#include "internal.hpp"

struct sht_kernel {
  std::string expr;
  std::vector<int> sa, sb, sc;
  double alpha = 1.0, beta = 0.0;
  sht_dtype dtype = SHT_FLOAT64;
  sht_kernel_type type = SHT_KERNEL_XMM;
  sht::ContractMaps maps;
  std::vector<int64_t> srow[3], scol[3];  // scatter vectors of A, B, C
  int M = 0, N = 0, K = 0;
  bool alpha_zero = false, beta_zero = false;
  int* d_offs = nullptr;
  hipFunction_t jit_fn = nullptr;  // owned by the process-wide cache
  bool jit_mfma = false;           // the specialised kernel is the MFMA form
  int64_t jit_wpo = 1;             // its workgroups per op
  int jit_nw = 1;                  // and waves per workgroup
  // device staging for host-memory operands (grown on demand, freed with the kernel)
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  std::mutex stage_mu;
};

namespace sht {

namespace {

template <class T>
bool all_bytes_zero(T v)
{
  unsigned char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  for (unsigned char x : b)
    if (x) return false;
  return true;
}

bool leading_then_trailing(const std::vector<int>& lead, const std::vector<int>& trail)
{
  for (size_t r = 0; r < lead.size(); ++r)
    if (lead[r] != static_cast<int>(r)) return false;
  for (size_t r = 0; r < trail.size(); ++r)
    if (trail[r] != static_cast<int>(lead.size() + r)) return false;
  return true;
}

std::string join(const std::vector<int64_t>& v)
{
  std::string s;
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s;
}
std::string join(const std::vector<int>& v)
{
  std::string s;
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s;
}

// ---- hipRTC specialisation ------------------------------------------------------
struct JitEntry {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex g_jit_mu;
std::unordered_map<std::string, JitEntry> g_jit;

// Launch shape of a specialised kernel.  FMA form: one lane per C element
// of one op (grid-stride).  MFMA form (FP64, at least 16 rows and 4
// contracted indices): one wave per group of R 16-row tiles x one 16-column
// tile of one op; the waves of a workgroup (at most four) cover WR row
// groups x WC column tiles of the same op and share the WC column tiles' B
// panel through LDS.
struct JitShape {
  bool mfma = false;
  int R = 1;      // 16-row tiles per wave
  int NG = 1;     // row groups per op
  int NU = 1;     // 16-column tiles per op
  int WR = 1;     // waves of a workgroup along the row groups
  int WC = 1;     // waves of a workgroup along the column tiles
  int NW = 1;     // waves per workgroup (WR x WC)
  int WPO = 1;    // workgroups per op
  bool bl = false;  // the workgroup's B panel (16 WC columns) staged in LDS
};

constexpr int kKcWaves = 4;  // most waves per workgroup of the MFMA form

// Device staging of host operands: A, B, C of nb_ops ops, 256-byte aligned.
struct StageLayout {
  size_t a_bytes = 0, b_bytes = 0, c_bytes = 0;
  size_t b_off = 0, c_off = 0, total = 0;
};

JitShape jit_shape(const sht_kernel& k)
{
  JitShape js;
  const char* e = std::getenv("SHT_KCALL_FMA");
  const bool force_fma = e && std::strcmp(e, "1") == 0;
  if (k.dtype != SHT_FLOAT64 || k.M < 16 || k.K < 4 || force_fma) return js;
  js.mfma = true;
  js.R = std::min(4, (k.M + 15) / 16);
  js.NG = (k.M + 16 * js.R - 1) / (16 * js.R);
  js.NU = (k.N + 15) / 16;
  // Workgroups hold only waves with work: WR row groups x WC column tiles.
  // Blocks of one row group take two column tiles per workgroup (four made
  // the staged B panel too large for residency: 64 x 64 x 64 ops 17.7 TF
  // with two, 13.1 with four, 16.5 with one); taller blocks fill four waves
  // (128 x 128 x 128: 17.8 -> 27.4 TF against four row-group waves per
  // column tile).
  js.WR = std::min(kKcWaves, js.NG);
  js.WC = std::min(js.NG == 1 ? 2 : kKcWaves / js.WR, js.NU);
  js.NW = js.WR * js.WC;
  js.WPO = (js.NG + js.WR - 1) / js.WR * ((js.NU + js.WC - 1) / js.WC);
  js.bl = static_cast<int64_t>(k.K) * 16 * js.WC * 8 <= 32768;
  return js;
}

std::string jit_source(const sht_kernel& k, bool ca, bool cb, bool cc)
{
  const JitShape js = jit_shape(k);
  const char* T = k.dtype == SHT_FLOAT64 ? "double" : "float";
  std::ostringstream s;
  s << "typedef " << T << " T;\n"
    << "#define M " << k.M << "\n#define N " << k.N << "\n#define K " << k.K << "\n"
    << "#define AZ " << (k.alpha_zero ? 1 : 0) << "\n#define BZ " << (k.beta_zero ? 1 : 0) << "\n"
    << "#define CA " << (ca ? 1 : 0) << "\n#define CB " << (cb ? 1 : 0) << "\n#define CC " << (cc ? 1 : 0) << "\n";
  if (!js.mfma) {
    const char* FMA = k.dtype == SHT_FLOAT64 ? "__fma_rn" : "__fmaf_rn";
    s << "extern \"C\" __global__ __launch_bounds__(256) void sht_blockprod(const T* __restrict__ A,\n"
         "    const T* __restrict__ B, T* C, long long nb_ops, const int* __restrict__ offs, T alpha, T beta)\n"
         "{\n"
         "  const int* ar = offs; const int* ac = ar + M; const int* br = ac + K;\n"
         "  const int* bc = br + K; const int* cr = bc + N; const int* cc = cr + M;\n"
         "  const long long per = (long long)M * N;\n"
         "  const long long total = nb_ops * per;\n"
         "  const long long stride = (long long)gridDim.x * blockDim.x;\n"
         "  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {\n"
         "    const long long op = g / per;\n"
         "    const int r = (int)(g - op * per);\n"
         "    const int j = r / M, i = r - j * M;\n"
         "    const T* a = A + op * (long long)(M * K);\n"
         "    const T* b = B + op * (long long)(K * N);\n"
         "    T* c = C + op * per;\n"
         "    const int oc = CC ? (i + j * M) : (cr[i] + cc[j]);\n"
         "    T acc = BZ ? T(0) : c[oc] * beta;\n"
         "    if (!AZ) {\n"
         "      const int ai = CA ? i : ar[i];\n"
         "      const int bj = CB ? j * K : bc[j];\n"
         "#pragma unroll 8\n"
         "      for (int k = 0; k < K; ++k) {\n"
         "        const T av = a[ai + (CA ? k * M : ac[k])] * alpha;\n"
         "        acc = "
      << FMA
      << "(av, b[(CB ? k : br[k]) + bj], acc);\n"
         "      }\n"
         "    }\n"
         "    c[oc] = acc;\n"
         "  }\n"
         "}\n";
    return s.str();
  }
  // FP64 MFMA form.  v_mfma_f64_16x16x4f64 accumulates its four products as
  // a sequential fused multiply-add chain in k order, so C(i,j) =
  // fma(alpha*A(i,k), B(k,j), .) over k ascending -- the reference JIT's chain
  // (src/Kernel/KernelImpl.cpp:594-710) -- when the tile is computed
  // transposed (D = B^T x (alpha A)^T: products commute exactly): lane l then
  // holds rows i0 + l%16 of columns j0 + 4v + l/16, so C loads and stores
  // cover 16 consecutive rows per column.  The K % 4 tail runs as scalar fma
  // (no zero padding: a padded product could turn a -0 chain into +0).  Rows
  // past M and columns past N read clamped (valid) addresses and are not
  // stored.
  s << "#define R " << js.R << "\n#define NG " << js.NG << "\n#define NU " << js.NU << "\n#define WPO " << js.WPO
    << "\n#define WR " << js.WR << "\n#define WC " << js.WC << "\n#define BL " << (js.bl ? 1 : 0)
    << "\n#define NW " << js.NW << "\n"
    << R"SRC(typedef double v4d __attribute__((ext_vector_type(4)));
#define K4 (K / 4 * 4)
__device__ __forceinline__ int aoff(const int* ar, const int* ac, int i, int k) { return CA ? i + k * M : ar[i] + ac[k]; }
__device__ __forceinline__ int boff(const int* br, const int* bc, int k, int j) { return CB ? k + j * K : br[k] + bc[j]; }
__device__ __forceinline__ int coff(const int* cr, const int* cc, int i, int j) { return CC ? i + j * M : cr[i] + cc[j]; }
extern "C" __global__ __launch_bounds__(64 * NW) void sht_blockprod(const double* __restrict__ A,
    const double* __restrict__ B, double* C, long long nb_ops, const int* __restrict__ offs, double alpha, double beta)
{
  const int* ar = offs; const int* ac = ar + M; const int* br = ac + K;
  const int* bc = br + K; const int* cr = bc + N; const int* cc = cr + M;
#define NUB ((NU + WC - 1) / WC)
#define BW (16 * WC)
#if BL
  __shared__ double bs[K * BW];
#endif
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int wr = wave % WR, wc = wave / WR;  // (waves past WR x WC only stage B)
  const long long nwg = nb_ops * WPO;
  for (long long wg = blockIdx.x; wg < nwg; wg += gridDim.x) {
    const long long op = wg / WPO;
    const int rem = (int)(wg - op * WPO);
    const int ub = rem % NUB, g = (rem / NUB) * WR + wr;
    const int u = ub * WC + wc;
    const double* a = A + op * (long long)(M * K);
    const double* b = B + op * (long long)(K * N);
    double* c = C + op * (long long)(M * N);
    const int j0 = 16 * u;
#if BL
    __syncthreads();  // the previous item's reads of bs are done
    for (int x = threadIdx.x; x < K * BW; x += 64 * NW) {
      const int kk = x % K, jj = x / K;
      bs[kk * BW + jj] = b[boff(br, bc, kk, min(BW * ub + jj, N - 1))];
    }
    __syncthreads();
#endif
    if (wave >= WR * WC || g >= NG || u >= NU) continue;
    const int i0 = 16 * R * g;
    int ri[R];
#pragma unroll
    for (int t = 0; t < R; ++t) ri[t] = min(i0 + 16 * t + li, M - 1);
    int cj[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) cj[v] = min(j0 + 4 * v + lk, N - 1);
    v4d acc[R];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[t][v] = BZ ? 0.0 : c[coff(cr, cc, ri[t], cj[v])] * beta;
    if (!AZ) {
      const int jb = min(j0 + li, N - 1);
#pragma unroll 4
      for (int k0 = 0; k0 < K4; k0 += 4) {
#if BL
        const double bf = bs[(k0 + lk) * BW + 16 * wc + li];
#else
        const double bf = b[boff(br, bc, k0 + lk, jb)];
#endif
#pragma unroll
        for (int t = 0; t < R; ++t) {
          const double af = a[aoff(ar, ac, ri[t], k0 + lk)] * alpha;
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf, af, acc[t], 0, 0, 0);
        }
      }
#pragma unroll
      for (int k = K4; k < K; ++k)
#pragma unroll
        for (int t = 0; t < R; ++t) {
          const double av = a[aoff(ar, ac, ri[t], k)] * alpha;
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[t][v] = __fma_rn(av, b[boff(br, bc, k, cj[v])], acc[t][v]);
        }
    }
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (i0 + 16 * t + li < M && j0 + 4 * v + lk < N) c[coff(cr, cc, ri[t], cj[v])] = acc[t][v];
  }
}
)SRC";
  return s.str();
}

std::string cache_dir()
{
  const char* e = std::getenv("SHT_JIT_CACHE");
  if (e && *e) return e;
  const char* h = std::getenv("HOME");
  return std::string(h ? h : "/tmp") + "/.cache/shtensor_amd";
}

uint64_t fnv1a(const std::string& s)
{
  uint64_t h = 1469598103934665603ULL;
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

std::vector<char> compile_rtc(const std::string& src)
{
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "sht_blockprod.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw Error(SHT_ERR_HIP, "Failed to compile kernel: hiprtcCreateProgram");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, 7, opts);
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(SHT_ERR_HIP, std::string("Failed to compile kernel: ") + hiprtcGetErrorString(r) + "\n" + log);
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return code;
}

// Returns the code object for `src`, from disk cache or by compiling.
std::vector<char> code_object(const std::string& src)
{
  const std::string dir = cache_dir();
  char name[64];
  std::snprintf(name, sizeof(name), "/bp_%016llx.co", static_cast<unsigned long long>(fnv1a(src)));
  const std::string path = dir + name;
  {
    std::ifstream f(path, std::ios::binary);
    if (f) {
      std::vector<char> code((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
      if (!code.empty()) return code;
    }
  }
  std::vector<char> code = compile_rtc(src);
  // best effort: create the directory chain and write atomically
  std::string partial;
  for (size_t i = 0; i < dir.size(); ++i) {
    partial += dir[i];
    if (dir[i] == '/' || i + 1 == dir.size()) mkdir(partial.c_str(), 0755);
  }
  const std::string tmp = path + ".tmp" + std::to_string(static_cast<long long>(fnv1a(path) ^ reinterpret_cast<uintptr_t>(&code)));
  {
    std::ofstream f(tmp, std::ios::binary);
    if (f) f.write(code.data(), static_cast<std::streamsize>(code.size()));
  }
  std::rename(tmp.c_str(), path.c_str());
  return code;
}

bool jit_enabled()
{
  const char* e = std::getenv("SHT_JIT");
  return !(e && std::strcmp(e, "0") == 0);
}

}  // namespace

// Compile (or fetch) the shape-specialised kernel; nullptr when JIT disabled.
hipFunction_t jit_blockprod(const sht_kernel& k, bool load)
{
  if (!jit_enabled()) return nullptr;
  const bool ca = leading_then_trailing(k.maps.a_row, k.maps.a_col);
  const bool cb = leading_then_trailing(k.maps.b_row, k.maps.b_col);
  const bool cc = leading_then_trailing(k.maps.c_row, k.maps.c_col);
  const std::string src = jit_source(k, ca, cb, cc);
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto it = g_jit.find(src);
  if (it != g_jit.end()) return it->second.fn;
  std::vector<char> code = code_object(src);
  if (!load) return nullptr;
  JitEntry e;
  SHT_HIP(hipModuleLoadData(&e.module, code.data()));
  SHT_HIP(hipModuleGetFunction(&e.fn, e.module, "sht_blockprod"));
  g_jit[src] = e;
  return e.fn;
}

// Enqueue the block products of nb_ops packed ops (device pointers) on s.
void kernel_launch(const sht_kernel& k, const void* a, const void* b, void* c, int64_t nb_ops, hipStream_t s)
{
  if (k.jit_fn) {
    long long nops = nb_ops;
    unsigned g, bs;
    if (k.jit_mfma) {
      const int64_t nwg = nb_ops * k.jit_wpo;
      g = static_cast<unsigned>(nwg < 1048576 ? nwg : 1048576);
      bs = 64 * k.jit_nw;
    } else {
      const int64_t total = nb_ops * static_cast<int64_t>(k.M) * k.N;
      const int64_t blocks = (total + 255) / 256;
      g = static_cast<unsigned>(blocks < 262144 ? blocks : 262144);
      bs = 256;
    }
    if (k.dtype == SHT_FLOAT64) {
      double al = k.alpha, be = k.beta;
      void* args[] = {&a, &b, &c, &nops, const_cast<int**>(&k.d_offs), &al, &be};
      SHT_HIP(hipModuleLaunchKernel(k.jit_fn, g, 1, 1, bs, 1, 1, 0, s, args, nullptr));
    } else {
      float al = static_cast<float>(k.alpha), be = static_cast<float>(k.beta);
      void* args[] = {&a, &b, &c, &nops, const_cast<int**>(&k.d_offs), &al, &be};
      SHT_HIP(hipModuleLaunchKernel(k.jit_fn, g, 1, 1, bs, 1, 1, 0, s, args, nullptr));
    }
    return;
  }
  BlockProdArgs p{a, b, c, nb_ops, k.M, k.N, k.K, k.d_offs, k.alpha, k.beta, k.alpha_zero ? 1 : 0,
                  k.beta_zero ? 1 : 0};
  launch_blockprod(p, k.dtype == SHT_FLOAT64, s);
}

// True for memory the device does not own (pageable or pinned host memory).
bool is_host_ptr(const void* p)
{
  hipPointerAttribute_t at{};
  const hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory
    return true;
  }
  return at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged;
}

StageLayout stage_layout(const sht_kernel& k, int64_t nb_ops)
{
  const size_t es = k.dtype == SHT_FLOAT64 ? 8 : 4;
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  StageLayout L;
  L.a_bytes = static_cast<size_t>(nb_ops) * k.M * k.K * es;
  L.b_bytes = static_cast<size_t>(nb_ops) * k.K * k.N * es;
  L.c_bytes = static_cast<size_t>(nb_ops) * k.M * k.N * es;
  L.b_off = up(L.a_bytes);
  L.c_off = L.b_off + up(L.b_bytes);
  L.total = L.c_off + up(L.c_bytes);
  return L;
}

// One call with every host operand copied through `stage` (device memory of
// at least L.total bytes); returns after C is back in host memory.
void kernel_call_staged(const sht_kernel& k, const void* a, const void* b, void* c, int64_t nb_ops, hipStream_t s,
                        void* stage, const StageLayout& L)
{
  char* base = static_cast<char*>(stage);
  const bool ha = is_host_ptr(a), hb = is_host_ptr(b), hc = is_host_ptr(c);
  const void* da = ha ? base : a;
  const void* db = hb ? base + L.b_off : b;
  void* dc = hc ? base + L.c_off : c;
  if (ha) SHT_HIP(hipMemcpyAsync(base, a, L.a_bytes, hipMemcpyHostToDevice, s));
  if (hb) SHT_HIP(hipMemcpyAsync(base + L.b_off, b, L.b_bytes, hipMemcpyHostToDevice, s));
  // C is read only when beta != 0 (KernelImpl.cpp:478-479)
  if (hc && !k.beta_zero) SHT_HIP(hipMemcpyAsync(base + L.c_off, c, L.c_bytes, hipMemcpyHostToDevice, s));
  kernel_launch(k, da, db, dc, nb_ops, s);
  if (hc) SHT_HIP(hipMemcpyAsync(c, base + L.c_off, L.c_bytes, hipMemcpyDeviceToHost, s));
  SHT_HIP(hipStreamSynchronize(s));
}

// ---- allocation-free staging (sht_kernel_call_static) ---------------------
// A fixed device array stages host operands for callers that cannot allocate
// device memory.  Host code never passes the array's name to a kernel or a
// copy: the name is a host-side stand-in; each device's own address comes
// from hipGetSymbolAddress (looked up once per device, with that device
// current) and only that address is used.  Calls on one device share the
// array: calls that overlap in time on one device race.
constexpr size_t kStaticStageBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticStageBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

void* static_stage_ptr()
{
  int dev = 0;
  SHT_HIP(hipGetDevice(&dev));
  SHT_REQUIRE(dev >= 0 && dev < kMaxDevices, SHT_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    SHT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SHT_REQUIRE(ptr != nullptr, SHT_ERR_HIP, "hipGetSymbolAddress returned no address for the staging array");
    g_scratch_dev[dev] = ptr;
  }
  return g_scratch_dev[dev];
}

}  // namespace sht

using sht::Error;
using sht::guard;

extern "C" {

sht_status sht_kernel_create(const char* expr, const int* sizes_a, int na, const int* sizes_b, int nb,
                             const int* sizes_c, int nc, double alpha, double beta, sht_dtype dtype,
                             sht_kernel_type kernel_type, sht_kernel** out)
{
  return guard([&] {
    SHT_REQUIRE(out && expr, SHT_ERR_INVALID, "null argument");
    SHT_REQUIRE(na >= 0 && nb >= 0 && nc >= 0 && (sizes_a || !na) && (sizes_b || !nb) && (sizes_c || !nc),
                SHT_ERR_INVALID, "bad size arrays");
    // KernelImpl.cpp:42-71: (float type, kernel type) must be known
    SHT_REQUIRE((dtype == SHT_FLOAT32 || dtype == SHT_FLOAT64) &&
                    (kernel_type == SHT_KERNEL_LAPACK || kernel_type == SHT_KERNEL_XMM),
                SHT_ERR_INVALID, "Unknown kernel type");
    auto k = std::make_unique<sht_kernel>();
    k->expr = expr;
    k->sa.assign(sizes_a, sizes_a + na);
    k->sb.assign(sizes_b, sizes_b + nb);
    k->sc.assign(sizes_c, sizes_c + nc);
    k->dtype = dtype;
    k->type = kernel_type;
    k->maps = sht::analyse_expression(k->expr, k->sa, k->sb, k->sc);
    for (int s : k->sa) SHT_REQUIRE(s >= 1, SHT_ERR_INVALID, "sizes must be positive");
    for (int s : k->sb) SHT_REQUIRE(s >= 1, SHT_ERR_INVALID, "sizes must be positive");
    for (int s : k->sc) SHT_REQUIRE(s >= 1, SHT_ERR_INVALID, "sizes must be positive");
    if (dtype == SHT_FLOAT64) {
      k->alpha = alpha;
      k->beta = beta;
      k->alpha_zero = sht::all_bytes_zero(alpha);
      k->beta_zero = sht::all_bytes_zero(beta);
    } else {
      const float fa = static_cast<float>(alpha), fb = static_cast<float>(beta);
      k->alpha = fa;
      k->beta = fb;
      k->alpha_zero = sht::all_bytes_zero(fa);
      k->beta_zero = sht::all_bytes_zero(fb);
    }
    const sht::ContractMaps& m = k->maps;
    const std::vector<int>* sizes[3] = {&k->sa, &k->sb, &k->sc};
    const std::vector<int>* rows[3] = {&m.a_row, &m.b_row, &m.c_row};
    const std::vector<int>* cols[3] = {&m.a_col, &m.b_col, &m.c_col};
    for (int t = 0; t < 3; ++t) {
      k->srow[t] = sht::scatter_offsets(*rows[t], *sizes[t]);
      k->scol[t] = sht::scatter_offsets(*cols[t], *sizes[t]);
    }
    const int64_t M = static_cast<int64_t>(k->srow[0].size());
    const int64_t K = static_cast<int64_t>(k->scol[0].size());
    const int64_t N = static_cast<int64_t>(k->scol[1].size());
    // KernelImpl.cpp:408-426: scatter offsets are int32
    SHT_REQUIRE(M * K < (int64_t(1) << 31) && K * N < (int64_t(1) << 31) && M * N < (int64_t(1) << 31),
                SHT_ERR_INVALID, "Scatter index overflow");
    k->M = static_cast<int>(M);
    k->N = static_cast<int>(N);
    k->K = static_cast<int>(K);
    std::vector<int> offs;
    offs.reserve(2 * (M + N + K));
    for (int64_t v : k->srow[0]) offs.push_back(static_cast<int>(v));
    for (int64_t v : k->scol[0]) offs.push_back(static_cast<int>(v));
    for (int64_t v : k->srow[1]) offs.push_back(static_cast<int>(v));
    for (int64_t v : k->scol[1]) offs.push_back(static_cast<int>(v));
    for (int64_t v : k->srow[2]) offs.push_back(static_cast<int>(v));
    for (int64_t v : k->scol[2]) offs.push_back(static_cast<int>(v));
    SHT_HIP(hipMalloc(&k->d_offs, offs.size() * sizeof(int)));
    SHT_HIP(hipMemcpy(k->d_offs, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice));
    k->jit_fn = sht::jit_blockprod(*k, true);
    if (k->jit_fn) {
      const sht::JitShape js = sht::jit_shape(*k);
      k->jit_mfma = js.mfma;
      k->jit_wpo = js.WPO;
      k->jit_nw = js.NW;
    }
    *out = k.release();
  });
}

sht_status sht_kernel_call(sht_kernel* k, const void* a, const void* b, void* c, int64_t nb_ops, void* stream)
{
  return guard([&] {
    SHT_REQUIRE(k, SHT_ERR_INVALID, "null kernel");
    if (nb_ops <= 0) return;  // the reference's do-while runs one op for 0; here 0 is a no-op
    SHT_REQUIRE(a && b && c, SHT_ERR_INVALID, "null data pointer");
    hipStream_t s = sht::as_stream(stream);
    const bool ha = sht::is_host_ptr(a), hb = sht::is_host_ptr(b), hc = sht::is_host_ptr(c);
    if (!ha && !hb && !hc) {
      sht::kernel_launch(*k, a, b, c, nb_ops, s);
      return;
    }
    // Host operands (the reference's callers pass std::vector data,
    // test/test_kernel.cpp:143-150): staged through device buffers owned by
    // the kernel; the call then completes before it returns, as the
    // reference's does.
    std::lock_guard<std::mutex> lk(k->stage_mu);
    const sht::StageLayout L = sht::stage_layout(*k, nb_ops);
    if (k->stage_bytes < L.total) {
      if (k->d_stage) SHT_HIP(hipFree(k->d_stage));
      k->d_stage = nullptr;
      k->stage_bytes = 0;
      SHT_HIP(hipMalloc(&k->d_stage, L.total));
      k->stage_bytes = L.total;
    }
    sht::kernel_call_staged(*k, a, b, c, nb_ops, s, k->d_stage, L);
  });
}

sht_status sht_kernel_call_static(sht_kernel* k, const void* a, const void* b, void* c, int64_t nb_ops,
                                  void* stream)
{
  return guard([&] {
    SHT_REQUIRE(k, SHT_ERR_INVALID, "null kernel");
    if (nb_ops <= 0) return;
    SHT_REQUIRE(a && b && c, SHT_ERR_INVALID, "null data pointer");
    hipStream_t s = sht::as_stream(stream);
    const sht::StageLayout L = sht::stage_layout(*k, nb_ops);
    SHT_REQUIRE(L.total <= sht::kStaticStageBytes, SHT_ERR_INVALID,
                "sht_kernel_call_static: operands of " + std::to_string(nb_ops) + " ops need " +
                    std::to_string(L.total) + " bytes of staging, above the fixed " +
                    std::to_string(sht::kStaticStageBytes) + "-byte buffer");
    sht::kernel_call_staged(*k, a, b, c, nb_ops, s, sht::static_stage_ptr(), L);
  });
}

sht_status sht_kernel_info(const sht_kernel* k, char* buf, size_t len)
{
  return guard([&] {
    SHT_REQUIRE(k && buf && len, SHT_ERR_INVALID, "null argument");
    // layout of KernelImpl::get_info (KernelImpl.cpp:773-794)
    std::string out;
    out += "Kernel Info\n";
    out += "  Type: " + std::to_string(static_cast<int>(k->type)) + "\n";
    out += "  Expression: " + k->expr + "\n";
    const char* names[3] = {"Info Tensor In 1\n", "Info Tensor In 2\n", "Info Tensor Out\n"};
    const std::vector<int>* sizes[3] = {&k->sa, &k->sb, &k->sc};
    const std::vector<int>* rows[3] = {&k->maps.a_row, &k->maps.b_row, &k->maps.c_row};
    const std::vector<int>* cols[3] = {&k->maps.a_col, &k->maps.b_col, &k->maps.c_col};
    for (int t = 0; t < 3; ++t) {
      out += names[t];
      out += "  Sizes: " + sht::join(*sizes[t]) + "\n";
      out += "map_row: " + sht::join(*rows[t]) + "\n";
      out += "map_col: " + sht::join(*cols[t]) + "\n";
      out += "scatter row: " + sht::join(k->srow[t]) + "\n";
      out += "scatter col: " + sht::join(k->scol[t]) + "\n";
    }
    out += std::string("  Backend: HIP gfx950, ") +
           (k->jit_fn ? (k->jit_mfma ? "hipRTC-specialised (FP64 MFMA)" : "hipRTC-specialised (FMA)") : "generic") +
           "\n";
    std::snprintf(buf, len, "%s", out.c_str());
  });
}

sht_status sht_kernel_destroy(sht_kernel* k)
{
  return guard([&] {
    if (!k) return;
    if (k->d_offs) SHT_HIP(hipFree(k->d_offs));
    if (k->d_stage) SHT_HIP(hipFree(k->d_stage));
    delete k;
  });
}

// Compile the code object of a kernel shape into the disk cache without a
// GPU (used to pre-warm the cache at build time).  Returns SHT_OK when the
// object is in the cache afterwards.
sht_status sht_kernel_prewarm(const char* expr, const int* sizes_a, int na, const int* sizes_b, int nb,
                              const int* sizes_c, int nc, double alpha, double beta, sht_dtype dtype)
{
  return guard([&] {
    sht_kernel k;
    k.expr = expr;
    k.sa.assign(sizes_a, sizes_a + na);
    k.sb.assign(sizes_b, sizes_b + nb);
    k.sc.assign(sizes_c, sizes_c + nc);
    k.dtype = dtype;
    k.maps = sht::analyse_expression(k.expr, k.sa, k.sb, k.sc);
    k.M = static_cast<int>(sht::prod_of(k.sa, k.maps.a_row));
    k.K = static_cast<int>(sht::prod_of(k.sa, k.maps.a_col));
    k.N = static_cast<int>(sht::prod_of(k.sb, k.maps.b_col));
    if (dtype == SHT_FLOAT64) {
      k.alpha_zero = sht::all_bytes_zero(alpha);
      k.beta_zero = sht::all_bytes_zero(beta);
    } else {
      k.alpha_zero = sht::all_bytes_zero(static_cast<float>(alpha));
      k.beta_zero = sht::all_bytes_zero(static_cast<float>(beta));
    }
    sht::jit_blockprod(k, false);
  });
}

}  // extern "C"
