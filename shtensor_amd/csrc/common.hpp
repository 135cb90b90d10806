// Shared host utilities: error type, status conversion, HIP checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "sht.h"

namespace sht {

struct Error : std::runtime_error {
  sht_status code;
  Error(sht_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define SHT_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      throw ::sht::Error(SHT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define SHT_REQUIRE(cond, code, msg)                                                              \
  do {                                                                                            \
    if (!(cond)) throw ::sht::Error((code), (msg));                                               \
  } while (0)

template <class F>
sht_status guard(F&& f)
{
  try {
    f();
    return SHT_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return SHT_ERR_OOM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return SHT_ERR_INVALID;
  }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t prod_of(const std::vector<int>& v, const std::vector<int>& idx)
{
  int64_t p = 1;
  for (int i : idx) p *= v[i];
  return p;
}

// Expression analysis (product implementation; semantics of the reference's
// compute_contract_info, src/Basics/ContractInfo.h:68-346).
struct ContractMaps {
  std::vector<int> a_row, a_col, b_row, b_col, c_row, c_col;
  int na = 0, nb = 0, nc = 0;  // tensor orders
};

// Throws Error(SHT_ERR_EXPR, <reference message>) on rejection.
ContractMaps analyse_expression(const std::string& expr, const std::vector<int>& sizes_a,
                                const std::vector<int>& sizes_b, const std::vector<int>& sizes_c);

// create_scatter_vector (ContractInfo.h:38-66): element offset of each
// linearised row/col index, first listed axis fastest.
std::vector<int64_t> scatter_offsets(const std::vector<int>& map, const std::vector<int>& sizes);

}  // namespace sht
