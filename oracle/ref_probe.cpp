// Probe driver compiled against the reference's own headers where they lie
// (src/Basics/ContractInfo.h, Utils.h, Logger.h), used to write golden
// vectors for the oracle's expression analysis and reshape.  Built by
// oracle/Makefile into oracle/_ref/; never shipped, never run on the GPU.
//
// stdin, one request per line:
//   I|<expr>|<sizes A>|<sizes B>|<sizes C>     sizes comma separated
//   R|<sizes>|<order>                          reshape of iota
// stdout: one JSON object per line.

#include "ContractInfo.h"

#include <iostream>
#include <sstream>
#include <string>
#include <vector>

static std::vector<int> ints(const std::string& s)
{
  std::vector<int> v;
  std::stringstream ss(s);
  std::string t;
  while (std::getline(ss, t, ',')) if (!t.empty()) v.push_back(std::stoi(t));
  return v;
}

static std::string jarr(const std::vector<int>& v)
{
  std::string o = "[";
  for (size_t i = 0; i < v.size(); ++i) o += (i ? "," : "") + std::to_string(v[i]);
  return o + "]";
}

static std::string jstr(const std::string& s)
{
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o + "\"";
}

int main()
{
  std::string line;
  while (std::getline(std::cin, line)) {
    if (line.empty()) continue;
    std::vector<std::string> f;
    std::stringstream ss(line);
    std::string t;
    while (std::getline(ss, t, '|')) f.push_back(t);
    if (line.back() == '|') f.push_back("");
    if (f[0] == "I" && f.size() == 5) {
      auto da = ints(f[2]), db = ints(f[3]), dc = ints(f[4]);
      Shtensor::ContractInfo ia, ib, ic;
      std::string err;
      bool ok = false;
      try {
        ok = Shtensor::compute_contract_info(f[1], da, db, dc, ia, ib, ic, err);
      } catch (const std::exception& e) {
        err = std::string("exception: ") + e.what();
      }
      std::cout << "{\"expr\":" << jstr(f[1]) << ",\"sa\":" << jarr(da) << ",\"sb\":" << jarr(db)
                << ",\"sc\":" << jarr(dc) << ",\"ok\":" << (ok ? "true" : "false");
      if (ok) {
        std::cout << ",\"a_row\":" << jarr(ia.map_row) << ",\"a_col\":" << jarr(ia.map_col)
                  << ",\"b_row\":" << jarr(ib.map_row) << ",\"b_col\":" << jarr(ib.map_col)
                  << ",\"c_row\":" << jarr(ic.map_row) << ",\"c_col\":" << jarr(ic.map_col)
                  << ",\"a_srow\":" << jarr(ia.scatter_row) << ",\"a_scol\":" << jarr(ia.scatter_col)
                  << ",\"b_srow\":" << jarr(ib.scatter_row) << ",\"b_scol\":" << jarr(ib.scatter_col)
                  << ",\"c_srow\":" << jarr(ic.scatter_row) << ",\"c_scol\":" << jarr(ic.scatter_col);
      } else {
        std::cout << ",\"err\":" << jstr(err);
      }
      std::cout << "}\n";
    } else if (f[0] == "R" && f.size() == 3) {
      auto sizes = ints(f[1]), order = ints(f[2]);
      int64_t n = 1;
      for (int s : sizes) n *= s;
      std::vector<int> in(n), out(n);
      for (int64_t i = 0; i < n; ++i) in[i] = (int)i;
      Shtensor::Utils::reshape(in.data(), sizes, order, out.data());
      std::cout << "{\"sizes\":" << jarr(sizes) << ",\"order\":" << jarr(order) << ",\"out\":" << jarr(out)
                << "}\n";
    }
  }
  return 0;
}
