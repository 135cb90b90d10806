// Expression analysis of "A, B -> C" einsum strings.
//
// Behaviour of src/Basics/ContractInfo.h:68-346: letters and '_' are axis
// characters, whitespace is skipped, the grammar is exactly
// <indices> , <indices> -> <indices>; axes are numbered by first appearance
// scanning A, B, C; contracted (con) and free (ncon1 of A, ncon2 of B) sets are
// ordered by that number; maps hold the local axis positions:
// A=(ncon1 | con), B=(con | ncon2), C=(ncon1 | ncon2).  The reference's error
// messages are kept verbatim.  Two deliberate differences: an output index
// that neither input carries is rejected as "Dimension mismatch" (the
// reference only compares counts, ContractInfo.h:257), and the number of sizes
// must equal the number of indices of each tensor (unchecked there).

#include <cctype>
#include <map>

#include "common.hpp"

namespace sht {

namespace {

enum class Kind { Indices, Comma, Arrow };

struct Token {
  Kind kind;
  int pos;
  std::string text;
};

class Parser {
 public:
  explicit Parser(const std::string& s) : src_(s) {}

  std::vector<Token> tokens()
  {
    std::vector<Token> out;
    size_t p = 0;
    while (p < src_.size()) {
      const char ch = src_[p];
      if (std::isspace(static_cast<unsigned char>(ch))) { ++p; continue; }
      if (axis_char(ch)) {
        const size_t start = p;
        while (p < src_.size() && axis_char(src_[p])) ++p;
        out.push_back({Kind::Indices, (int)start, src_.substr(start, p - start)});
        continue;
      }
      if (ch == '-' && p + 1 < src_.size() && src_[p + 1] == '>') {
        out.push_back({Kind::Arrow, (int)p, "->"});
        p += 2;
        continue;
      }
      if (ch == ',') {
        out.push_back({Kind::Comma, (int)p, ","});
        ++p;
        continue;
      }
      throw Error(SHT_ERR_EXPR, "Unknown character in equation at position " + std::to_string(p) + ": " +
                                    std::string(1, ch));
    }
    return out;
  }

 private:
  static bool axis_char(char c) { return std::isalpha(static_cast<unsigned char>(c)) || c == '_'; }
  const std::string& src_;
};

const char* kind_name(Kind k)
{
  switch (k) {
    case Kind::Indices: return "index";
    case Kind::Comma: return "\",\"";
    default: return "->";
  }
}

std::vector<int> local_positions(const std::vector<int>& axes, const std::vector<int>& ordered_subset)
{
  std::vector<int> out;
  out.reserve(ordered_subset.size());
  for (int g : ordered_subset) {
    for (size_t i = 0; i < axes.size(); ++i) {
      if (axes[i] == g) { out.push_back((int)i); break; }
    }
  }
  return out;
}

void check_sizes(const std::vector<int>& m1, const std::vector<int>& m2, const std::vector<int>& d1,
                 const std::vector<int>& d2, const char* n1, const char* n2)
{
  for (size_t i = 0; i < m1.size(); ++i) {
    if (d1[m1[i]] != d2[m2[i]]) {
      throw Error(SHT_ERR_EXPR, "Sizes for dimensions " + std::to_string(m1[i]) + " in " + n1 +
                                    " and for dimension " + std::to_string(m2[i]) + " in " + n2 +
                                    " do not match.");
    }
  }
}

}  // namespace

ContractMaps analyse_expression(const std::string& expr, const std::vector<int>& sa, const std::vector<int>& sb,
                                const std::vector<int>& sc)
{
  if (expr.empty()) throw Error(SHT_ERR_EXPR, "Equation is empty");
  Parser parser(expr);
  const std::vector<Token> tok = parser.tokens();

  static const Kind grammar[5] = {Kind::Indices, Kind::Comma, Kind::Indices, Kind::Arrow, Kind::Indices};
  for (size_t i = 0; i < 5; ++i) {
    if (i >= tok.size())
      throw Error(SHT_ERR_EXPR, std::string("Unexpected end of expression, expected ") + kind_name(grammar[i]));
    if (tok[i].kind != grammar[i])
      throw Error(SHT_ERR_EXPR,
                  std::string("Expected ") + kind_name(grammar[i]) + " at position " + std::to_string(tok[i].pos));
  }
  if (tok.size() != 5) throw Error(SHT_ERR_EXPR, "Invalid syntax");

  std::map<char, int> axis_id;
  auto number = [&axis_id](const std::string& s) {
    std::vector<int> v;
    for (char ch : s) {
      auto it = axis_id.emplace(ch, (int)axis_id.size()).first;
      v.push_back(it->second);
    }
    return v;
  };
  const std::vector<int> xa = number(tok[0].text);
  const std::vector<int> xb = number(tok[2].text);
  const std::vector<int> xc = number(tok[4].text);
  const int naxes = (int)axis_id.size();

  std::vector<int> in_a(naxes, 0), in_b(naxes, 0), in_c(naxes, 0);
  bool dup = false;
  for (int x : xa) dup |= (in_a[x]++ > 0);
  for (int x : xb) dup |= (in_b[x]++ > 0);
  for (int x : xc) dup |= (in_c[x]++ > 0);
  if (dup) throw Error(SHT_ERR_EXPR, "Duplicate tensor indices");

  std::vector<int> con, ncon1, ncon2;  // ascending axis id
  for (int g = 0; g < naxes; ++g) {
    if (in_a[g] && in_b[g]) con.push_back(g);
    else if (in_a[g]) ncon1.push_back(g);
    else if (in_b[g]) ncon2.push_back(g);
  }
  if (ncon1.empty() || ncon2.empty()) throw Error(SHT_ERR_EXPR, "Element-wise multiplication not supported");
  if (con.empty()) throw Error(SHT_ERR_EXPR, "No indices found over which to contract");
  bool out_ok = (ncon1.size() + ncon2.size() == xc.size());
  for (int x : xc) out_ok = out_ok && !(in_a[x] && in_b[x]) && (in_a[x] || in_b[x]);
  if (!out_ok) throw Error(SHT_ERR_EXPR, "Dimension mismatch");
  if (sa.size() != xa.size())
    throw Error(SHT_ERR_EXPR, "Number of sizes does not match number of indices of tensor A");
  if (sb.size() != xb.size())
    throw Error(SHT_ERR_EXPR, "Number of sizes does not match number of indices of tensor B");
  if (sc.size() != xc.size())
    throw Error(SHT_ERR_EXPR, "Number of sizes does not match number of indices of tensor C");

  ContractMaps m;
  m.na = (int)xa.size();
  m.nb = (int)xb.size();
  m.nc = (int)xc.size();
  m.a_row = local_positions(xa, ncon1);
  m.a_col = local_positions(xa, con);
  m.b_row = local_positions(xb, con);
  m.b_col = local_positions(xb, ncon2);
  m.c_row = local_positions(xc, ncon1);
  m.c_col = local_positions(xc, ncon2);
  check_sizes(m.a_row, m.c_row, sa, sc, "tensor A", "tensor C");
  check_sizes(m.b_col, m.c_col, sb, sc, "tensor B", "tensor C");
  check_sizes(m.a_col, m.b_row, sa, sb, "tensor A", "tensor B");
  return m;
}

std::vector<int64_t> scatter_offsets(const std::vector<int>& map, const std::vector<int>& sizes)
{
  std::vector<int64_t> stride(sizes.size(), 1);
  for (size_t d = 1; d < sizes.size(); ++d) stride[d] = stride[d - 1] * sizes[d - 1];
  // odometer over the mapped axes, first listed fastest
  int64_t count = 1;
  for (int ax : map) count *= sizes[ax];
  std::vector<int64_t> out(static_cast<size_t>(count), 0);
  std::vector<int> ctr(map.size(), 0);
  int64_t off = 0;
  for (int64_t idx = 0; idx < count; ++idx) {
    out[static_cast<size_t>(idx)] = off;
    for (size_t r = 0; r < map.size(); ++r) {
      const int ax = map[r];
      if (++ctr[r] < sizes[ax]) { off += stride[ax]; break; }
      off -= (int64_t)(sizes[ax] - 1) * stride[ax];
      ctr[r] = 0;
    }
  }
  return out;
}

}  // namespace sht

extern "C" sht_status sht_contract_info(const char* expr, const int* sa, int na, const int* sb, int nb, const int* sc,
                                        int nc, int* maps_out)
{
  return sht::guard([&] {
    SHT_REQUIRE(expr && maps_out && na >= 0 && nb >= 0 && nc >= 0, SHT_ERR_INVALID, "null argument");
    const sht::ContractMaps m = sht::analyse_expression(expr, std::vector<int>(sa, sa + na),
                                                        std::vector<int>(sb, sb + nb), std::vector<int>(sc, sc + nc));
    const std::vector<int>* v[6] = {&m.a_row, &m.a_col, &m.b_row, &m.b_col, &m.c_row, &m.c_col};
    for (int r = 0; r < 6; ++r) {
      maps_out[r * 17] = static_cast<int>(v[r]->size());
      for (size_t i = 0; i < v[r]->size() && i < 16; ++i) maps_out[r * 17 + 1 + i] = (*v[r])[i];
    }
  });
}
