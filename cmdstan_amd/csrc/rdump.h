// Reader for R dump text, the second data format CmdStan accepts for data,
// init, metric and log_prob parameter files (src/cmdstan/command_helper.hpp:
// 116-133 get_var_context: a file that does not end in .json is read as R dump
// by stan::io::dump).  Accepted forms, as R's dump() writes them:
//   name <- 3            name <- 2L            name <- -Inf / Inf / NaN / NA
//   name <- c(1, 2.5, -3e4)                    name <- 1:10     name <- integer(0)
//   name <- structure(c(...) or a:b, .Dim = c(2L, 3L))
// A name may be quoted ("name") and `=` may replace `<-`.  R stores arrays in
// column-major order; the values are returned in row-major order, the layout
// nuts_data uses for every array.
#pragma once
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace nutsgpu {

struct RDumpVar {
  std::string name;
  std::vector<int64_t> dims;  // empty: scalar
  std::vector<double> vals;   // row-major
  bool is_int = true;
};

class RDumpParser {
 public:
  explicit RDumpParser(const std::string& t) : s(t), i(0) {}

  std::vector<RDumpVar> parse() {
    std::vector<RDumpVar> out;
    for (;;) {
      ws();
      if (i >= s.size()) break;
      RDumpVar v;
      v.name = name();
      ws();
      if (s.compare(i, 2, "<-") == 0) i += 2;
      else if (i < s.size() && s[i] == '=') ++i;
      else err("expected '<-' after " + v.name);
      value(v);
      out.push_back(std::move(v));
      ws();
      if (i < s.size() && s[i] == ';') ++i;
    }
    return out;
  }

 private:
  const std::string& s;
  size_t i;

  [[noreturn]] void err(const std::string& m) const {
    throw std::invalid_argument("R dump parse error at offset " + std::to_string(i) + ": " + m);
  }
  void ws() {
    while (i < s.size()) {
      const char c = s[i];
      if (c == ' ' || c == '\n' || c == '\t' || c == '\r') { ++i; continue; }
      if (c == '#') { while (i < s.size() && s[i] != '\n') ++i; continue; }
      break;
    }
  }
  bool eat(char c) {
    ws();
    if (i < s.size() && s[i] == c) { ++i; return true; }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) err(std::string("expected '") + c + "'");
  }
  bool word(const char* w) {
    ws();
    const size_t n = std::char_traits<char>::length(w);
    if (s.compare(i, n, w) != 0) return false;
    const size_t j = i + n;
    if (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_' || s[j] == '.')) return false;
    i = j;
    return true;
  }
  std::string name() {
    ws();
    std::string n;
    if (i < s.size() && (s[i] == '"' || s[i] == '\'' || s[i] == '`')) {
      const char q = s[i++];
      while (i < s.size() && s[i] != q) n += s[i++];
      if (i >= s.size()) err("unterminated name");
      ++i;
    } else {
      while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '.')) n += s[i++];
    }
    if (n.empty()) err("expected a variable name");
    return n;
  }
  // one number; is_int stays true only for integer literals (digits, optional L)
  double number(bool& is_int) {
    ws();
    bool neg = false;
    if (i < s.size() && (s[i] == '-' || s[i] == '+')) { neg = s[i] == '-'; ++i; ws(); }
    if (word("Inf")) { is_int = false; return neg ? -INFINITY : INFINITY; }
    if (word("NaN") || word("NA")) { is_int = false; return NAN; }
    const size_t b = i;
    bool frac = false;
    while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.' || s[i] == 'e' || s[i] == 'E' ||
                            ((s[i] == '-' || s[i] == '+') && i > b && (s[i - 1] == 'e' || s[i - 1] == 'E')))) {
      if (!std::isdigit((unsigned char)s[i])) frac = true;
      ++i;
    }
    if (i == b) err("expected a number");
    const std::string tok = s.substr(b, i - b);
    char* e = nullptr;
    const double x = std::strtod(tok.c_str(), &e);
    if (*e) err("bad number '" + tok + "'");
    if (i < s.size() && s[i] == 'L') {
      if (frac && x != std::floor(x)) err("non-integer with an L suffix");
      ++i;
      frac = false;
    }
    if (frac) is_int = false;
    return neg ? -x : x;
  }
  // a number, or a:b (integer sequence, either direction)
  void element(std::vector<double>& out, bool& is_int) {
    bool a_int = true;
    const double a = number(a_int);
    ws();
    if (i < s.size() && s[i] == ':') {
      ++i;
      bool b_int = true;
      const double b = number(b_int);
      if (a != std::floor(a) || b != std::floor(b)) err("sequence bounds must be integers");
      if (a <= b) for (double x = a; x <= b; x += 1) out.push_back(x);
      else for (double x = a; x >= b; x -= 1) out.push_back(x);
      return;
    }
    if (!a_int) is_int = false;
    out.push_back(a);
  }
  // c(...), a:b, integer(n), double(n), numeric(n) or one number
  void vector_value(std::vector<double>& out, bool& is_int, bool& is_vector) {
    ws();
    is_vector = true;
    if (word("c")) {
      expect('(');
      if (eat(')')) return;
      do { element(out, is_int); } while (eat(','));
      expect(')');
      return;
    }
    for (const char* f : {"integer", "double", "numeric"}) {
      if (word(f)) {
        expect('(');
        bool n_int = true;
        const double n = number(n_int);
        expect(')');
        if (n < 0 || n != std::floor(n)) err("bad length");
        if (f[0] != 'i') is_int = false;
        out.assign((size_t)n, 0.0);
        return;
      }
    }
    const size_t before = out.size();
    element(out, is_int);
    is_vector = out.size() - before != 1;  // a:b is a vector, a single number a scalar
  }
  void value(RDumpVar& v) {
    ws();
    if (word("structure")) {
      expect('(');
      bool vec = true;
      vector_value(v.vals, v.is_int, vec);
      expect(',');
      ws();
      if (!(word(".Dim") || word(".dim"))) err("expected .Dim in structure()");
      expect('=');
      std::vector<double> d;
      bool d_int = true, d_vec = true;
      vector_value(d, d_int, d_vec);
      expect(')');
      int64_t n = 1;
      for (double x : d) {
        if (x < 0 || x != std::floor(x)) err("bad .Dim");
        v.dims.push_back((int64_t)x);
        n *= (int64_t)x;
      }
      if (n != (int64_t)v.vals.size())
        err("variable " + v.name + ": .Dim has " + std::to_string(n) + " elements, the data " +
            std::to_string(v.vals.size()));
      v.vals = column_to_row_major(v.vals, v.dims);
      return;
    }
    bool vec = true;
    vector_value(v.vals, v.is_int, vec);
    if (vec) v.dims = {(int64_t)v.vals.size()};
  }
  static std::vector<double> column_to_row_major(const std::vector<double>& cm, const std::vector<int64_t>& dims) {
    const size_t nd = dims.size();
    if (nd < 2) return cm;
    std::vector<double> rm(cm.size());
    std::vector<int64_t> idx(nd, 0);
    for (size_t r = 0; r < rm.size(); ++r) {
      // idx is the row-major multi-index of r; its column-major offset
      int64_t off = 0, stride = 1;
      for (size_t k = 0; k < nd; ++k) { off += idx[k] * stride; stride *= dims[k]; }
      rm[r] = cm[(size_t)off];
      for (size_t k = nd; k-- > 0;) {
        if (++idx[k] < dims[k]) break;
        idx[k] = 0;
      }
    }
    return rm;
  }
};

}  // namespace nutsgpu
