// Minimal JSON reader for CmdStan data, init and metric files
// (the formats read by src/cmdstan/command_helper.hpp:116-245 through
// stan::json::json_data): numbers, nested rectangular arrays, objects, and the
// strings "inf", "+inf", "-inf", "Infinity", "-Infinity", "NaN" for reals.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nutsgpu {

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  double num = 0;
  bool is_int = false;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
};

class JsonParser {
 public:
  explicit JsonParser(const std::string& t) : s(t), i(0) {}
  JVal parse() {
    JVal v = value();
    ws();
    if (i != s.size()) err("trailing characters");
    return v;
  }

 private:
  const std::string& s;
  size_t i;
  [[noreturn]] void err(const std::string& m) {
    throw std::invalid_argument("JSON parse error at offset " + std::to_string(i) + ": " + m);
  }
  void ws() { while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i; }
  JVal value() {
    ws();
    if (i >= s.size()) err("unexpected end");
    const char c = s[i];
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') { JVal v; v.kind = JVal::STR; v.str = string(); return v; }
    if (s.compare(i, 4, "true") == 0) { i += 4; JVal v; v.kind = JVal::BOOL; v.num = 1; return v; }
    if (s.compare(i, 5, "false") == 0) { i += 5; JVal v; v.kind = JVal::BOOL; v.num = 0; return v; }
    if (s.compare(i, 4, "null") == 0) { i += 4; return JVal(); }
    return number();
  }
  std::string string() {
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\') {
        ++i;
        if (i >= s.size()) err("bad escape");
        const char e = s[i];
        if (e == 'n') out += '\n'; else if (e == 't') out += '\t'; else if (e == 'r') out += '\r';
        else if (e == 'b') out += '\b'; else if (e == 'f') out += '\f';
        else if (e == 'u') { i += 4; out += '?'; }
        else out += e;
        ++i;
      } else {
        out += s[i++];
      }
    }
    if (i >= s.size()) err("unterminated string");
    ++i;
    return out;
  }
  JVal number() {
    const size_t b = i;
    bool integral = true;
    if (s[i] == '-' || s[i] == '+') ++i;
    while (i < s.size() && ((s[i] >= '0' && s[i] <= '9') || s[i] == '.' || s[i] == 'e' || s[i] == 'E' ||
                            s[i] == '-' || s[i] == '+')) {
      if (s[i] == '.' || s[i] == 'e' || s[i] == 'E') integral = false;
      ++i;
    }
    if (i == b) err("unexpected character");
    JVal v;
    v.kind = JVal::NUM;
    v.num = std::strtod(s.substr(b, i - b).c_str(), nullptr);
    v.is_int = integral;
    return v;
  }
  JVal array() {
    ++i;
    JVal v;
    v.kind = JVal::ARR;
    ws();
    if (i < s.size() && s[i] == ']') { ++i; return v; }
    for (;;) {
      v.arr.push_back(value());
      ws();
      if (i < s.size() && s[i] == ',') { ++i; continue; }
      if (i < s.size() && s[i] == ']') { ++i; return v; }
      err("expected , or ]");
    }
  }
  JVal object() {
    ++i;
    JVal v;
    v.kind = JVal::OBJ;
    ws();
    if (i < s.size() && s[i] == '}') { ++i; return v; }
    for (;;) {
      ws();
      if (i >= s.size() || s[i] != '"') err("expected key");
      std::string k = string();
      ws();
      if (i >= s.size() || s[i] != ':') err("expected :");
      ++i;
      v.obj.emplace_back(k, value());
      ws();
      if (i < s.size() && s[i] == ',') { ++i; continue; }
      if (i < s.size() && s[i] == '}') { ++i; return v; }
      err("expected , or }");
    }
  }
};

// Flattens a (nested, rectangular) JSON value into dims + row-major values.
inline void json_flatten(const JVal& v, std::vector<int64_t>& dims, std::vector<double>& vals, bool& all_int,
                         const std::string& name) {
  auto scalar = [&](const JVal& x) -> double {
    if (x.kind == JVal::NUM) { if (!x.is_int) all_int = false; return x.num; }
    if (x.kind == JVal::BOOL) return x.num;
    if (x.kind == JVal::STR) {
      all_int = false;
      const std::string& t = x.str;
      if (t == "inf" || t == "+inf" || t == "Inf" || t == "Infinity" || t == "+Infinity") return INFINITY;
      if (t == "-inf" || t == "-Inf" || t == "-Infinity") return -INFINITY;
      if (t == "NaN" || t == "nan") return NAN;
    }
    throw std::invalid_argument("variable " + name + ": unsupported JSON value");
  };
  // dims from the first element at each depth
  dims.clear();
  const JVal* p = &v;
  while (p->kind == JVal::ARR) {
    dims.push_back((int64_t)p->arr.size());
    if (p->arr.empty()) break;
    p = &p->arr[0];
  }
  std::vector<const JVal*> stack{&v};
  // depth-first, row-major
  struct Rec {
    static void go(const JVal& x, size_t depth, const std::vector<int64_t>& dims, std::vector<double>& out,
                   const std::function<double(const JVal&)>& sc, const std::string& nm) {
      if (depth == dims.size()) { out.push_back(sc(x)); return; }
      if (x.kind != JVal::ARR || (int64_t)x.arr.size() != dims[depth])
        throw std::invalid_argument("variable " + nm + ": ragged array");
      for (const auto& e : x.arr) go(e, depth + 1, dims, out, sc, nm);
    }
  };
  vals.clear();
  if (!dims.empty() && dims.back() == 0) return;
  Rec::go(v, 0, dims, vals, scalar, name);
}

}  // namespace nutsgpu
