// Stan CSV files as written by the sampler (one file per chain): "# key = value"
// comment lines, the column header, draw rows and the elapsed-time footer.
// Shared by stansummary and diagnose.
#pragma once
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace stancsv {

struct ChainFile {
  std::vector<std::string> names;
  std::vector<std::vector<double>> rows;
  std::map<std::string, std::string> meta;  // "key = value" comment lines
  double warmup_s = 0, sampling_s = 0;
};

inline std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

inline std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  std::string cur;
  std::istringstream is(s);
  while (std::getline(is, cur, sep)) out.push_back(cur);
  return out;
}

inline ChainFile read_csv(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("Can't open file " + path);
  ChainFile f;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    if (line[0] == '#') {
      const std::string body = trim(line.substr(1));
      const size_t eq = body.find('=');
      if (eq != std::string::npos) {
        std::string k = trim(body.substr(0, eq)), v = trim(body.substr(eq + 1));
        const size_t sp = v.find(' ');
        if (sp != std::string::npos && v.find("(Default)") != std::string::npos) v = v.substr(0, sp);
        if (!f.meta.count(k)) f.meta[k] = v;
      }
      // "#  Elapsed Time: 0.005 seconds (Warm-up)"
      const size_t et = body.find("Elapsed Time:");
      const std::string tb = et != std::string::npos ? body.substr(et + 13) : body;
      if (tb.find("(Warm-up)") != std::string::npos) f.warmup_s = std::atof(trim(tb).c_str());
      else if (tb.find("(Sampling)") != std::string::npos) f.sampling_s = std::atof(trim(tb).c_str());
      continue;
    }
    if (f.names.empty()) {
      f.names = split(trim(line), ',');
      continue;
    }
    const auto cells = split(trim(line), ',');
    if (cells.size() != f.names.size()) continue;
    std::vector<double> r(cells.size());
    for (size_t i = 0; i < cells.size(); ++i) r[i] = std::strtod(cells[i].c_str(), nullptr);
    f.rows.push_back(std::move(r));
  }
  if (f.names.empty()) throw std::runtime_error("no header in " + path);
  // saved warmup rows come first: keep the sampling draws only
  auto it = f.meta.find("save_warmup");
  if (it != f.meta.end() && (it->second == "1" || it->second == "true")) {
    const long nw = f.meta.count("num_warmup") ? std::atol(f.meta["num_warmup"].c_str()) : 0;
    const long th = f.meta.count("thin") ? std::max(1L, std::atol(f.meta["thin"].c_str())) : 1;
    const size_t warm = (size_t)((nw + th - 1) / th);
    f.rows.erase(f.rows.begin(), f.rows.begin() + std::min(warm, f.rows.size()));
  }
  return f;
}

// a.1.2 -> a[1,2]
inline std::string display_name(const std::string& n) {
  const size_t dot = n.find('.');
  if (dot == std::string::npos) return n;
  std::string r = n.substr(0, dot) + "[";
  std::string rest = n.substr(dot + 1);
  for (char& ch : rest) if (ch == '.') ch = ',';
  return r + rest + "]";
}

}  // namespace stancsv
