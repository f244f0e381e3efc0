// stansummary over Stan CSV files written by the sampler (one file per chain):
// the statistics of stansummary_helper.hpp:397-428 (get_stats) through
// nuts_summary, the console table and the --csv_filename output
// (src/cmdstan/stansummary.cpp:29-120 options).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cmdstan_amd.h"
#include "stan_csv.h"

namespace {

using stancsv::ChainFile;
using stancsv::display_name;
using stancsv::read_csv;
using stancsv::split;
using stancsv::trim;

double quantile7(std::vector<double> v, double p) {
  std::sort(v.begin(), v.end());
  const double h = (v.size() - 1) * p;
  const size_t lo = (size_t)std::floor(h);
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (h - lo) * (v[hi] - v[lo]);
}

std::string fmt_sig(double v, int sig) {
  std::ostringstream os;
  if (!std::isfinite(v)) {
    os << v;
    return os.str();
  }
  const double a = std::fabs(v);
  if (a != 0 && (a >= 1e6 || a < 1e-3)) {
    os << std::scientific << std::setprecision(std::max(0, sig - 1)) << v;
  } else {
    const int ip = a >= 1 ? (int)std::floor(std::log10(a)) + 1 : 0;
    const int prec = std::max(0, sig - ip);
    os << std::fixed << std::setprecision(ip >= sig ? 0 : prec) << v;
  }
  return os.str();
}

int usage() {
  std::cout << "Usage: stansummary [OPTIONS] stan_csv_file(s)\n"
               "  -c, --csv_filename [file]   Write statistics to a csv file.\n"
               "  -p, --percentiles [values]  Percentiles to report, default \"5,50,95\".\n"
               "  -a, --autocorr [n]          Display the chain autocorrelation for the n-th input file.\n"
               "  -s, --sig_figs [n]          Significant figures reported. Default is 2.\n"
               "  -i, --include_param [name]  Include the named parameter in the summary output.\n";
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> files, include;
  std::string csv_filename, pct_spec = "5,50,95";
  bool have_csv = false;
  int sig_figs = 2;
  int autocorr = 0;  // 1-based input file whose autocorrelations are printed
  try {
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      auto value = [&](const std::string& longname, const std::string& shortname) -> std::string {
        if (a.rfind(longname + "=", 0) == 0) return a.substr(longname.size() + 1);
        if (a == longname || a == shortname) {
          if (i + 1 >= argc) throw std::invalid_argument(longname + ": requires a value");
          return argv[++i];
        }
        return std::string("\x01");
      };
      if (a == "-h" || a == "--help") return usage();
      std::string v;
      if ((v = value("--csv_filename", "-c")) != "\x01") { csv_filename = v; have_csv = true; }
      else if ((v = value("--percentiles", "-p")) != "\x01") pct_spec = v;
      else if ((v = value("--sig_figs", "-s")) != "\x01") {
        sig_figs = std::atoi(v.c_str());
        if (sig_figs < 1 || sig_figs > 18) {
          std::cout << "--sig_figs: Value " << v << " not in range 1 to 18" << std::endl;
          return 1;
        }
      } else if ((v = value("--autocorr", "-a")) != "\x01") {
        autocorr = std::atoi(v.c_str());
        if (autocorr <= 0) {
          std::cout << "--autocorr: Number less or equal to 0" << std::endl;
          return 1;
        }
      } else if ((v = value("--include_param", "-i")) != "\x01") include.push_back(v);
      else if (!a.empty() && a[0] == '-') throw std::invalid_argument("unknown option " + a);
      else files.push_back(a);
    }
    if (files.empty()) return usage(), 1;
    if (autocorr > (int)files.size()) {
      std::cout << "Option --autocorr: " << autocorr << " not a valid chain id." << std::endl;
      return 1;
    }
    // percentiles: numbers in [0, 100], strictly increasing (stansummary.cpp:98-110)
    std::vector<std::string> pcts;
    double prev = -1.0;
    for (auto& p : split(pct_spec, ',')) {
      std::string t = trim(p);
      t.erase(std::remove(t.begin(), t.end(), '"'), t.end());
      if (t.empty()) continue;
      char* end = nullptr;
      const double x = std::strtod(t.c_str(), &end);
      if (*end != '\0' || !(x >= 0 && x <= 100)) {
        std::cout << "Option --percentiles " << pct_spec << ": values must be in range [0, 100]" << std::endl;
        return 1;
      }
      if (!(x > prev)) {
        std::cout << "Option --percentiles " << pct_spec << ": values must be in ascending order" << std::endl;
        return 1;
      }
      prev = x;
      pcts.push_back(t);
    }
    for (auto& f : files) {
      if (!std::ifstream(f).good()) {
        std::cout << "File does not exist: " << f << std::endl;
        return 1;
      }
    }
    std::vector<ChainFile> chains;
    for (auto& f : files) chains.push_back(read_csv(f));
    const auto& names = chains[0].names;
    size_t n = chains[0].rows.size();
    for (auto& c : chains) {
      if (c.names != names) throw std::runtime_error("the files have different columns");
      n = std::min(n, c.rows.size());
    }
    if (n < 2) throw std::runtime_error("need at least 2 draws per chain");
    // column selection: sampler columns, then requested (or all) model columns
    std::vector<std::string> unknown;
    for (auto& inc : include) {
      bool found = false;
      for (auto& nm : names) {
        const std::string dn = display_name(nm);
        found = found || nm == inc || dn == inc || dn.rfind(inc + "[", 0) == 0;
      }
      if (!found) unknown.push_back(inc);
    }
    if (!unknown.empty()) {
      std::cout << "--include_param: Unrecognized parameter(s): ";
      for (auto& u : unknown) std::cout << "'" << u << "' ";
      std::cout << std::endl;
      return 1;
    }
    std::vector<int> cols;
    for (size_t j = 0; j < names.size(); ++j) {
      const bool sampler = names[j].size() > 2 && names[j].compare(names[j].size() - 2, 2, "__") == 0;
      bool keep = include.empty() || sampler;
      for (auto& inc : include) {
        const std::string dn = display_name(names[j]);
        if (names[j] == inc || dn == inc || dn.rfind(inc + "[", 0) == 0) keep = true;
      }
      if (keep) cols.push_back((int)j);
    }
    const int P = (int)cols.size(), M = (int)chains.size();
    std::vector<double> dr((size_t)P * M * n);
    for (int p = 0; p < P; ++p)
      for (int c = 0; c < M; ++c)
        for (size_t k = 0; k < n; ++k) dr[((size_t)p * M + c) * n + k] = chains[c].rows[k][cols[p]];
    std::vector<double> st((size_t)P * 10);
    if (nuts_summary(dr.data(), (int)n, M, P, 8, st.data()) != NUTS_OK) throw std::runtime_error(nuts_last_error());
    // rows: Mean MCSE StdDev MAD pcts... ESS_bulk ESS_tail R_hat
    std::vector<std::string> header = {"Mean", "MCSE", "StdDev", "MAD"};
    for (auto& p : pcts) header.push_back(p + "%");
    header.insert(header.end(), {"ESS_bulk", "ESS_tail", "R_hat"});
    std::vector<std::vector<double>> table(P);
    for (int p = 0; p < P; ++p) {
      const double* s = &st[(size_t)p * 10];
      std::vector<double> all(dr.begin() + (size_t)p * M * n, dr.begin() + (size_t)(p + 1) * M * n);
      std::vector<double> row = {s[0], s[1], s[2], s[3]};
      for (auto& q : pcts) {
        const double pr = std::atof(q.c_str()) / 100.0;
        row.push_back(pr == 0.05 ? s[4] : pr == 0.5 ? s[5] : pr == 0.95 ? s[6] : quantile7(all, pr));
      }
      row.insert(row.end(), {s[7], s[8], s[9]});
      table[p] = row;
    }
    // console
    const ChainFile& c0 = chains[0];
    auto meta = [&](const std::string& k, const std::string& d) { return c0.meta.count(k) ? c0.meta.at(k) : d; };
    std::cout << "Inference for Stan model: " << meta("model", "model") << "\n"
              << M << " chains: each with iter=" << meta("num_samples", std::to_string(n))
              << "; warmup=" << meta("num_warmup", "0") << "; thin=" << meta("thin", "1") << "; " << n * M
              << " iterations saved.\n\n";
    double tw = 0, ts = 0;
    for (auto& c : chains) { tw += c.warmup_s; ts += c.sampling_s; }
    std::cout << std::fixed << std::setprecision(3) << "Warmup took " << tw << " seconds\nSampling took " << ts
              << " seconds\n\n" << std::defaultfloat;
    size_t wname = 4;
    for (int p = 0; p < P; ++p) wname = std::max(wname, display_name(names[cols[p]]).size());
    std::vector<std::vector<std::string>> cells(P, std::vector<std::string>(header.size()));
    std::vector<size_t> w(header.size());
    for (size_t h = 0; h < header.size(); ++h) w[h] = header[h].size();
    for (int p = 0; p < P; ++p)
      for (size_t h = 0; h < header.size(); ++h) {
        const bool ess = header[h] == "ESS_bulk" || header[h] == "ESS_tail";
        cells[p][h] = ess ? (std::isfinite(table[p][h]) ? std::to_string((long long)std::llround(table[p][h])) : "nan")
                          : fmt_sig(table[p][h], sig_figs);
        w[h] = std::max(w[h], cells[p][h].size());
      }
    std::cout << std::string(wname, ' ');
    for (size_t h = 0; h < header.size(); ++h) std::cout << "  " << std::setw((int)w[h]) << header[h];
    std::cout << "\n\n";
    for (int p = 0; p < P; ++p) {
      std::cout << std::left << std::setw((int)wname) << display_name(names[cols[p]]) << std::right;
      for (size_t h = 0; h < header.size(); ++h) std::cout << "  " << std::setw((int)w[h]) << cells[p][h];
      std::cout << "\n";
      if (names[cols[p]] == "energy__") std::cout << "\n";
    }
    std::cout << "\nSamples were drawn using " << meta("algorithm", "hmc") << " with " << meta("engine", "nuts")
              << ".\nFor each parameter, ESS_bulk and ESS_tail measure the effective sample size for the entire "
                 "sample (bulk) and for the .05 and .95 tails (tail), \nand R_hat measures the potential scale "
                 "reduction on split chains. At convergence R_hat will be very close to 1.00.\n";
    if (autocorr > 0) {
      // normalised autocorrelation of every reported column of one chain,
      // lags 0 .. n-1 (stansummary_helper.hpp:700-745 layout)
      const auto& ch = chains[autocorr - 1];
      std::cout << "Displaying the autocorrelations for chain " << autocorr << ":" << std::endl << std::endl;
      const int nd = (int)n;
      int lag_w = 1;
      for (int x = nd; x != 0; x /= 10) ++lag_w;
      size_t wn = 0;
      for (int p = 0; p < P; ++p) wn = std::max(wn, names[cols[p]].size());
      std::vector<std::vector<double>> ac(P, std::vector<double>(nd, 0.0));
      for (int p = 0; p < P; ++p) {
        double m = 0.0, lo = ch.rows[0][cols[p]], hi = lo;
        for (int k = 0; k < nd; ++k) {
          const double x = ch.rows[k][cols[p]];
          m += x;
          lo = std::min(lo, x);
          hi = std::max(hi, x);
        }
        m /= nd;
        if (hi == lo) {  // a constant column has no autocorrelation (0 / 0)
          for (int lag = 0; lag < nd; ++lag) ac[p][lag] = std::nan("");
          continue;
        }
        std::vector<double> d(nd);
        for (int k = 0; k < nd; ++k) d[k] = ch.rows[k][cols[p]] - m;
        double c0 = 0.0;
        for (int k = 0; k < nd; ++k) c0 += d[k] * d[k];
        for (int lag = 0; lag < nd; ++lag) {
          double c = 0.0;
          for (int k = 0; k + lag < nd; ++k) c += d[k] * d[k + lag];
          ac[p][lag] = c / c0;
        }
      }
      std::cout << std::setprecision(6) << std::defaultfloat;
      std::cout << std::setw(lag_w > 4 ? lag_w : 4) << "Lag";
      for (int p = 0; p < P; ++p) std::cout << std::setw((int)wn + 1) << std::right << names[cols[p]];
      std::cout << std::endl;
      for (int lag = 0; lag < nd; ++lag) {
        std::cout << std::setw(lag_w) << std::right << lag;
        for (int p = 0; p < P; ++p) std::cout << std::setw((int)wn + 1) << std::right << ac[p][lag];
        std::cout << std::endl;
      }
    }
    if (have_csv) {
      std::ofstream out(csv_filename);
      if (!out) {
        std::cout << "Cannot save to csv_filename: " << csv_filename << "." << std::endl;
        return 1;
      }
      out << "name";
      for (auto& h : header) out << "," << h;
      out << "\n";
      out << std::setprecision(6);
      for (int p = 0; p < P; ++p) {
        out << "\"" << display_name(names[cols[p]]) << "\"";
        for (double v : table[p]) out << "," << v;
        out << "\n";
      }
    }
  } catch (const std::exception& e) {
    std::cout << e.what() << std::endl;
    return 1;
  }
  return 0;
}
