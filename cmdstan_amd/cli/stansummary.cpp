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
#include <set>
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

// Fixed-point slot of one value at `sig` significant digits: its width and
// decimal places (the column rule of stansummary_helper.hpp:20-48).
void fixed_slot(double v, int sig, int& width, int& prec) {
  const double a = std::fabs(v);
  if (v == 0) {
    width = sig;
    prec = sig;
  } else if (std::isnan(v)) {
    width = 3;
    prec = sig;
  } else if (a >= 1) {
    const int ip = (int)std::ceil(std::log10(a) + 1e-6);
    width = ip >= sig ? ip : sig + 1;
    prec = ip >= sig ? 0 : sig - ip;
  } else {
    const int fp = (int)std::fabs(std::floor(std::log10(a)));
    width = 1 + fp + sig;
    prec = fp + sig - 1;
  }
  if (std::signbit(v)) ++width;
}

// A column prints fixed-point when its widest value plus two spaces of padding
// stays under 8 characters, else in scientific notation
// (stansummary_helper.hpp:94-124); returns the column width.
int column_format(const std::vector<double>& x, const std::string& name, int sig, bool& scientific) {
  int wmax = 0;
  bool neg = false;
  for (double v : x) {
    int w, pr;
    fixed_slot(v, sig, w, pr);
    wmax = std::max(wmax, w);
    neg = neg || v < 0;
  }
  if (wmax + 2 < 8) {
    scientific = false;
    return std::max<int>(wmax, (int)name.size()) + 2;
  }
  scientific = true;
  const int ws = sig + 1 + 4 + (neg ? 1 : 0);
  return std::max<int>(ws, (int)name.size()) + 2;
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
  bool sig_figs_given = false;  // the csv file keeps 6 digits unless --sig_figs is given
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
        sig_figs_given = true;
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
    // percentiles: the spec is trimmed and split on commas and spaces (runs of
    // separators count once); each value is read as a number prefix, must be
    // finite, within [0, 100] and not below the integer part of the one
    // before it; any violation is one message (stansummary.cpp:98-110,
    // stansummary_helper.hpp:300-320)
    std::vector<std::string> pcts;
    {
      const std::string spec = trim(pct_spec);
      std::string tok;
      auto flush = [&]() {
        if (!tok.empty()) pcts.push_back(tok);
        tok.clear();
      };
      for (char ch : spec) {
        if (ch == ',' || ch == ' ') flush();
        else tok += ch;
      }
      flush();
      int cur = 0;
      for (auto& t : pcts) {
        char* end = nullptr;
        const double x = std::strtod(t.c_str(), &end);
        if (end == t.c_str() || !std::isfinite(x) || x < 0.0 || x > 100.0 || x < cur) {
          std::cout << "Option --percentiles " << spec
                    << ": values must be in range (0, 100), inclusive, and strictly increasing." << std::endl;
          return 1;
        }
        cur = (int)x;
      }
    }
    if (have_csv) {  // the summary csv is never written over an existing file
      std::ifstream probe(csv_filename);
      if (probe.good()) {
        std::cout << "--csv_filename: Path already exists: " << csv_filename << std::endl;
        return 1;
      }
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
    // column selection: lp__ and the sampler columns always; the model
    // columns all, or those whose display name (theta[1]; a request written
    // theta.1 is converted first) equals a requested one exactly. Requests
    // are a set; any left unmatched are reported in sorted order
    // (stansummary.cpp:158-195).
    auto is_sampler_name = [](const std::string& nm) {
      return nm.size() >= 2 && nm.compare(nm.size() - 2, 2, "__") == 0;
    };
    std::set<std::string> requested;
    for (auto& inc : include) requested.insert(display_name(inc));
    std::vector<int> cols;
    for (size_t j = 0; j < names.size(); ++j) {
      if (is_sampler_name(names[j])) cols.push_back((int)j);
      else if (include.empty() || requested.erase(display_name(names[j])) > 0) cols.push_back((int)j);
    }
    if (!requested.empty()) {
      std::cout << "--include_param: Unrecognized parameter(s): ";
      for (auto& u : requested) std::cout << "'" << u << "' ";
      std::cout << std::endl;
      return 1;
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
    // model, chain counts and timing (stansummary_helper.hpp:595-680), then the
    // closing notes (:682-697); the csv file repeats them as "# " comments
    auto write_timing = [&](const std::string& pre, std::ostream& o) {
      o << pre << "Inference for Stan model: " << meta("model", "model") << "\n"
        << pre << M << " chains: each with iter=" << meta("num_samples", std::to_string(n)) << "; warmup="
        << meta("num_warmup", "0") << "; thin=" << meta("thin", "1") << "; " << n * M << " iterations saved.\n"
        << pre << "\n";
      auto took = [&](const char* what, bool warm) {
        double tot = 0;
        for (auto& c : chains) tot += warm ? c.warmup_s : c.sampling_s;
        std::string unit = "seconds";
        if (tot / 3600 > 1) { tot /= 3600; unit = "hours"; }
        else if (tot / 60 > 1) { tot /= 60; unit = "minutes"; }
        auto fx = [&](double v) {
          int w, pr;
          fixed_slot(v, 2, w, pr);
          o << std::fixed << std::setprecision(pr) << v;
        };
        o << pre << what << " took ";
        if (chains.size() == 1) {
          fx(tot);
          o << " " << unit << "\n";
        } else {
          o << "(";
          for (size_t c = 0; c < chains.size(); ++c) {
            if (c) o << ", ";
            fx(warm ? chains[c].warmup_s : chains[c].sampling_s);
          }
          o << ") seconds, ";
          fx(tot);
          o << " " << unit << " total\n";
        }
        o << std::defaultfloat;
      };
      took("Warmup", true);
      took("Sampling", false);
    };
    auto write_sampler_info = [&](const std::string& pre, std::ostream& o) {
      o << pre << "Samples were drawn using " << meta("algorithm", "hmc") << " with " << meta("engine", "nuts")
        << ".\n"
        << pre << "For each parameter, ESS_bulk and ESS_tail measure the effective sample size for the entire "
                  "sample (bulk) and for the .05 and .95 tails (tail), \n"
        << pre << "and R_hat measures the potential scale reduction on split chains. At convergence R_hat will "
                  "be very close to 1.00.\n";
    };
    write_timing("", std::cout);
    std::cout << "\n";
    // formats per column are chosen separately for the sampler rows and the
    // model rows (lp__ prints with the model rows' formats); a column is as
    // wide as the wider of the two (stansummary.cpp:205-246)
    size_t wname = 0;  // over every column of the file, reported or not
    for (auto& nm : names) wname = std::max(wname, display_name(nm).size());
    auto is_sampler = [&](int p) {
      const std::string& nm = names[cols[p]];
      return nm != "lp__" && nm.size() > 2 && nm.compare(nm.size() - 2, 2, "__") == 0;
    };
    auto is_model = [&](int p) { return names[cols[p]].find("__") == std::string::npos; };
    std::vector<int> cw(header.size());
    std::vector<char> sci_s(header.size()), sci_m(header.size());
    for (size_t h = 0; h < header.size(); ++h) {
      std::vector<double> cs, cm;
      for (int p = 0; p < P; ++p) {
        if (is_sampler(p)) cs.push_back(table[p][h]);
        else if (is_model(p)) cm.push_back(table[p][h]);
      }
      bool a = false, b = false;
      const int ws = column_format(cs, header[h], sig_figs, a);
      const int wm = column_format(cm, header[h], sig_figs, b);
      cw[h] = std::max(ws, wm);
      sci_s[h] = a;
      sci_m[h] = b;
    }
    auto row = [&](int p, const std::vector<char>& sci) {
      std::cout << std::setw((int)wname + 1) << std::left << display_name(names[cols[p]]) << std::right;
      for (size_t h = 0; h < header.size(); ++h) {
        const double v = table[p][h];
        int w, pr;
        fixed_slot(v, sig_figs, w, pr);
        if (sci[h]) std::cout << std::scientific << std::setprecision(sig_figs - 1);
        else std::cout << std::fixed << std::setprecision(pr);
        std::cout << std::setw(cw[h]) << v;
      }
      std::cout << std::defaultfloat << "\n";
    };
    std::cout << std::setw((int)wname + 1) << "";
    for (size_t h = 0; h < header.size(); ++h) std::cout << std::setw(cw[h]) << header[h];
    std::cout << "\n\n";
    for (int p = 0; p < P; ++p)
      if (names[cols[p]] == "lp__") row(p, sci_m);
    for (int p = 0; p < P; ++p)
      if (is_sampler(p)) row(p, sci_s);
    std::cout << "\n";
    for (int p = 0; p < P; ++p)
      if (is_model(p)) row(p, sci_m);
    std::cout << "\n";
    write_sampler_info("", std::cout);
    if (autocorr > 0) {
      // normalised autocorrelation of every column of one chain (lp__, the
      // sampler columns and all model columns, whatever --include_param
      // selects), lags 0 .. n-1, under the display names
      // (stansummary_helper.hpp:700-745 layout)
      const auto& ch = chains[autocorr - 1];
      std::cout << "Displaying the autocorrelations for chain " << autocorr << ":" << std::endl << std::endl;
      const int nd = (int)n, NC = (int)names.size();
      int lag_w = 1;
      for (int x = nd; x != 0; x /= 10) ++lag_w;
      std::vector<std::vector<double>> ac(NC, std::vector<double>(nd, 0.0));
      for (int p = 0; p < NC; ++p) {
        double m = 0.0, lo = ch.rows[0][p], hi = lo;
        for (int k = 0; k < nd; ++k) {
          const double x = ch.rows[k][p];
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
        for (int k = 0; k < nd; ++k) d[k] = ch.rows[k][p] - m;
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
      for (int p = 0; p < NC; ++p) std::cout << std::setw((int)wname + 1) << std::right << display_name(names[p]);
      std::cout << std::endl;
      for (int lag = 0; lag < nd; ++lag) {
        std::cout << std::setw(lag_w) << std::right << lag;
        for (int p = 0; p < NC; ++p) std::cout << std::setw((int)wname + 1) << std::right << ac[p][lag];
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
      out << std::setprecision(sig_figs_given ? sig_figs : 6);
      for (int p = 0; p < P; ++p) {
        out << "\"" << display_name(names[cols[p]]) << "\"";
        for (double v : table[p]) out << "," << v;
        out << "\n";
      }
      write_timing("# ", out);
      write_sampler_info("# ", out);
    }
  } catch (const std::exception& e) {
    std::cout << e.what() << std::endl;
    return 1;
  }
  return 0;
}
