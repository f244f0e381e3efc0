// diagnose over Stan CSV files written by the sampler (src/cmdstan/diagnose.cpp):
// tree-depth saturation (:84-112), divergences (:113-132), E-BFMI (:133-162),
// bulk/tail ESS per draw below 0.001 and split R-hat above 1.01499 (:163-173),
// with the messages CmdStan prints.  ESS and R-hat come from nuts_summary.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/cmdstan_amd.h"
#include "stan_csv.h"

namespace {

constexpr double kRhatMax = 1.01499;  // printed as 1.01
constexpr double kEbfmiMin = 0.3;
constexpr double kEssPerDraw = 0.001;

void print_list(const std::vector<std::string>& v) {
  std::cout << "  ";
  for (size_t i = 0; i < v.size(); ++i) std::cout << (i ? ", " : "") << v[i];
  std::cout << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 1) {
    std::cout << "USAGE:  diagnose <filename 1> [<filename 2> ... <filename N>]" << std::endl << std::endl;
    return 0;
  }
  std::vector<std::string> files;
  for (int i = 1; i < argc; ++i) {
    if (std::ifstream(argv[i]).good()) files.push_back(argv[i]);
    else std::cout << "File " << argv[i] << " not found" << std::endl;
  }
  if (files.empty()) {
    std::cout << "No valid input files, exiting." << std::endl;
    return 1;
  }
  std::cout << std::fixed << std::setprecision(2);
  std::vector<stancsv::ChainFile> chains;
  for (auto& f : files) {
    try {
      chains.push_back(stancsv::read_csv(f));
    } catch (const std::exception& e) {
      std::cout << "Cannot parse input csv file: " << f << e.what() << "." << std::endl;
      return 1;
    }
  }
  const auto& names = chains[0].names;
  size_t per = chains[0].rows.size();
  for (auto& c : chains) {
    if (c.names != names) {
      std::cout << "Cannot parse input csv file: the files have different columns." << std::endl;
      return 1;
    }
    per = std::min(per, c.rows.size());
  }
  const int M = (int)chains.size();
  const long total = (long)per * M;
  // column j, chain after chain
  auto column = [&](size_t j) {
    std::vector<double> v;
    v.reserve(total);
    for (auto& c : chains)
      for (size_t k = 0; k < per; ++k) v.push_back(c.rows[k][j]);
    return v;
  };
  const auto& meta = chains[0].meta;
  const int max_depth = meta.count("max_depth") ? std::atoi(meta.at("max_depth").c_str()) : 10;
  std::vector<std::string> bad_ess, bad_rhat;
  bool problems = false;
  for (size_t j = 0; j < names.size(); ++j) {
    const std::string& n = names[j];
    if (n == "treedepth__") {
      std::cout << "Checking sampler transitions treedepth." << std::endl;
      long hit = 0;
      for (double t : column(j)) hit += t >= max_depth ? 1 : 0;
      if (hit > 0) {
        problems = true;
        std::cout << hit << " of " << total << " (" << 100.0 * hit / total << "%)"
                  << " transitions hit the maximum treedepth limit of " << max_depth << ", or 2^" << max_depth
                  << " leapfrog steps." << std::endl
                  << "Trajectories that are prematurely terminated due to this limit will result in slow "
                     "exploration."
                  << std::endl
                  << "For optimal performance, increase this limit." << std::endl
                  << std::endl;
      } else {
        std::cout << "Treedepth satisfactory for all transitions." << std::endl << std::endl;
      }
    } else if (n == "divergent__") {
      std::cout << "Checking sampler transitions for divergences." << std::endl;
      long div = 0;
      for (double d : column(j)) div += (long)d;
      if (div > 0) {
        problems = true;
        std::cout << div << " of " << total << " (" << 100.0 * div / total
                  << "%) transitions ended with a divergence." << std::endl
                  << "These divergent transitions indicate that HMC is not fully able to explore the posterior "
                     "distribution."
                  << std::endl
                  << "Try increasing adapt delta closer to 1." << std::endl
                  << "If this doesn't remove all divergences, try to reparameterize the model." << std::endl
                  << std::endl;
      } else {
        std::cout << "No divergent transitions found." << std::endl << std::endl;
      }
    } else if (n == "energy__") {
      // E-BFMI over the draws of all chains taken in sequence: mean squared
      // step of the energy over its variance
      std::cout << "Checking E-BFMI - sampler transitions HMC potential energy." << std::endl;
      const auto e = column(j);
      double step2 = 0.0;
      for (size_t k = 1; k < e.size(); ++k) {
        const double d = e[k] - e[k - 1];
        step2 += (d * d - step2) / (double)k;
      }
      double mean = 0.0;
      for (double x : e) mean += x;
      mean /= (double)e.size();
      double var = 0.0;
      for (double x : e) var += (x - mean) * (x - mean);
      var /= (double)(e.size() - 1);
      const double ebfmi = step2 / var;
      if (ebfmi < kEbfmiMin) {
        problems = true;
        std::cout << "The E-BFMI, " << ebfmi << ", is below the nominal threshold of " << kEbfmiMin
                  << " which suggests that HMC may have trouble exploring the target distribution." << std::endl
                  << "If possible, try to reparameterize the model." << std::endl
                  << std::endl;
      } else {
        std::cout << "E-BFMI satisfactory." << std::endl << std::endl;
      }
    } else if (n.find("__") == std::string::npos) {
      std::vector<double> dr((size_t)M * per);
      for (int c = 0; c < M; ++c)
        for (size_t k = 0; k < per; ++k) dr[(size_t)c * per + k] = chains[c].rows[k][j];
      double st[10];
      if (nuts_summary(dr.data(), (int)per, M, 1, 1, st) != NUTS_OK) {
        std::cout << nuts_last_error() << std::endl;
        return 1;
      }
      const double ess = std::min(st[7], st[8]);
      if (ess / total < kEssPerDraw) bad_ess.push_back(stancsv::display_name(n));
      if (st[9] > kRhatMax) bad_rhat.push_back(stancsv::display_name(n));
    }
  }
  if (!bad_ess.empty()) {
    problems = true;
    std::cout << "The following parameters had fewer than 0.001 effective draws per transition:" << std::endl;
    print_list(bad_ess);
    std::cout << "Such low values indicate that the effective sample size estimators may be biased high and "
                 "actual performance may be substantially lower than quoted."
              << std::endl
              << std::endl;
  } else {
    std::cout << "Rank-normalized split effective sample size satisfactory for all parameters." << std::endl
              << std::endl;
  }
  if (!bad_rhat.empty()) {
    problems = true;
    std::cout << "The following parameters had rank-normalized split R-hat greater than " << kRhatMax << ":"
              << std::endl;
    print_list(bad_rhat);
    std::cout << "Such high values indicate incomplete mixing and biased estimation." << std::endl
              << "You should consider regularizating your model with additional prior information or a more "
                 "effective parameterization."
              << std::endl
              << std::endl;
  } else {
    std::cout << "Rank-normalized split R-hat values satisfactory for all parameters." << std::endl << std::endl;
  }
  std::cout << (problems ? "Processing complete." : "Processing complete, no problems detected.") << std::endl;
  return 0;
}
