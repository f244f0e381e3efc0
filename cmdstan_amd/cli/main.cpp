// Command-line front end with CmdStan's `sample` grammar and Stan CSV output.
//
// Usage (the model family is the executable's name, or --model=<family>):
//   bernoulli sample num_chains=4 data file=bernoulli.data.json output file=out.csv
//
// Mirrors, for method=sample algorithm=hmc engine=nuts:
//   argument tree + parsing ..... src/cmdstan/arguments/argument_parser.hpp:22-135,
//                                 categorical_argument.hpp:69-120, list_argument.hpp:50-96
//   defaults ..................... arg_sample.hpp:18-29, arg_adapt.hpp:19-37, arg_hmc.hpp:19-27,
//                                 arg_nuts.hpp:15-16, arg_metric.hpp:17-22, arg_output.hpp:20-29
//   command orchestration ........ src/cmdstan/command.hpp:84-292, 552-697
//   per-chain file names ......... src/cmdstan/file.hpp:189-212 (make_filenames)
//   config header / CSV layout ... write_config.hpp:19-29; example_output/bernoulli_chain_1.csv
//   save_metric JSON ............. command.hpp:249-264
// Exit codes follow return_codes.hpp:21-23 (0 OK, 1 NOT_OK).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cmdstan_amd.h"
#include "stan_csv.h"

namespace {

// ------------------------------------------------------------ argument tree
struct Node {
  std::string name;
  enum Kind { GROUP, LIST, VALUE } kind = VALUE;
  std::string value, deflt;
  std::vector<std::string> choices;  // LIST: names of child groups
  std::vector<std::unique_ptr<Node>> kids;
  Node* parent = nullptr;
  bool set = false;
  Node* add(const std::string& n, Kind k, const std::string& d = "") {
    kids.emplace_back(new Node());
    Node* c = kids.back().get();
    c->name = n; c->kind = k; c->value = d; c->deflt = d; c->parent = this;
    return c;
  }
  Node* kid(const std::string& n) const {
    for (auto& k : kids) if (k->name == n) return k.get();
    return nullptr;
  }
};

std::unique_ptr<Node> make_tree() {
  std::unique_ptr<Node> root(new Node());
  root->kind = Node::GROUP;
  Node* method = root->add("method", Node::LIST, "sample");
  method->choices = {"sample", "optimize", "variational", "diagnose", "generate_quantities", "pathfinder",
                     "log_prob", "laplace"};
  Node* lp = method->add("log_prob", Node::GROUP);  // arg_log_prob.hpp
  lp->add("unconstrained_params", Node::VALUE, "");
  lp->add("constrained_params", Node::VALUE, "");
  lp->add("jacobian", Node::VALUE, "true");
  Node* sample = method->add("sample", Node::GROUP);
  sample->add("num_samples", Node::VALUE, "1000");
  sample->add("num_warmup", Node::VALUE, "1000");
  sample->add("save_warmup", Node::VALUE, "false");
  sample->add("thin", Node::VALUE, "1");
  Node* adapt = sample->add("adapt", Node::GROUP);
  adapt->add("engaged", Node::VALUE, "true");
  adapt->add("gamma", Node::VALUE, "0.05");
  adapt->add("delta", Node::VALUE, "0.8");
  adapt->add("kappa", Node::VALUE, "0.75");
  adapt->add("t0", Node::VALUE, "10");
  adapt->add("init_buffer", Node::VALUE, "75");
  adapt->add("term_buffer", Node::VALUE, "50");
  adapt->add("window", Node::VALUE, "25");
  adapt->add("save_metric", Node::VALUE, "false");
  Node* algo = sample->add("algorithm", Node::LIST, "hmc");
  algo->choices = {"hmc", "fixed_param"};
  Node* hmc = algo->add("hmc", Node::GROUP);
  algo->add("fixed_param", Node::GROUP);
  Node* engine = hmc->add("engine", Node::LIST, "nuts");
  engine->choices = {"nuts", "static"};
  Node* nuts = engine->add("nuts", Node::GROUP);
  nuts->add("max_depth", Node::VALUE, "10");
  Node* stat = engine->add("static", Node::GROUP);
  stat->add("int_time", Node::VALUE, "6.28319");
  hmc->add("metric", Node::VALUE, "diag_e");
  hmc->add("metric_file", Node::VALUE, "");
  hmc->add("stepsize", Node::VALUE, "1");
  hmc->add("stepsize_jitter", Node::VALUE, "0");
  sample->add("num_chains", Node::VALUE, "1");
  root->add("id", Node::VALUE, "1");
  Node* data = root->add("data", Node::GROUP);
  data->add("file", Node::VALUE, "");
  root->add("init", Node::VALUE, "2");
  Node* rnd = root->add("random", Node::GROUP);
  rnd->add("seed", Node::VALUE, "-1");
  Node* out = root->add("output", Node::GROUP);
  out->add("file", Node::VALUE, "output.csv");
  out->add("diagnostic_file", Node::VALUE, "");
  out->add("refresh", Node::VALUE, "100");
  out->add("sig_figs", Node::VALUE, "-1");
  out->add("profile_file", Node::VALUE, "profile.csv");
  out->add("save_cmdstan_config", Node::VALUE, "false");
  root->add("num_threads", Node::VALUE, "1");
  return root;
}

// Subarguments follow their parent; within a level the order is free.  A name
// not found at the current level is looked up in the enclosing levels.
void parse(Node* root, const std::vector<std::string>& toks) {
  Node* ctx = root;
  for (const std::string& t : toks) {
    std::string name = t, value;
    const size_t eq = t.find('=');
    bool has_value = eq != std::string::npos;
    if (has_value) { name = t.substr(0, eq); value = t.substr(eq + 1); }
    if (!has_value && ctx == root && (name == "sample" || name == "optimize" || name == "variational" ||
                                      name == "diagnose" || name == "generate_quantities" ||
                                      name == "pathfinder" || name == "log_prob" || name == "laplace")) {
      has_value = true; value = name; name = "method";
    }
    Node* found = nullptr;
    for (Node* c = ctx; c; c = c->parent) {
      if ((found = c->kid(name))) break;
      // a bare list choice (e.g. "hmc", "nuts") names the selected group
      if (!has_value) {
        for (auto& k : c->kids)
          if (k->kind == Node::LIST && k->value == name && k->kid(name)) { found = k->kid(name); break; }
        if (found) break;
      }
    }
    if (!found) throw std::invalid_argument(name + " is either mistyped or misplaced.");
    if (found->kind == Node::VALUE) {
      if (!has_value) throw std::invalid_argument(name + " requires a value");
      found->value = value;
      found->set = true;
      ctx = found->parent;
    } else if (found->kind == Node::LIST) {
      if (!has_value) throw std::invalid_argument(name + " requires a value");
      bool ok = false;
      for (auto& ch : found->choices) ok |= ch == value;
      if (!ok) throw std::invalid_argument(value + " is not a valid value for \"" + name + "\"");
      found->value = value;
      found->set = true;
      ctx = found->kid(value) ? found->kid(value) : found->parent;
    } else {
      ctx = found;
    }
  }
}

void print_tree(const Node* n, int depth, std::ostream& o, const std::string& prefix) {
  for (auto& k : n->kids) {
    if (n->kind == Node::LIST && k->name != n->value) continue;  // unselected choice
    const std::string ind(2 * depth, ' ');
    if (k->kind == Node::GROUP) {
      o << prefix << ind << k->name << "\n";
      print_tree(k.get(), depth + 1, o, prefix);
    } else {
      o << prefix << ind << k->name << " = " << k->value << (k->set ? "" : " (Default)") << "\n";
      if (k->kind == Node::LIST) print_tree(k.get(), depth + 1, o, prefix);
    }
  }
}

// JSON form of the argument tree (write_config.hpp:31-47 with the structured
// writer): groups are objects, a list is {"value": choice, choice: {...}},
// values are numbers, booleans or strings.
std::string json_value(const std::string& v) {
  if (v == "true" || v == "false") return v;
  static const std::regex number("-?(0|[1-9][0-9]*)(\\.[0-9]+)?([eE][+-]?[0-9]+)?");
  if (std::regex_match(v, number)) return v;  // a JSON number as written
  std::string o = "\"";
  for (char ch : v) {
    if (ch == '"' || ch == '\\') o += '\\';
    o += ch;
  }
  return o + "\"";
}

void print_json_members(const Node* n, std::ostream& o, int depth, bool& first) {
  const std::string ind(2 * depth, ' ');
  for (auto& k : n->kids) {
    if (n->kind == Node::LIST) continue;  // a list prints its chosen member itself
    o << (first ? "" : ",\n") << ind << "\"" << k->name << "\": ";
    first = false;
    if (k->kind == Node::VALUE) {
      o << json_value(k->value);
    } else if (k->kind == Node::GROUP) {
      o << "{\n";
      bool f = true;
      print_json_members(k.get(), o, depth + 1, f);
      o << "\n" << ind << "}";
    } else {
      o << "{\n" << ind << "  \"value\": " << json_value(k->value);
      const Node* chosen = k->kid(k->value);
      if (chosen && chosen->kind == Node::GROUP) {
        o << ",\n" << ind << "  \"" << chosen->name << "\": {\n";
        bool f = true;
        print_json_members(chosen, o, depth + 2, f);
        o << "\n" << ind << "  }";
      }
      o << "\n" << ind << "}";
    }
  }
}

const Node* find(const Node* n, const std::vector<std::string>& path) {
  for (auto& p : path) {
    n = n->kid(p);
    if (!n) throw std::logic_error("internal: no argument " + p);
  }
  return n;
}
std::string val(const Node* r, const std::vector<std::string>& path) { return find(r, path)->value; }

bool to_bool(const std::string& v, const std::string& name) {
  if (v == "1" || v == "true") return true;
  if (v == "0" || v == "false") return false;
  throw std::invalid_argument(name + " must be a boolean (0/1/true/false)");
}
long long to_int(const std::string& v, const std::string& name) {
  char* e = nullptr;
  const long long x = std::strtoll(v.c_str(), &e, 10);
  if (v.empty() || *e) throw std::invalid_argument(name + " must be an integer, got " + v);
  return x;
}
double to_real(const std::string& v, const std::string& name) {
  char* e = nullptr;
  const double x = std::strtod(v.c_str(), &e);
  if (v.empty() || *e) throw std::invalid_argument(name + " must be a number, got " + v);
  return x;
}

std::string read_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::invalid_argument("Can't open specified file, \"" + path + "\"");
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// file.hpp get_basename_suffix / make_filenames
std::vector<std::string> make_filenames(const std::string& filename, const std::string& tag, const std::string& type,
                                        int num_chains, int id) {
  std::string base = filename, sfx;
  const size_t dot = filename.find_last_of('.');
  const size_t sep = filename.find_last_of('/');
  if (dot != std::string::npos && (sep == std::string::npos || dot > sep) && dot > 0) {
    base = filename.substr(0, dot);
    sfx = filename.substr(dot);
  }
  if (type != ".csv" || sfx.empty()) sfx = type;
  std::vector<std::string> names;
  for (int i = 0; i < num_chains; ++i)
    names.push_back(base + tag + (num_chains == 1 ? "" : "_" + std::to_string(i + id)) + sfx);
  return names;
}

void fmt_row(std::ostream& o, const double* v, int n) {
  for (int i = 0; i < n; ++i) {
    if (i) o << ",";
    o << v[i];
  }
  o << "\n";
}

// get_var_context (command_helper.hpp:116-133): a .json file is JSON, any other
// file is read as R dump, with CmdStan's deprecation warning.
void load_var_context(const std::string& path, nuts_data* d, const std::string& what) {
  const std::string text = read_file(path);
  const size_t dot = path.find_last_of('.');
  const bool json = dot != std::string::npos && path.substr(dot) == ".json";
  int rc;
  if (json) {
    rc = nuts_data_parse_json(d, text.c_str());
  } else {
    std::cerr << "Warning: file '" << path << "' is being read as an 'RDump' file.\n"
              << "\tThis format is deprecated and will not receive new features.\n"
              << "\tConsider saving your data in JSON format instead." << std::endl;
    rc = nuts_data_parse_rdump(d, text.c_str());
  }
  if (rc != NUTS_OK) throw std::invalid_argument("Error reading " + what + " file: " + nuts_last_error());
}

std::string file_header(const std::string& family) {
  std::ostringstream f;
  f << "# stan_version_major = 2\n# stan_version_minor = 36\n# stan_version_patch = 0\n";
  f << "# model = " << family << "_model\n";
  std::time_t now = std::time(nullptr);
  char tb[64];
  std::strftime(tb, sizeof tb, "%Y-%m-%d %H:%M:%S UTC", std::gmtime(&now));
  f << "# start_datetime = " << tb << "\n";
  return f.str();
}

// Constrained parameter values of one Stan CSV row to a var_context: columns
// "name" or "name.i.j..." (Stan CSV order) become arrays in row-major order.
nuts_data* row_context(const std::vector<std::string>& cols, const std::vector<double>& row,
                       const std::vector<std::string>& wanted) {
  struct Arr { std::vector<std::vector<int64_t>> idx; std::vector<double> v; };
  std::map<std::string, Arr> by;
  for (size_t c = 0; c < cols.size(); ++c) {
    const std::string& n = cols[c];
    const size_t dot = n.find('.');
    const std::string base = n.substr(0, dot);
    if (std::find(wanted.begin(), wanted.end(), base) == wanted.end()) continue;
    std::vector<int64_t> ix;
    if (dot != std::string::npos) {
      std::stringstream ss(n.substr(dot + 1));
      std::string t;
      while (std::getline(ss, t, '.')) ix.push_back(std::atoll(t.c_str()) - 1);
    }
    by[base].idx.push_back(ix);
    by[base].v.push_back(row[c]);
  }
  nuts_data* d = nuts_data_create();
  for (auto& kv : by) {
    const Arr& a = kv.second;
    const size_t nd = a.idx[0].size();
    std::vector<int64_t> dims(nd, 0);
    for (auto& ix : a.idx) {
      if (ix.size() != nd) throw std::invalid_argument("constrained_params: inconsistent columns for " + kv.first);
      for (size_t k = 0; k < nd; ++k) dims[k] = std::max(dims[k], ix[k] + 1);
    }
    int64_t n = 1;
    for (auto x : dims) n *= x;
    if (n != (int64_t)a.v.size()) throw std::invalid_argument("constrained_params: missing columns for " + kv.first);
    std::vector<double> vals((size_t)n);
    for (size_t e = 0; e < a.v.size(); ++e) {
      int64_t off = 0;
      for (size_t k = 0; k < nd; ++k) off = off * dims[k] + a.idx[e][k];
      vals[(size_t)off] = a.v[e];
    }
    nuts_data_add(d, kv.first.c_str(), (int)nd, dims.data(), vals.data(), 0);
  }
  return d;
}

// method=log_prob (command.hpp:414-458, services_log_prob_grad at
// command_helper.hpp:569-600): lp__ and g_<unconstrained name> per point
int run_log_prob(const Node* R, const std::string& family, const std::string& cfgtxt) {
  const std::vector<std::string> L = {"method", "log_prob"};
  auto lv = [&](const std::string& k) { std::vector<std::string> p = L; p.push_back(k); return val(R, p); };
  const std::string upars = lv("unconstrained_params"), cpars = lv("constrained_params");
  const bool jacobian = to_bool(lv("jacobian"), "jacobian");
  const std::string out_file = val(R, {"output", "file"});
  const int sig_figs = (int)to_int(val(R, {"output", "sig_figs"}), "sig_figs");
  if (sig_figs < -1 || sig_figs > 18) throw std::invalid_argument("sig_figs must be in [-1, 18]");
  if (!cpars.empty() && cpars == out_file)
    throw std::invalid_argument("Filename conflict, constrained_params file " + cpars +
                                " and output file names are identical, must be different.");
  if (upars.empty() && cpars.empty())
    throw std::invalid_argument("No input parameter files provided, cannot calculate log probability density.");
  if (!upars.empty() && !cpars.empty())
    throw std::invalid_argument("Cannot specify both input files of both constrained and unconstrained parameter "
                                "values.");
  std::cout << cfgtxt << std::endl;
  nuts_data* data = nuts_data_create();
  const std::string data_file = val(R, {"data", "file"});
  if (!data_file.empty()) load_var_context(data_file, data, "data");
  const nuts_data* dl[1] = {data};
  nuts_model* model = nuts_model_create(family.c_str(), dl, 1);
  if (!model) throw std::invalid_argument(nuts_last_error());
  int D = 0, nout = 0;
  nuts_model_dims(model, 0, &D, &nout);
  char buf[256];
  std::vector<std::string> un, pn;
  for (int i = 0; i < D; ++i) { nuts_model_unconstrained_name(model, 0, i, buf, sizeof buf); un.push_back(buf); }
  for (int i = 0; i < nout; ++i) {
    nuts_model_param_name(model, 0, i, buf, sizeof buf);
    const std::string n = buf;
    const std::string base = n.substr(0, n.find('.'));
    if (std::find(pn.begin(), pn.end(), base) == pn.end()) pn.push_back(base);
  }
  std::vector<double> q;  // [points][D]
  if (!upars.empty()) {
    // get_uparams_r (command_helper.hpp:520-555): params_r, a vector or an array of vectors
    nuts_data* u = nuts_data_create();
    load_var_context(upars, u, "unconstrained parameters");
    int nd = 0;
    int64_t dims[8];
    const double* v = nullptr;
    if (nuts_data_get(u, "params_r", &nd, dims, &v) != NUTS_OK || nd < 1 || dims[0] == 0)
      throw std::invalid_argument("Unconstrained parameters file has no variable 'params_r' with unconstrained "
                                  "parameter values!");
    const int64_t rows = nd == 2 ? dims[0] : 1, cols = nd == 2 ? dims[1] : dims[0];
    if (cols != D)
      throw std::invalid_argument("Incorrect number of unconstrained parameters provided! Model has " +
                                  std::to_string(D) + " parameters but " + std::to_string(cols) + " were found.");
    q.assign(v, v + rows * cols);
    nuts_data_destroy(u);
  } else {
    const size_t dot = cpars.find_last_of('.');
    const bool csv = dot != std::string::npos && cpars.substr(dot) == ".csv";
    std::vector<nuts_data*> ctx;
    if (csv) {
      const stancsv::ChainFile f = stancsv::read_csv(cpars);
      for (const auto& row : f.rows) ctx.push_back(row_context(f.names, row, pn));
      if (ctx.empty()) throw std::invalid_argument("constrained_params: no draws in " + cpars);
    } else {
      nuts_data* c = nuts_data_create();
      load_var_context(cpars, c, "constrained parameters");
      ctx.push_back(c);
    }
    q.resize(ctx.size() * (size_t)D);
    for (size_t k = 0; k < ctx.size(); ++k) {
      const int rc = nuts_model_transform_inits(model, 0, ctx[k], q.data() + k * D);
      nuts_data_destroy(ctx[k]);
      if (rc != NUTS_OK) throw std::invalid_argument(nuts_last_error());
      for (int i = 0; i < D; ++i)
        if (std::isnan(q[k * D + i]))
          throw std::invalid_argument("constrained_params: no value for parameter " + un[i]);
    }
  }
  const int n = (int)(q.size() / D);
  std::vector<double> lp(n), g((size_t)n * D);
  std::ofstream f(out_file);
  if (!f) throw std::invalid_argument("Can't open output file " + out_file);
  f << file_header(family);
  std::istringstream ct(cfgtxt);
  std::string line;
  while (std::getline(ct, line)) f << "# " << line << "\n";
  if (nuts_log_prob_grad(model, 0, q.data(), n, jacobian ? 1 : 0, lp.data(), g.data()) != NUTS_OK) {
    std::cerr << "Error during log_prob calculation:\n" << nuts_last_error() << std::endl;
    return 1;
  }
  f << std::setprecision(sig_figs < 0 ? 6 : sig_figs) << "lp__";
  for (auto& nm : un) f << ",g_" << nm;
  f << "\n";
  for (int k = 0; k < n; ++k) {
    f << lp[k];
    for (int i = 0; i < D; ++i) f << "," << g[(size_t)k * D + i];
    f << "\n";
  }
  nuts_model_destroy(model);
  nuts_data_destroy(data);
  return 0;
}

int usage(const std::string& exe) {
  std::cout << "Usage: " << exe << " <arg1> <subarg1_1> ... <subarg1_m> ... <arg_n> <subarg_n_1> ... <subarg_n_m>\n\n"
            << "Begin by selecting amongst the following inference methods and diagnostics,\n"
            << "  sample      Bayesian inference with Markov Chain Monte Carlo\n\n"
            << "Only method=sample algorithm=hmc engine=nuts runs on this build (MI355X batched NUTS).\n"
            << "Model families: bernoulli logistic_regression eight_schools_cp hier_normal_nc hier_logistic_cp gauss\n";
  return 0;
}

int run(int argc, char** argv) {
  std::string exe = argv[0];
  std::string family = exe.substr(exe.find_last_of('/') + 1);
  std::vector<std::string> toks;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--model=", 0) == 0) { family = a.substr(8); continue; }
    if (a == "help" || a == "help-all" || a == "--help") return usage(exe);
    toks.push_back(a);
  }
  if (const char* fam = std::getenv("CMDSTAN_AMD_MODEL")) if (family == "cmdstan_amd") family = fam;
  if (toks.empty()) return usage(exe);

  auto tree = make_tree();
  parse(tree.get(), toks);
  const Node* R = tree.get();
  const std::string method = val(R, {"method"});
  if (method == "log_prob") {
    std::ostringstream lpt;
    print_tree(R, 0, lpt, "");
    return run_log_prob(R, family, lpt.str());
  }
  if (method != "sample")
    throw std::invalid_argument("method=" + method + " is not provided by this build; method=sample and "
                                "method=log_prob run here");
  const std::vector<std::string> S = {"method", "sample"};
  auto sv = [&](std::vector<std::string> p) { std::vector<std::string> q = S; q.insert(q.end(), p.begin(), p.end()); return val(R, q); };
  if (sv({"algorithm"}) != "hmc")
    throw std::invalid_argument("algorithm=" + sv({"algorithm"}) + " is not provided by this build");
  if (sv({"algorithm", "hmc", "engine"}) != "nuts")
    throw std::invalid_argument("engine=" + sv({"algorithm", "hmc", "engine"}) + " is not provided by this build");

  nuts_config cfg;
  nuts_config_default(&cfg);
  cfg.num_samples = (int)to_int(sv({"num_samples"}), "num_samples");
  cfg.num_warmup = (int)to_int(sv({"num_warmup"}), "num_warmup");
  cfg.save_warmup = to_bool(sv({"save_warmup"}), "save_warmup");
  cfg.thin = (int)to_int(sv({"thin"}), "thin");
  cfg.adapt_engaged = to_bool(sv({"adapt", "engaged"}), "engaged");
  cfg.gamma = to_real(sv({"adapt", "gamma"}), "gamma");
  cfg.delta = to_real(sv({"adapt", "delta"}), "delta");
  cfg.kappa = to_real(sv({"adapt", "kappa"}), "kappa");
  cfg.t0 = to_real(sv({"adapt", "t0"}), "t0");
  cfg.init_buffer = (int)to_int(sv({"adapt", "init_buffer"}), "init_buffer");
  cfg.term_buffer = (int)to_int(sv({"adapt", "term_buffer"}), "term_buffer");
  cfg.window = (int)to_int(sv({"adapt", "window"}), "window");
  const bool save_metric = to_bool(sv({"adapt", "save_metric"}), "save_metric");
  cfg.max_depth = (int)to_int(sv({"algorithm", "hmc", "engine", "nuts", "max_depth"}), "max_depth");
  const std::string metric = sv({"algorithm", "hmc", "metric"});
  if (metric == "unit_e") cfg.metric = NUTS_METRIC_UNIT_E;
  else if (metric == "diag_e") cfg.metric = NUTS_METRIC_DIAG_E;
  else if (metric == "dense_e") cfg.metric = NUTS_METRIC_DENSE_E;
  else throw std::invalid_argument(metric + " is not a valid value for \"metric\"");
  const std::string metric_file = sv({"algorithm", "hmc", "metric_file"});
  cfg.stepsize = to_real(sv({"algorithm", "hmc", "stepsize"}), "stepsize");
  cfg.stepsize_jitter = to_real(sv({"algorithm", "hmc", "stepsize_jitter"}), "stepsize_jitter");
  const long long nch = to_int(sv({"num_chains"}), "num_chains");
  if (nch < 1) throw std::invalid_argument("num_chains must be >= 1");
  cfg.num_chains = (int)nch;
  const long long id = to_int(val(R, {"id"}), "id");
  if (id < 0) throw std::invalid_argument("id must be >= 0");
  cfg.id = (int)id;
  long long seed = to_int(val(R, {"random", "seed"}), "seed");
  if (seed < -1 || seed > 4294967295LL)
    throw std::invalid_argument("seed must be in [-1, 4294967295]");  // arg_seed.hpp:19-47
  if (seed == -1) {
    // arg_seed: a time-based seed, echoed in the configuration and the CSV header
    seed = (long long)(std::chrono::system_clock::now().time_since_epoch().count() % 4294967296LL);
    tree->kid("random")->kid("seed")->value = std::to_string(seed);
  }
  cfg.seed = (uint64_t)seed;
  const std::string init = val(R, {"init"});
  char* ie = nullptr;
  const double radius = std::strtod(init.c_str(), &ie);
  const bool init_is_file = init.empty() || *ie;
  if (!init_is_file) cfg.init_radius = radius;
  const long long refresh = to_int(val(R, {"output", "refresh"}), "refresh");
  if (refresh < 0) throw std::invalid_argument("refresh must be >= 0");
  cfg.refresh = (int)refresh;
  if (save_metric && !cfg.adapt_engaged)
    throw std::invalid_argument("Saving the adapted metric (save_metric=1) can only be enabled when adaptation is "
                                "engaged (adapt engaged=1)");  // command.hpp:249-259
  if (nuts_config_validate(&cfg) != NUTS_OK) throw std::invalid_argument(nuts_last_error());

  const std::string out_file = val(R, {"output", "file"});
  const std::string data_file = val(R, {"data", "file"});
  if (!data_file.empty() && data_file == out_file)
    throw std::invalid_argument("Output file name must be different from the data file name");  // command_helper.hpp:650-685
  const int sig_figs = (int)to_int(val(R, {"output", "sig_figs"}), "sig_figs");
  if (sig_figs < -1 || sig_figs > 18) throw std::invalid_argument("sig_figs must be in [-1, 18]");

  // echo the configuration (command.hpp:153-156)
  std::ostringstream cfgtxt;
  print_tree(R, 0, cfgtxt, "");
  std::cout << cfgtxt.str() << std::endl;
  if (to_bool(val(R, {"output", "save_cmdstan_config"}), "save_cmdstan_config")) {
    // <output basename>_config.json (command.hpp:279-285)
    std::string base = out_file;
    const size_t dot = base.find_last_of('.'), sep = base.find_last_of('/');
    if (dot != std::string::npos && (sep == std::string::npos || dot > sep) && dot > 0) base = base.substr(0, dot);
    std::ofstream jf(base + "_config.json");
    if (!jf) throw std::invalid_argument("Can't open config file " + base + "_config.json");
    std::time_t now = std::time(nullptr);
    char tb[64];
    std::strftime(tb, sizeof tb, "%Y-%m-%d %H:%M:%S UTC", std::gmtime(&now));
    jf << "{\n  \"stan_major_version\": 2,\n  \"stan_minor_version\": 36,\n  \"stan_patch_version\": 0,\n"
       << "  \"model_name\": " << json_value(family + "_model") << ",\n  \"start_datetime\": " << json_value(tb)
       << ",\n";
    bool first = true;
    print_json_members(R, jf, 1, first);
    jf << ",\n  \"mpi_enabled\": false,\n  \"engine\": " << json_value(nuts_version()) << "\n}\n";
  }

  nuts_data* data = nuts_data_create();
  if (!data_file.empty()) load_var_context(data_file, data, "data");
  const nuts_data* dl[1] = {data};
  nuts_model* model = nuts_model_create(family.c_str(), dl, 1);
  if (!model) throw std::invalid_argument(nuts_last_error());
  int D = 0, nout = 0;
  nuts_model_dims(model, 0, &D, &nout);
  std::vector<std::string> names;
  char buf[256];
  for (int i = 0; i < nout; ++i) { nuts_model_param_name(model, 0, i, buf, sizeof buf); names.push_back(buf); }

  std::vector<double> inv_metric;
  if (!metric_file.empty()) {
    nuts_data* md = nuts_data_create();
    load_var_context(metric_file, md, "metric");
    int nd = 0;
    int64_t dims[8];
    const double* v = nullptr;
    if (nuts_data_get(md, "inv_metric", &nd, dims, &v) != NUTS_OK) throw std::invalid_argument(nuts_last_error());
    const int64_t need = cfg.metric == NUTS_METRIC_DENSE_E ? (int64_t)D * D : D;
    int64_t have = 1;
    for (int i = 0; i < nd; ++i) have *= dims[i];
    if (have != need) throw std::invalid_argument("metric_file: inv_metric has the wrong size");
    for (int k = 0; k < cfg.num_chains; ++k) inv_metric.insert(inv_metric.end(), v, v + need);
    nuts_data_destroy(md);
  }

  const int rows = nuts_output_rows(&cfg);
  const int cols = 7 + nout;
  const int C = cfg.num_chains;
  std::vector<double> draws((size_t)C * rows * cols), step(C);
  std::vector<double> imet((size_t)C * (cfg.metric == NUTS_METRIC_DENSE_E ? (size_t)D * D : D));
  std::vector<int64_t> nlf(C), nge(C);
  std::vector<int> status(C);
  nuts_output out;
  std::memset(&out, 0, sizeof out);
  out.draws = draws.data(); out.stepsize = step.data(); out.inv_metric = imet.data();
  out.n_leapfrog = nlf.data(); out.grad_evals = nge.data(); out.status = status.data();
  // diagnostic_file: q, p, g of every saved draw (mcmc_writer::write_diagnostic_params)
  const std::string diag_file = val(R, {"output", "diagnostic_file"});
  const int dcols = 7 + 3 * D;
  std::vector<double> diag;
  if (!diag_file.empty()) {
    diag.resize((size_t)C * rows * dcols);
    out.diagnostics = diag.data();
  }
  std::cout << "Running " << C << " chain(s) of NUTS on the GPU (" << nuts_version() << ")" << std::endl;
  // init=<file>: per-chain files {name}_{id+i}{ext} when the first exists, else
  // the named file for every chain (command_helper.hpp:144-245)
  std::vector<double> inits;
  if (init_is_file) {
    // get_vec_var_context (command_helper.hpp:144-245): one chain reads the file
    // as given; several chains need a .json or .R ending and use per-chain files
    // {name}_{id+i}{ext} when the first exists, else the named file for all
    std::vector<std::string> files(cfg.num_chains, init);
    if (cfg.num_chains > 1) {
      const size_t dot = init.find_last_of('.');
      if (dot == std::string::npos)
        throw std::invalid_argument("Found: \"" + init + "\" but user specified files must end in .json or .R");
      const std::string ext = init.substr(dot);
      if (ext != ".json" && ext != ".R")
        throw std::invalid_argument("file ending of " + ext + " is not supported by cmdstan");
      const auto per = make_filenames(init.substr(0, dot), "", ext, cfg.num_chains, cfg.id);
      if (std::ifstream(per[0]).good()) {
        files = per;
        for (int c = 1; c < cfg.num_chains; ++c)
          if (!std::ifstream(per[c]).good())
            throw std::invalid_argument("Found \"" + per[0] + "\" but cannot open \"" + per[c] + "\"");
      } else if (!std::ifstream(init).good()) {
        throw std::invalid_argument("Can't open either of specified files,\"" + per[0] + "\" and base file \"" +
                                    init + "\"");
      }
    }
    inits.resize((size_t)cfg.num_chains * D);
    for (int c = 0; c < cfg.num_chains; ++c) {
      if (!std::ifstream(files[c]).good()) throw std::invalid_argument("Can't open init file " + files[c]);
      nuts_data* idata = nuts_data_create();
      load_var_context(files[c], idata, "init");
      const bool ok = nuts_model_transform_inits(model, 0, idata, inits.data() + (size_t)c * D) == NUTS_OK;
      nuts_data_destroy(idata);
      if (!ok) throw std::invalid_argument(nuts_last_error());
    }
  }
  std::vector<int> rejections((size_t)C * 2, 0);
  out.init_rejections = rejections.data();
  const int rc = nuts_sample(model, &cfg, inits.empty() ? nullptr : inits.data(),
                             inv_metric.empty() ? nullptr : inv_metric.data(), &out);
  // services::util::initialize's messages, one block per rejected initial value
  for (int k = 0; k < C; ++k) {
    const std::string pre = C > 1 ? "Chain [" + std::to_string(cfg.id + k) + "] " : "";
    for (int r = 0; r < rejections[2 * k]; ++r)
      std::cout << pre << "Rejecting initial value:\n" << pre
                << "  Log probability evaluates to log(0), i.e. negative infinity.\n" << pre
                << "  Stan can't start sampling from this initial value.\n";
    for (int r = 0; r < rejections[2 * k + 1]; ++r)
      std::cout << pre << "Rejecting initial value:\n" << pre
                << "  Gradient evaluated at the initial value is not finite.\n" << pre
                << "  Stan can't start sampling from this initial value.\n";
    if (status[k] == 1 && !init_is_file && cfg.init_radius > 0)
      std::cout << pre << "\n" << pre << "Initialization between (-" << cfg.init_radius << ", " << cfg.init_radius
                << ") failed after 100 attempts. \n" << pre
                << " Try specifying initial values, reducing ranges of constrained values, or reparameterizing the "
                   "model.\n";
  }
  std::cout << std::flush;
  if (rc != NUTS_OK) {
    std::cerr << nuts_last_error() << std::endl;
    return 1;
  }

  const auto files = make_filenames(out_file, "", ".csv", C, cfg.id);
  const auto mfiles = make_filenames(out_file, "_metric", ".json", C, cfg.id);
  for (int k = 0; k < C; ++k) {
    std::ofstream f(files[k]);
    if (!f) throw std::invalid_argument("Can't open output file " + files[k]);
    if (sig_figs > -1) f << std::setprecision(sig_figs);
    f << "# stan_version_major = 2\n# stan_version_minor = 36\n# stan_version_patch = 0\n";
    f << "# model = " << family << "_model\n";
    std::time_t now = std::time(nullptr);
    char tb[64];
    std::strftime(tb, sizeof tb, "%Y-%m-%d %H:%M:%S UTC", std::gmtime(&now));
    f << "# start_datetime = " << tb << "\n";
    std::istringstream ct(cfgtxt.str());
    std::string line;
    while (std::getline(ct, line)) f << "# " << line << "\n";
    f << "# engine = " << nuts_version() << "\n";
    f << "lp__,accept_stat__,stepsize__,treedepth__,n_leapfrog__,divergent__,energy__";
    for (auto& n : names) f << "," << n;
    f << "\n";
    const double* d = draws.data() + (size_t)k * rows * cols;
    const int warm_rows = cfg.save_warmup ? (cfg.num_warmup + cfg.thin - 1) / cfg.thin : 0;
    for (int r = 0; r < warm_rows; ++r) fmt_row(f, d + (size_t)r * cols, cols);
    // write_adapt_finish / write_sampler_state
    if (cfg.adapt_engaged) f << "# Adaptation terminated\n";
    f << "# Step size = " << step[k] << "\n";
    if (cfg.metric == NUTS_METRIC_DENSE_E) {
      f << "# Elements of inverse mass matrix:\n";
      const double* m = imet.data() + (size_t)k * D * D;
      for (int i = 0; i < D; ++i) {
        f << "# ";
        for (int j = 0; j < D; ++j) f << (j ? ", " : "") << m[(size_t)i * D + j];
        f << "\n";
      }
    } else if (cfg.metric == NUTS_METRIC_DIAG_E) {
      f << "# Diagonal elements of inverse mass matrix:\n# ";
      const double* m = imet.data() + (size_t)k * D;
      for (int i = 0; i < D; ++i) f << (i ? ", " : "") << m[i];
      f << "\n";
    }
    for (int r = warm_rows; r < rows; ++r) fmt_row(f, d + (size_t)r * cols, cols);
    const double tw = out.warmup_seconds, ts = out.sampling_seconds;
    f << "# \n#  Elapsed Time: " << tw << " seconds (Warm-up)\n#                " << ts
      << " seconds (Sampling)\n#                " << tw + ts << " seconds (Total)\n# \n";
    if (save_metric) {
      std::ofstream mf(mfiles[k]);
      mf << std::setprecision(17) << "{\n  \"stepsize\" : " << step[k] << ",\n  \"metric_type\" : \"" << metric
         << "\",\n  \"inv_metric\" : [";
      if (cfg.metric == NUTS_METRIC_DENSE_E) {
        const double* m = imet.data() + (size_t)k * D * D;
        for (int i = 0; i < D; ++i) {
          mf << (i ? ", " : "") << "[";
          for (int j = 0; j < D; ++j) mf << (j ? ", " : "") << m[(size_t)i * D + j];
          mf << "]";
        }
      } else {
        const double* m = imet.data() + (size_t)k * D;
        for (int i = 0; i < D; ++i) mf << (i ? ", " : "") << m[i];
      }
      mf << "]\n}\n";
    }
  }
  if (!diag_file.empty()) {
    std::vector<std::string> un;
    for (int i = 0; i < D; ++i) { nuts_model_unconstrained_name(model, 0, i, buf, sizeof buf); un.push_back(buf); }
    const auto dfiles = make_filenames(diag_file, "", ".csv", C, cfg.id);
    for (int k = 0; k < C; ++k) {
      std::ofstream f(dfiles[k]);
      if (!f) throw std::invalid_argument("Can't open diagnostic file " + dfiles[k]);
      if (sig_figs > -1) f << std::setprecision(sig_figs);
      std::istringstream ct(cfgtxt.str());
      std::string line;
      while (std::getline(ct, line)) f << "# " << line << "\n";
      f << "lp__,accept_stat__,stepsize__,treedepth__,n_leapfrog__,divergent__,energy__";
      for (auto& n : un) f << "," << n;
      for (auto& n : un) f << ",p_" << n;
      for (auto& n : un) f << ",g_" << n;
      f << "\n";
      const double* d = diag.data() + (size_t)k * rows * dcols;
      for (int r = 0; r < rows; ++r) fmt_row(f, d + (size_t)r * dcols, dcols);
    }
  }
  std::cout << "Elapsed Time: " << out.warmup_seconds << " seconds (Warm-up), " << out.sampling_seconds
            << " seconds (Sampling)" << std::endl;
  nuts_model_destroy(model);
  nuts_data_destroy(data);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
}
