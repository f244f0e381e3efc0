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
  if (method != "sample")
    throw std::invalid_argument("method=" + method + " is not provided by this build; only method=sample runs here");
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
    throw std::invalid_argument("Cannot save adapted metric if adaptation is disabled.");  // command.hpp:249-259
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
  if (!data_file.empty() && nuts_data_parse_json(data, read_file(data_file).c_str()) != NUTS_OK)
    throw std::invalid_argument(std::string("Error reading data file: ") + nuts_last_error());
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
    if (nuts_data_parse_json(md, read_file(metric_file).c_str()) != NUTS_OK)
      throw std::invalid_argument(std::string("Error reading metric file: ") + nuts_last_error());
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
    const size_t dot = init.find_last_of('.');
    if (dot == std::string::npos || init.substr(dot) != ".json")
      throw std::invalid_argument("Found: \"" + init + "\" but user specified files must end in .json");
    std::vector<std::string> files(cfg.num_chains, init);
    if (cfg.num_chains > 1) {
      const auto per = make_filenames(init.substr(0, dot), "", ".json", cfg.num_chains, cfg.id);
      if (std::ifstream(per[0]).good()) files = per;
      else if (!std::ifstream(init).good())
        throw std::invalid_argument("Can't open either of specified files, \"" + per[0] + "\" and base file \"" +
                                    init + "\"");
    }
    inits.resize((size_t)cfg.num_chains * D);
    for (int c = 0; c < cfg.num_chains; ++c) {
      if (!std::ifstream(files[c]).good()) throw std::invalid_argument("Can't open init file " + files[c]);
      nuts_data* idata = nuts_data_create();
      const bool ok = nuts_data_parse_json(idata, read_file(files[c]).c_str()) == NUTS_OK &&
                      nuts_model_transform_inits(model, 0, idata, inits.data() + (size_t)c * D) == NUTS_OK;
      nuts_data_destroy(idata);
      if (!ok) throw std::invalid_argument(nuts_last_error());
    }
  }
  const int rc = nuts_sample(model, &cfg, inits.empty() ? nullptr : inits.data(),
                             inv_metric.empty() ? nullptr : inv_metric.data(), &out);
  if (rc != NUTS_OK) {
    std::cerr << nuts_last_error() << std::endl;
    for (int k = 0; k < C; ++k)
      if (status[k] == 1) std::cerr << "Chain [" << (cfg.id + k) << "] Rejecting initial value: initialization failed." << std::endl;
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
