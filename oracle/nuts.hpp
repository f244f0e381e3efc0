// ORACLE — test infrastructure only (tests/, smoke check, bench cpu_baseline).
//
// FP64 CPU restatement of Stan's NUTS sampling path, written literally (recursive
// tree, Eigen-free loops) so that it can check the batched HIP implementation.
// The upstream code (stan-dev/stan 2.36, src/stan/mcmc/ and src/stan/services/)
// is not in the reference tree (the `stan` submodule is empty, .gitmodules:1-3);
// this follows the published algorithm (Hoffman & Gelman 2014; Betancourt 2017,
// arXiv:1701.02434) and Stan's documented defaults, entered through
// cmdstan's call at src/cmdstan/command.hpp:603-697.  Pieces that no reference
// fixture can pin are marked "per Stan docs, unpinned here".
//
//   base_nuts::transition / build_tree ........ NutsSampler::transition / build_tree
//   expl_leapfrog::evolve ..................... NutsSampler::evolve
//   {unit,diag,dense}_e_metric ................ Hamiltonian
//   base_hmc::init_stepsize / sample_stepsize . NutsSampler::init_stepsize / sample_stepsize
//   stepsize_adaptation ....................... StepsizeAdaptation
//   windowed_adaptation + var/covar_adaptation  WindowedAdaptation + Welford{Var,Covar}
//   services::util::initialize ................ initialize()
//   services::util::run_{adaptive_,}sampler ... run_chain()
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "models.hpp"
#include "philox.hpp"

namespace oracle {

using vec = std::vector<double>;
static const double INF = std::numeric_limits<double>::infinity();

// --------------------------------------------------------------- draws
struct Draws {
  virtual ~Draws() {}
  virtual double uniform() = 0;
  virtual void normals(int n, double* z) = 0;
  virtual double init_uniform(int try_idx, int i, int D) = 0;
};

// The project's per-chain Philox streams (philox.hpp).
struct PhiloxDraws : Draws {
  uint64_t seed; uint32_t chain; uint64_t uctr = 0, nctr = 0;
  PhiloxDraws(uint64_t s, uint32_t c) : seed(s), chain(c) {}
  double uniform() override { return stream_uniform(seed, chain, PURPOSE_UNIFORM, uctr++); }
  void normals(int n, double* z) override {
    for (int i = 0; i < n; i += 2) {
      double a, b;
      stream_normal_pair(seed, chain, nctr + (uint64_t)(i / 2), &a, &b);
      z[i] = a;
      if (i + 1 < n) z[i + 1] = b;
    }
    nctr += (uint64_t)((n + 1) / 2);
  }
  double init_uniform(int t, int i, int D) override {
    return stream_uniform(seed, chain, PURPOSE_INIT, (uint64_t)t * (uint64_t)D + (uint64_t)i);
  }
};

// Injected draws for single-transition replay: consumed strictly in order.
struct InjectedDraws : Draws {
  const double* u; size_t nu; const double* z; size_t nz;
  size_t iu = 0, iz = 0;
  InjectedDraws(const double* u_, size_t nu_, const double* z_, size_t nz_) : u(u_), nu(nu_), z(z_), nz(nz_) {}
  double uniform() override {
    if (iu >= nu) throw std::runtime_error("injected uniforms exhausted");
    return u[iu++];
  }
  void normals(int n, double* out) override {
    if (iz + (size_t)n > nz) throw std::runtime_error("injected normals exhausted");
    for (int i = 0; i < n; ++i) out[i] = z[iz + i];
    iz += n;
  }
  double init_uniform(int, int, int) override { throw std::runtime_error("no init draws injected"); }
};

// ----------------------------------------------------------- utilities
// stan::math::log_sum_exp(double, double) and log1p_exp
inline double log_sum_exp(double a, double b) {
  if (a == -INF) return b;
  if (a == INF && b == INF) return INF;
  if (a > b) return a + log1p_exp(b - a);
  return b + log1p_exp(a - b);
}
inline double dot(const vec& a, const vec& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
inline bool all_finite(const vec& v) {
  for (double x : v) if (!std::isfinite(x)) return false;
  return true;
}

// ---------------------------------------------------------- hamiltonian
enum MetricKind { UNIT_E = 0, DIAG_E = 1, DENSE_E = 2 };

struct PS { vec q, p, g; double V = 0; };

struct Hamiltonian {
  const Model* m; int D; int kind;
  vec minv;  // diag: D, dense: D*D row-major, unit: unused
  int64_t grad_calls = 0;

  Hamiltonian(const Model* m_, int kind_) : m(m_), D(m_->dim()), kind(kind_) {
    if (kind == DIAG_E) minv.assign(D, 1.0);
    if (kind == DENSE_E) { minv.assign((size_t)D * D, 0.0); for (int i = 0; i < D; ++i) minv[(size_t)i * D + i] = 1.0; }
  }
  void dtau_dp(const vec& p, vec& out) const {
    out.resize(D);
    if (kind == UNIT_E) { out = p; return; }
    if (kind == DIAG_E) { for (int i = 0; i < D; ++i) out[i] = minv[i] * p[i]; return; }
    for (int i = 0; i < D; ++i) {
      double s = 0;
      for (int j = 0; j < D; ++j) s += minv[(size_t)i * D + j] * p[j];
      out[i] = s;
    }
  }
  double T(const PS& z) const {
    if (kind == UNIT_E) return 0.5 * dot(z.p, z.p);
    vec t; dtau_dp(z.p, t);
    return 0.5 * dot(z.p, t);
  }
  double H(const PS& z) const { return T(z) + z.V; }
  // Minv = L L^T; returns U = L^T row-major (dense_e_metric::sample_p uses llt().matrixU()).
  vec chol_upper() const {
    vec L((size_t)D * D, 0.0);
    for (int j = 0; j < D; ++j) {
      double s = minv[(size_t)j * D + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)j * D + k] * L[(size_t)j * D + k];
      if (!(s > 0)) throw std::domain_error("inverse metric not positive definite");
      double d = std::sqrt(s);
      L[(size_t)j * D + j] = d;
      for (int i = j + 1; i < D; ++i) {
        double t = minv[(size_t)i * D + j];
        for (int k = 0; k < j; ++k) t -= L[(size_t)i * D + k] * L[(size_t)j * D + k];
        L[(size_t)i * D + j] = t / d;
      }
    }
    vec U((size_t)D * D, 0.0);
    for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) U[(size_t)i * D + j] = L[(size_t)j * D + i];
    return U;
  }
  void sample_p(PS& z, Draws& rng) const {
    vec u(D);
    rng.normals(D, u.data());
    z.p.resize(D);
    if (kind == UNIT_E) { z.p = u; return; }
    if (kind == DIAG_E) { for (int i = 0; i < D; ++i) z.p[i] = u[i] / std::sqrt(minv[i]); return; }
    vec U = chol_upper();  // solve U p = u (back substitution)
    for (int i = D - 1; i >= 0; --i) {
      double s = u[i];
      for (int j = i + 1; j < D; ++j) s -= U[(size_t)i * D + j] * z.p[j];
      z.p[i] = s / U[(size_t)i * D + i];
    }
  }
  // base_hamiltonian::update_potential_gradient: V = -lp, g = -grad lp; a
  // throwing density gives V = +inf.
  void update_potential_gradient(PS& z) {
    ++grad_calls;
    z.g.resize(D);
    try {
      double lp = m->log_prob_grad(z.q.data(), z.g.data(), true);
      z.V = -lp;
      for (int i = 0; i < D; ++i) z.g[i] = -z.g[i];
    } catch (const std::exception&) {
      z.V = INF;
    }
  }
};

// ------------------------------------------------------------- NUTS
struct Sample { vec q; double lp; double accept_stat; };

struct NutsSampler {
  Hamiltonian ham;
  Draws* rng;
  PS z;
  double nom_eps = 1.0, eps = 1.0, jitter = 0.0;
  int max_depth = 10;
  double max_deltaH = 1000;  // per Stan docs (divergence threshold), unpinned here
  int depth = 0, n_leapfrog = 0; bool divergent = false; double energy = 0;
  int64_t total_leapfrogs = 0;

  NutsSampler(const Model* m, int kind, Draws* r) : ham(m, kind), rng(r) {
    int D = m->dim();
    z.q.assign(D, 0); z.p.assign(D, 0); z.g.assign(D, 0);
  }

  void sample_stepsize() {
    eps = nom_eps;
    if (jitter) eps *= 1.0 + jitter * (2.0 * rng->uniform() - 1.0);
  }

  // expl_leapfrog / base_leapfrog::evolve
  void evolve(PS& s, double e) {
    const int D = (int)s.q.size();
    const double h = 0.5 * e;
    for (int i = 0; i < D; ++i) s.p[i] -= h * s.g[i];
    vec t; ham.dtau_dp(s.p, t);
    for (int i = 0; i < D; ++i) s.q[i] += e * t[i];
    ham.update_potential_gradient(s);
    for (int i = 0; i < D; ++i) s.p[i] -= h * s.g[i];
  }

  static bool criterion(const vec& ps_minus, const vec& ps_plus, const vec& rho) {
    return dot(ps_plus, rho) > 0 && dot(ps_minus, rho) > 0;
  }

  bool build_tree(int d, PS& z_propose, vec& ps_beg, vec& ps_end, vec& rho, vec& p_beg, vec& p_end,
                  double H0, double sign, int& n_lf, double& log_sum_weight, double& sum_metro) {
    const int D = (int)z.q.size();
    if (d == 0) {
      evolve(z, sign * eps);
      ++n_lf;
      double h = ham.H(z);
      if (std::isnan(h)) h = INF;
      if ((h - H0) > max_deltaH) divergent = true;
      log_sum_weight = log_sum_exp(log_sum_weight, H0 - h);
      if (H0 - h > 0) sum_metro += 1;
      else sum_metro += std::exp(H0 - h);
      z_propose = z;
      ham.dtau_dp(z.p, ps_beg);
      ps_end = ps_beg;
      for (int i = 0; i < D; ++i) rho[i] += z.p[i];
      p_beg = z.p;
      p_end = p_beg;
      return !divergent;
    }
    double lsw_init = -INF;
    vec p_init_end(D), ps_init_end(D), rho_init(D, 0.0);
    bool valid_init = build_tree(d - 1, z_propose, ps_beg, ps_init_end, rho_init, p_beg, p_init_end,
                                 H0, sign, n_lf, lsw_init, sum_metro);
    if (!valid_init) return false;
    PS z_propose_final(z);
    double lsw_final = -INF;
    vec p_final_beg(D), ps_final_beg(D), rho_final(D, 0.0);
    bool valid_final = build_tree(d - 1, z_propose_final, ps_final_beg, ps_end, rho_final, p_final_beg, p_end,
                                  H0, sign, n_lf, lsw_final, sum_metro);
    if (!valid_final) return false;
    double lsw_sub = log_sum_exp(lsw_init, lsw_final);
    log_sum_weight = log_sum_exp(log_sum_weight, lsw_sub);
    if (lsw_final > lsw_sub) {
      z_propose = z_propose_final;
    } else {
      double ap = std::exp(lsw_final - lsw_sub);
      if (rng->uniform() < ap) z_propose = z_propose_final;
    }
    vec rho_sub(D);
    for (int i = 0; i < D; ++i) { rho_sub[i] = rho_init[i] + rho_final[i]; rho[i] += rho_sub[i]; }
    bool persist = criterion(ps_beg, ps_end, rho_sub);
    vec rho_ext(D);
    for (int i = 0; i < D; ++i) rho_ext[i] = rho_init[i] + p_final_beg[i];
    persist &= criterion(ps_beg, ps_final_beg, rho_ext);
    for (int i = 0; i < D; ++i) rho_ext[i] = rho_final[i] + p_init_end[i];
    persist &= criterion(ps_init_end, ps_end, rho_ext);
    return persist;
  }

  // base_nuts::transition
  Sample transition(const vec& q_init) {
    const int D = (int)q_init.size();
    sample_stepsize();
    z.q = q_init;
    ham.sample_p(z, *rng);
    ham.update_potential_gradient(z);  // hamiltonian.init(z): recomputes V, g at q

    PS z_fwd(z), z_bck(z), z_sample(z), z_propose(z);
    vec p_fwd_fwd = z.p; vec ps_fwd_fwd; ham.dtau_dp(z.p, ps_fwd_fwd);
    vec p_fwd_bck = z.p; vec ps_fwd_bck = ps_fwd_fwd;
    vec p_bck_fwd = z.p; vec ps_bck_fwd = ps_fwd_fwd;
    vec p_bck_bck = z.p; vec ps_bck_bck = ps_fwd_fwd;
    vec rho = z.p;
    double log_sum_weight = 0;
    double H0 = ham.H(z);
    int n_lf = 0;
    double sum_metro = 0;
    depth = 0;
    divergent = false;

    while (depth < max_depth) {
      vec rho_fwd(D, 0.0), rho_bck(D, 0.0);
      bool valid = false;
      double lsw_sub = -INF;
      if (rng->uniform() > 0.5) {
        z = z_fwd;
        rho_bck = rho;
        p_bck_fwd = p_fwd_bck;
        ps_bck_fwd = ps_fwd_bck;
        valid = build_tree(depth, z_propose, ps_fwd_bck, ps_fwd_fwd, rho_fwd, p_fwd_bck, p_fwd_fwd,
                           H0, 1, n_lf, lsw_sub, sum_metro);
        z_fwd = z;
      } else {
        z = z_bck;
        rho_fwd = rho;
        p_fwd_bck = p_bck_fwd;
        ps_fwd_bck = ps_bck_fwd;
        valid = build_tree(depth, z_propose, ps_bck_fwd, ps_bck_bck, rho_bck, p_bck_fwd, p_bck_bck,
                           H0, -1, n_lf, lsw_sub, sum_metro);
        z_bck = z;
      }
      if (!valid) break;
      ++depth;
      if (lsw_sub > log_sum_weight) {
        z_sample = z_propose;
      } else {
        double ap = std::exp(lsw_sub - log_sum_weight);
        if (rng->uniform() < ap) z_sample = z_propose;
      }
      log_sum_weight = log_sum_exp(log_sum_weight, lsw_sub);
      for (int i = 0; i < D; ++i) rho[i] = rho_bck[i] + rho_fwd[i];
      bool persist = criterion(ps_bck_bck, ps_fwd_fwd, rho);
      vec rho_ext(D);
      for (int i = 0; i < D; ++i) rho_ext[i] = rho_bck[i] + p_fwd_bck[i];
      persist &= criterion(ps_bck_bck, ps_fwd_bck, rho_ext);
      for (int i = 0; i < D; ++i) rho_ext[i] = rho_fwd[i] + p_bck_fwd[i];
      persist &= criterion(ps_bck_fwd, ps_fwd_fwd, rho_ext);
      if (!persist) break;
    }
    n_leapfrog = n_lf;
    total_leapfrogs += n_lf;
    double accept = sum_metro / (double)n_lf;
    z = z_sample;
    energy = ham.H(z);
    return Sample{z.q, -z.V, accept};
  }

  // base_hmc::init_stepsize (loop limits 1e7 / 0 per Stan docs, unpinned here)
  void init_stepsize() {
    PS z_init(z);
    if (nom_eps == 0 || nom_eps > 1e7 || std::isnan(nom_eps)) return;
    ham.sample_p(z, *rng);
    ham.update_potential_gradient(z);
    double H0 = ham.H(z);
    evolve(z, nom_eps);
    double h = ham.H(z);
    if (std::isnan(h)) h = INF;
    double delta_H = H0 - h;
    int direction = delta_H > std::log(0.8) ? 1 : -1;
    while (true) {
      z = z_init;
      ham.sample_p(z, *rng);
      ham.update_potential_gradient(z);
      double H0b = ham.H(z);
      evolve(z, nom_eps);
      double hb = ham.H(z);
      if (std::isnan(hb)) hb = INF;
      double dH = H0b - hb;
      if ((direction == 1) && !(dH > std::log(0.8))) break;
      else if ((direction == -1) && !(dH < std::log(0.8))) break;
      else direction == 1 ? nom_eps *= 2 : nom_eps *= 0.5;
      if (nom_eps > 1e7) throw std::runtime_error("Posterior is improper. Please check your model.");
      if (nom_eps == 0) throw std::runtime_error("No acceptably small step size could be found. Perhaps the posterior is not continuous?");
    }
    z = z_init;
  }
};

// ------------------------------------------------------- adaptation
// stepsize_adaptation: Nesterov dual averaging.
struct StepsizeAdaptation {
  double mu = 0.5, delta = 0.5, gamma = 0.05, kappa = 0.75, t0 = 10;
  double counter = 0, s_bar = 0, x_bar = 0;
  void restart() { counter = 0; s_bar = 0; x_bar = 0; }
  void learn_stepsize(double& epsilon, double adapt_stat) {
    ++counter;
    adapt_stat = adapt_stat > 1 ? 1 : adapt_stat;
    const double eta = 1.0 / (counter + t0);
    s_bar = (1.0 - eta) * s_bar + eta * (delta - adapt_stat);
    const double x = mu - s_bar * std::sqrt(counter) / gamma;
    const double x_eta = std::pow(counter, -kappa);
    x_bar = (1.0 - x_eta) * x_bar + x_eta * x;
    epsilon = std::exp(x);
  }
  void complete_adaptation(double& epsilon) { epsilon = std::exp(x_bar); }
};

// windowed_adaptation (unsigned arithmetic as upstream).  The 15%/75%/10%
// branch sets the buffers but does not restart the window schedule, so the
// first window end stays at UINT_MAX and the metric is never updated; the
// reference pins that outcome: src/test/interface/metric_test.cpp:17-92 expects
// the metric_file values back after adapt engaged=1 num_warmup=100.
struct WindowedAdaptation {
  unsigned num_warmup = 0, init_buffer = 0, term_buffer = 0, base_window = 0;
  unsigned counter = 0, window_size = 0, next_window = 0;
  WindowedAdaptation() { restart(); }
  void restart() {
    counter = 0;
    window_size = base_window;
    next_window = init_buffer + window_size - 1;
  }
  // returns 0 normal, 1 = "num_warmup < 20" (no estimation), 2 = 15/75/10 fallback
  int set_window_params(unsigned nw, unsigned ib, unsigned tb, unsigned bw) {
    if (nw < 20) return 1;
    if (ib + bw + tb > nw) {
      num_warmup = nw;
      init_buffer = (unsigned)(0.15 * nw);
      term_buffer = (unsigned)(0.1 * nw);
      base_window = nw - (init_buffer + term_buffer);
      return 2;
    }
    num_warmup = nw; init_buffer = ib; term_buffer = tb; base_window = bw;
    restart();
    return 0;
  }
  bool adaptation_window() const {
    return (counter >= init_buffer) && (counter < num_warmup - term_buffer) && (counter != num_warmup);
  }
  bool end_adaptation_window() const { return (counter == next_window) && (counter != num_warmup); }
  void compute_next_window() {
    if (next_window == num_warmup - term_buffer - 1) return;
    window_size *= 2;
    next_window = counter + window_size;
    if (next_window != num_warmup - term_buffer - 1) {
      unsigned boundary = next_window + 2 * window_size;
      if (boundary >= num_warmup - term_buffer) next_window = num_warmup - term_buffer - 1;
    }
  }
};

struct WelfordVar {
  int n = 0; vec m, m2;
  explicit WelfordVar(int D = 0) : m(D, 0.0), m2(D, 0.0) {}
  void restart() { n = 0; std::fill(m.begin(), m.end(), 0.0); std::fill(m2.begin(), m2.end(), 0.0); }
  void add_sample(const vec& q) {
    ++n;
    for (size_t i = 0; i < q.size(); ++i) {
      double delta = q[i] - m[i];
      m[i] += delta / n;
      m2[i] += (q[i] - m[i]) * delta;
    }
  }
  void sample_variance(vec& var) const { if (n > 1) for (size_t i = 0; i < m.size(); ++i) var[i] = m2[i] / (n - 1.0); }
};

struct WelfordCovar {
  int n = 0; int D; vec m, m2;
  explicit WelfordCovar(int D_ = 0) : D(D_), m(D_, 0.0), m2((size_t)D_ * D_, 0.0) {}
  void restart() { n = 0; std::fill(m.begin(), m.end(), 0.0); std::fill(m2.begin(), m2.end(), 0.0); }
  void add_sample(const vec& q) {
    ++n;
    vec delta(D);
    for (int i = 0; i < D; ++i) { delta[i] = q[i] - m[i]; m[i] += delta[i] / n; }
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) m2[(size_t)i * D + j] += (q[i] - m[i]) * delta[j];
  }
  void sample_covariance(vec& cov) const { if (n > 1) for (size_t i = 0; i < m2.size(); ++i) cov[i] = m2[i] / (n - 1.0); }
};

// adapt_{unit,diag,dense}_e_nuts::transition
struct AdaptiveNuts : NutsSampler {
  bool adapt_flag = false;
  StepsizeAdaptation sa;
  WindowedAdaptation wa;
  WelfordVar wv; WelfordCovar wc;
  int metric_updates = 0;

  AdaptiveNuts(const Model* m, int kind, Draws* r) : NutsSampler(m, kind, r), wv(m->dim()), wc(kind == DENSE_E ? m->dim() : 0) {}

  bool learn_metric(const vec& q) {
    if (ham.kind == UNIT_E) return false;
    if (wa.adaptation_window()) { if (ham.kind == DIAG_E) wv.add_sample(q); else wc.add_sample(q); }
    if (wa.end_adaptation_window()) {
      wa.compute_next_window();
      const int D = ham.D;
      if (ham.kind == DIAG_E) {
        wv.sample_variance(ham.minv);
        double n = (double)wv.n;
        for (int i = 0; i < D; ++i) ham.minv[i] = (n / (n + 5.0)) * ham.minv[i] + 1e-3 * (5.0 / (n + 5.0));
        if (!all_finite(ham.minv)) throw std::runtime_error("Numerical overflow in metric adaptation.");
        wv.restart();
      } else {
        wc.sample_covariance(ham.minv);
        double n = (double)wc.n;
        for (int i = 0; i < D; ++i)
          for (int j = 0; j < D; ++j)
            ham.minv[(size_t)i * D + j] = (n / (n + 5.0)) * ham.minv[(size_t)i * D + j] + (i == j ? 1e-3 * (5.0 / (n + 5.0)) : 0.0);
        if (!all_finite(ham.minv)) throw std::runtime_error("Numerical overflow in metric adaptation.");
        wc.restart();
      }
      ++wa.counter;
      return true;
    }
    ++wa.counter;
    return false;
  }

  Sample adapt_transition(const vec& q) {
    Sample s = transition(q);
    if (adapt_flag) {
      sa.learn_stepsize(nom_eps, s.accept_stat);
      bool update = learn_metric(z.q);
      if (update) {
        ++metric_updates;
        init_stepsize();
        sa.mu = std::log(10 * nom_eps);
        sa.restart();
      }
    }
    return s;
  }
};

// ----------------------------------------------------------- services
struct RunConfig {
  int num_warmup = 1000, num_samples = 1000, thin = 1;
  bool save_warmup = false, adapt = true;
  double delta = 0.8, gamma = 0.05, kappa = 0.75, t0 = 10;
  int init_buffer = 75, term_buffer = 50, window = 25;
  double stepsize = 1, jitter = 0;
  int max_depth = 10, metric = DIAG_E;
  double init_radius = 2;
};

struct RunResult {
  int rows = 0, cols = 0;
  vec draws;                // rows x cols: 7 sampler columns + model outputs
  double stepsize = 0;
  vec inv_metric;
  int64_t grad_evals = 0, leapfrogs = 0;
  int init_tries = 0;
  std::string error;
};

// services::util::initialize: up to 100 random draws in U(-R, R) on the
// unconstrained scale until lp and its gradient are finite (retry count per
// Stan docs; rejection behaviour in src/test/interface/command_test.cpp:151-250).
inline bool initialize(const Model& m, Draws& rng, double R, const double* user_init, vec& q, int& tries,
                       int64_t& grad_evals, std::string& err) {
  const int D = m.dim();
  q.assign(D, 0.0);
  vec g(D);
  const int max_tries = (user_init || R == 0) ? 1 : 100;
  for (int t = 0; t < max_tries; ++t) {
    tries = t + 1;
    if (user_init) for (int i = 0; i < D; ++i) q[i] = user_init[i];
    else if (R == 0) for (int i = 0; i < D; ++i) q[i] = 0;
    else for (int i = 0; i < D; ++i) q[i] = R * (2.0 * rng.init_uniform(t, i, D) - 1.0);
    double lp;
    ++grad_evals;
    try { lp = m.log_prob_grad(q.data(), g.data(), true); } catch (const std::exception&) { continue; }
    if (!std::isfinite(lp)) continue;
    if (!all_finite(g)) continue;
    return true;
  }
  err = "Initialization failed.";
  return false;
}

inline void write_row(const Model& m, const NutsSampler& s, const Sample& smp, RunResult& r) {
  r.draws.push_back(smp.lp);
  r.draws.push_back(smp.accept_stat);
  r.draws.push_back(s.eps);
  r.draws.push_back(s.depth);
  r.draws.push_back(s.n_leapfrog);
  r.draws.push_back(s.divergent ? 1 : 0);
  r.draws.push_back(s.energy);
  vec out(m.num_out());
  m.write_array(smp.q.data(), out.data());
  for (double v : out) r.draws.push_back(v);
  ++r.rows;
}

// run_adaptive_sampler / run_sampler for one chain.  `inv_metric_init` may be
// null (unit/identity default) — the metric_file path.
inline RunResult run_chain(const Model& m, const RunConfig& c, Draws& rng, const double* user_init,
                           const double* inv_metric_init) {
  RunResult r;
  r.cols = 7 + m.num_out();
  AdaptiveNuts s(&m, c.metric, &rng);
  s.nom_eps = c.stepsize;
  s.jitter = c.jitter;
  s.max_depth = c.max_depth;
  const int D = m.dim();
  if (inv_metric_init && c.metric != UNIT_E)
    for (size_t i = 0; i < s.ham.minv.size(); ++i) s.ham.minv[i] = inv_metric_init[i];
  vec q;
  if (!initialize(m, rng, c.init_radius, user_init, q, r.init_tries, r.grad_evals, r.error)) return r;
  if (c.adapt) {
    s.sa.mu = std::log(10 * c.stepsize);
    s.sa.delta = c.delta; s.sa.gamma = c.gamma; s.sa.kappa = c.kappa; s.sa.t0 = c.t0;
    s.wa.set_window_params(c.num_warmup, c.init_buffer, c.term_buffer, c.window);
    s.adapt_flag = true;
    s.z.q = q;
    try { s.init_stepsize(); } catch (const std::exception& e) { r.error = e.what(); return r; }
  }
  Sample smp{q, 0, 0};
  try {
    for (int it = 0; it < c.num_warmup; ++it) {
      smp = s.adapt_transition(smp.q);
      if (c.save_warmup && it % c.thin == 0) write_row(m, s, smp, r);
    }
    if (c.adapt) { s.adapt_flag = false; s.sa.complete_adaptation(s.nom_eps); }
    for (int it = 0; it < c.num_samples; ++it) {
      smp = s.adapt_transition(smp.q);
      if (it % c.thin == 0) write_row(m, s, smp, r);
    }
  } catch (const std::exception& e) {
    r.error = e.what();
    return r;
  }
  r.stepsize = s.nom_eps;
  r.inv_metric = s.ham.minv;
  r.grad_evals += s.ham.grad_calls;
  r.leapfrogs = s.total_leapfrogs;
  (void)D;
  return r;
}

}  // namespace oracle
