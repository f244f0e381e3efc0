// ORACLE — test infrastructure only (tests/, smoke check, bench cpu_baseline).
//
// FP64 restatement of the log densities the sampler targets, in Stan's
// `log_prob<propto=true, jacobian>` convention: constants that do not depend on
// parameters are dropped, and the log-Jacobian of each constraining transform is
// added when `jacobian` is set (cmdstan command_helper.hpp:569-600 calls
// stan::model::log_prob_grad<true, jacobian>; the fixture lp__ columns, e.g.
// src/test/interface/example_output/bernoulli_chain_1.csv:44-1043, confirm that
// the sampler's lp__ is the propto + Jacobian density).
//
// Every density is written once as a template over the scalar type so the
// complex-step derivative Im f(q + i h e_k) / h can check the hand-derived
// gradients to machine precision.
//
// Transforms (Stan reference manual, "Constraint transforms"):
//   lower=0        x = exp(u),          log|J| = u
//   lower=0,upper=1 x = inv_logit(u),    log|J| = log x + log(1-x)
#pragma once
#include <cmath>
#include <complex>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;

inline double re(double x) { return x; }
inline double re(const cplx& x) { return x.real(); }

// log(1 + exp(x)), branch on the real part (analytic on each branch).
template <class T> inline T log1p_exp(const T& x) {
  using std::exp; using std::log;
  if (re(x) > 0) return x + log(T(1.0) + exp(-x));
  return log(T(1.0) + exp(x));
}
inline double log1p_exp(double x) {
  return x > 0 ? x + std::log1p(std::exp(-x)) : std::log1p(std::exp(x));
}
inline double inv_logit(double x) {
  if (x >= 0) return 1.0 / (1.0 + std::exp(-x));
  double e = std::exp(x);
  return e / (1.0 + e);
}

// Named real arrays with dims, like stan::io::var_context.
struct Data {
  struct Field { std::vector<int64_t> dims; std::vector<double> v; };
  std::map<std::string, Field> f;
  bool has(const std::string& n) const { return f.count(n) != 0; }
  const std::vector<double>& vec(const std::string& n) const {
    auto it = f.find(n);
    if (it == f.end()) throw std::invalid_argument("missing data field: " + n);
    return it->second.v;
  }
  double scalar(const std::string& n) const { return vec(n).at(0); }
  double scalar_or(const std::string& n, double d) const { return has(n) ? vec(n).at(0) : d; }
  int integer(const std::string& n) const { return (int)std::llround(scalar(n)); }
};

struct Model {
  virtual ~Model() {}
  virtual std::string name() const = 0;
  virtual int dim() const = 0;                 // unconstrained dimension
  virtual int num_out() const = 0;             // constrained + transformed outputs
  virtual std::vector<std::string> out_names() const = 0;
  // returns lp, writes gradient of lp (not of V) into g
  virtual double log_prob_grad(const double* q, double* g, bool jacobian) const = 0;
  virtual cplx log_prob_c(const cplx* q, bool jacobian) const = 0;
  virtual double log_prob(const double* q, bool jacobian) const = 0;
  virtual void write_array(const double* q, double* out) const = 0;
};

// ---------------------------------------------------------------- bernoulli
// examples/bernoulli/bernoulli.stan:1-11: theta ~ beta(1,1); y ~ bernoulli(theta).
struct Bernoulli : Model {
  int N; double s;  // number of ones
  explicit Bernoulli(const Data& d) {
    N = d.integer("N");
    const auto& y = d.vec("y");
    if ((int)y.size() != N) throw std::invalid_argument("bernoulli: y size != N");
    s = 0; for (double v : y) { if (v != 0 && v != 1) throw std::domain_error("bernoulli: y not in {0,1}"); s += v; }
  }
  std::string name() const override { return "bernoulli"; }
  int dim() const override { return 1; }
  int num_out() const override { return 1; }
  std::vector<std::string> out_names() const override { return {"theta"}; }
  template <class T> T lp(const T* q, bool jac) const {
    T u = q[0];
    T log_t = -log1p_exp(-u), log_1mt = -log1p_exp(u);
    T r = s * log_t + (N - s) * log_1mt;
    if (jac) r += log_t + log_1mt;
    return r;
  }
  double log_prob(const double* q, bool jac) const override { return lp(q, jac); }
  cplx log_prob_c(const cplx* q, bool jac) const override { return lp(q, jac); }
  double log_prob_grad(const double* q, double* g, bool jac) const override {
    double t = inv_logit(q[0]);
    double j = jac ? 1.0 : 0.0;
    g[0] = (s + j) * (1.0 - t) - (N - s + j) * t;
    return lp(q, jac);
  }
  void write_array(const double* q, double* out) const override { out[0] = inv_logit(q[0]); }
};

// ------------------------------------------------------- logistic regression
// y_n ~ bernoulli_logit(x_n . beta), beta ~ normal(0, prior_scale).
// X row-major N x D.
struct LogisticRegression : Model {
  int N, D; std::vector<double> X, y; double s2inv;
  explicit LogisticRegression(const Data& d) {
    N = d.integer("N"); D = d.integer("D");
    X = d.vec("X"); y = d.vec("y");
    double sc = d.scalar_or("prior_scale", 2.5);
    s2inv = 1.0 / (sc * sc);
    if ((int64_t)X.size() != (int64_t)N * D || (int)y.size() != N)
      throw std::invalid_argument("logistic_regression: bad data sizes");
  }
  std::string name() const override { return "logistic_regression"; }
  int dim() const override { return D; }
  int num_out() const override { return D; }
  std::vector<std::string> out_names() const override {
    std::vector<std::string> r; for (int i = 0; i < D; ++i) r.push_back("beta." + std::to_string(i + 1)); return r;
  }
  template <class T> T lp(const T* q, bool) const {
    T r = 0;
    for (int n = 0; n < N; ++n) {
      T eta = 0;
      const double* x = &X[(size_t)n * D];
      for (int k = 0; k < D; ++k) eta += x[k] * q[k];
      r += y[n] * eta - log1p_exp(eta);
    }
    T pr = 0;
    for (int k = 0; k < D; ++k) pr += q[k] * q[k];
    return r - 0.5 * s2inv * pr;
  }
  double log_prob(const double* q, bool jac) const override { return lp(q, jac); }
  cplx log_prob_c(const cplx* q, bool jac) const override { return lp(q, jac); }
  double log_prob_grad(const double* q, double* g, bool) const override {
    double r = 0, pr = 0;
    for (int k = 0; k < D; ++k) { g[k] = -s2inv * q[k]; pr += q[k] * q[k]; }
    for (int n = 0; n < N; ++n) {
      const double* x = &X[(size_t)n * D];
      double eta = 0;
      for (int k = 0; k < D; ++k) eta += x[k] * q[k];
      r += y[n] * eta - log1p_exp(eta);
      double res = y[n] - inv_logit(eta);
      for (int k = 0; k < D; ++k) g[k] += res * x[k];
    }
    return r - 0.5 * s2inv * pr;
  }
  void write_array(const double* q, double* out) const override { for (int k = 0; k < D; ++k) out[k] = q[k]; }
};

// ------------------------------------------------------ eight schools (CP)
// Parameters in the order of src/test/interface/example_output/eight_schools_output.csv:42
// (mu, tau, theta[J]); theta ~ normal(mu, tau); y ~ normal(theta, sigma)
// (src/test/test-models/eight_schools.stan:11-14).  Optional priors:
// mu ~ normal(0, mu_scale) if mu_scale is given; tau ~ cauchy(0, tau_scale) if given.
struct EightSchoolsCP : Model {
  int J; std::vector<double> y, sigma; double mu_s, tau_s;
  explicit EightSchoolsCP(const Data& d) {
    J = d.integer("J"); y = d.vec("y"); sigma = d.vec("sigma");
    mu_s = d.scalar_or("mu_scale", 0); tau_s = d.scalar_or("tau_scale", 0);
    if ((int)y.size() != J || (int)sigma.size() != J) throw std::invalid_argument("eight_schools: bad sizes");
  }
  std::string name() const override { return "eight_schools_cp"; }
  int dim() const override { return J + 2; }
  int num_out() const override { return J + 2; }
  std::vector<std::string> out_names() const override {
    std::vector<std::string> r{"mu", "tau"};
    for (int j = 0; j < J; ++j) r.push_back("theta." + std::to_string(j + 1));
    return r;
  }
  template <class T> T lp(const T* q, bool jac) const {
    using std::exp; using std::log;
    T mu = q[0], v = q[1], tau = exp(v);
    T r = 0;
    if (mu_s > 0) r -= 0.5 * mu * mu / (mu_s * mu_s);
    if (tau_s > 0) r -= log(T(1.0) + tau * tau / (tau_s * tau_s));
    if (jac) r += v;
    for (int j = 0; j < J; ++j) {
      T th = q[2 + j];
      T z = (th - mu) / tau;
      r += -v - 0.5 * z * z;
      T e = (y[j] - th) / sigma[j];
      r += -0.5 * e * e;
    }
    return r;
  }
  double log_prob(const double* q, bool jac) const override { return lp(q, jac); }
  cplx log_prob_c(const cplx* q, bool jac) const override { return lp(q, jac); }
  double log_prob_grad(const double* q, double* g, bool jac) const override {
    double mu = q[0], v = q[1], tau = std::exp(v);
    double gmu = 0, gtau = 0;  // gtau: d/dtau
    if (mu_s > 0) gmu -= mu / (mu_s * mu_s);
    if (tau_s > 0) gtau -= 2.0 * tau / (tau_s * tau_s + tau * tau);
    double gv = jac ? 1.0 : 0.0;
    for (int j = 0; j < J; ++j) {
      double th = q[2 + j];
      double dz = (th - mu) / (tau * tau);
      gmu += dz;
      gtau += -1.0 / tau + (th - mu) * (th - mu) / (tau * tau * tau);
      g[2 + j] = -dz + (y[j] - th) / (sigma[j] * sigma[j]);
    }
    g[0] = gmu;
    g[1] = gv + tau * gtau;
    return lp(q, jac);
  }
  void write_array(const double* q, double* out) const override {
    out[0] = q[0]; out[1] = std::exp(q[1]);
    for (int j = 0; j < J; ++j) out[2 + j] = q[2 + j];
  }
};

// --------------------------------------------- hierarchical normal, non-centered
// mu ~ normal(0, 5); tau ~ cauchy(0, 5), tau > 0; eta ~ std_normal();
// theta = mu + tau * eta; y ~ normal(theta, sigma).  Unconstrained (mu, log tau, eta).
// Outputs mu, tau, eta[J], theta[J].
struct HierNormalNC : Model {
  int J; std::vector<double> y, sigma; double mu_s, tau_s;
  explicit HierNormalNC(const Data& d) {
    J = d.integer("J"); y = d.vec("y"); sigma = d.vec("sigma");
    mu_s = d.scalar_or("mu_scale", 5.0); tau_s = d.scalar_or("tau_scale", 5.0);
    if ((int)y.size() != J || (int)sigma.size() != J) throw std::invalid_argument("hier_normal_nc: bad sizes");
  }
  std::string name() const override { return "hier_normal_nc"; }
  int dim() const override { return J + 2; }
  int num_out() const override { return 2 + 2 * J; }
  std::vector<std::string> out_names() const override {
    std::vector<std::string> r{"mu", "tau"};
    for (int j = 0; j < J; ++j) r.push_back("eta." + std::to_string(j + 1));
    for (int j = 0; j < J; ++j) r.push_back("theta." + std::to_string(j + 1));
    return r;
  }
  template <class T> T lp(const T* q, bool jac) const {
    using std::exp; using std::log;
    T mu = q[0], v = q[1], tau = exp(v);
    T r = -0.5 * mu * mu / (mu_s * mu_s) - log(T(1.0) + tau * tau / (tau_s * tau_s));
    if (jac) r += v;
    for (int j = 0; j < J; ++j) {
      T eta = q[2 + j];
      T th = mu + tau * eta;
      T e = (y[j] - th) / sigma[j];
      r += -0.5 * eta * eta - 0.5 * e * e;
    }
    return r;
  }
  double log_prob(const double* q, bool jac) const override { return lp(q, jac); }
  cplx log_prob_c(const cplx* q, bool jac) const override { return lp(q, jac); }
  double log_prob_grad(const double* q, double* g, bool jac) const override {
    double mu = q[0], v = q[1], tau = std::exp(v);
    double gmu = -mu / (mu_s * mu_s);
    double gtau = -2.0 * tau / (tau_s * tau_s + tau * tau);
    for (int j = 0; j < J; ++j) {
      double eta = q[2 + j];
      double th = mu + tau * eta;
      double r = (y[j] - th) / (sigma[j] * sigma[j]);
      gmu += r;
      gtau += r * eta;
      g[2 + j] = -eta + r * tau;
    }
    g[0] = gmu;
    g[1] = (jac ? 1.0 : 0.0) + tau * gtau;
    return lp(q, jac);
  }
  void write_array(const double* q, double* out) const override {
    double tau = std::exp(q[1]);
    out[0] = q[0]; out[1] = tau;
    for (int j = 0; j < J; ++j) { out[2 + j] = q[2 + j]; out[2 + J + j] = q[0] + tau * q[2 + j]; }
  }
};

// ------------------------------------------ centered hierarchical logistic
// mu[k] ~ normal(0, mu_scale=5); tau[k] ~ normal+(0, 1);
// beta[g][k] ~ normal(mu[k], tau[k]); y_n ~ bernoulli_logit(x_n . beta[group[n]]).
// Unconstrained (mu[K], log tau[K], beta[G][K] row-major).  group is 1-based.
struct HierLogisticCP : Model {
  int N, K, G; std::vector<double> X, y; std::vector<int> grp; double mu_s;
  explicit HierLogisticCP(const Data& d) {
    N = d.integer("N"); K = d.integer("K"); G = d.integer("G");
    X = d.vec("X"); y = d.vec("y");
    const auto& gg = d.vec("group");
    mu_s = d.scalar_or("mu_scale", 5.0);
    if ((int64_t)X.size() != (int64_t)N * K || (int)y.size() != N || (int)gg.size() != N)
      throw std::invalid_argument("hier_logistic_cp: bad sizes");
    grp.resize(N);
    for (int n = 0; n < N; ++n) {
      grp[n] = (int)std::llround(gg[n]) - 1;
      if (grp[n] < 0 || grp[n] >= G) throw std::invalid_argument("hier_logistic_cp: group out of range");
    }
  }
  std::string name() const override { return "hier_logistic_cp"; }
  int dim() const override { return 2 * K + G * K; }
  int num_out() const override { return 2 * K + G * K; }
  std::vector<std::string> out_names() const override {
    std::vector<std::string> r;
    for (int k = 0; k < K; ++k) r.push_back("mu." + std::to_string(k + 1));
    for (int k = 0; k < K; ++k) r.push_back("tau." + std::to_string(k + 1));
    for (int k = 0; k < K; ++k)
      for (int g = 0; g < G; ++g) r.push_back("beta." + std::to_string(g + 1) + "." + std::to_string(k + 1));
    return r;
  }
  template <class T> T lp(const T* q, bool jac) const {
    using std::exp;
    T r = 0;
    std::vector<T> tau(K);
    for (int k = 0; k < K; ++k) {
      T mu = q[k], v = q[K + k];
      tau[k] = exp(v);
      r += -0.5 * mu * mu / (mu_s * mu_s) - 0.5 * tau[k] * tau[k];
      if (jac) r += v;
    }
    const T* beta = q + 2 * K;
    for (int g = 0; g < G; ++g)
      for (int k = 0; k < K; ++k) {
        T z = (beta[g * K + k] - q[k]) / tau[k];
        r += -q[K + k] - 0.5 * z * z;
      }
    for (int n = 0; n < N; ++n) {
      const T* b = beta + (size_t)grp[n] * K;
      T eta = 0;
      for (int k = 0; k < K; ++k) eta += X[(size_t)n * K + k] * b[k];
      r += y[n] * eta - log1p_exp(eta);
    }
    return r;
  }
  double log_prob(const double* q, bool jac) const override { return lp(q, jac); }
  cplx log_prob_c(const cplx* q, bool jac) const override { return lp(q, jac); }
  double log_prob_grad(const double* q, double* g, bool jac) const override {
    const double* beta = q + 2 * K;
    double* gb = g + 2 * K;
    std::vector<double> tau(K), gtau(K, 0.0);
    for (int k = 0; k < K; ++k) {
      tau[k] = std::exp(q[K + k]);
      g[k] = -q[k] / (mu_s * mu_s);
      gtau[k] = -tau[k];
    }
    for (int gi = 0; gi < G; ++gi)
      for (int k = 0; k < K; ++k) {
        double d = beta[gi * K + k] - q[k];
        double t2 = tau[k] * tau[k];
        g[k] += d / t2;
        gtau[k] += -1.0 / tau[k] + d * d / (t2 * tau[k]);
        gb[gi * K + k] = -d / t2;
      }
    for (int k = 0; k < K; ++k) g[K + k] = (jac ? 1.0 : 0.0) + tau[k] * gtau[k];
    for (int n = 0; n < N; ++n) {
      const double* b = beta + (size_t)grp[n] * K;
      const double* x = &X[(size_t)n * K];
      double eta = 0;
      for (int k = 0; k < K; ++k) eta += x[k] * b[k];
      double res = y[n] - inv_logit(eta);
      double* gg = gb + (size_t)grp[n] * K;
      for (int k = 0; k < K; ++k) gg[k] += res * x[k];
    }
    return lp(q, jac);
  }
  void write_array(const double* q, double* out) const override {
    for (int k = 0; k < K; ++k) { out[k] = q[k]; out[K + k] = std::exp(q[K + k]); }
    // beta in column-major output order (first index fastest), as Stan writes containers
    for (int k = 0; k < K; ++k)
      for (int g = 0; g < G; ++g) out[2 * K + k * G + g] = q[2 * K + g * K + k];
  }
};

// ------------------------------------------------------------ gaussian
// q ~ multi_normal_prec(mu, P) with P given (D x D, row-major):
// lp = -1/2 (q-mu)' P (q-mu).  Covers src/test/test-models/proper.stan (P=[1])
// and test_model.stan (P=diag(1/100, 1)).
struct Gauss : Model {
  int D; std::vector<double> mu, P;
  explicit Gauss(const Data& d) {
    D = d.integer("D");
    mu = d.has("mu") ? d.vec("mu") : std::vector<double>(D, 0.0);
    if (d.has("P")) P = d.vec("P");
    else { P.assign((size_t)D * D, 0.0); for (int i = 0; i < D; ++i) P[(size_t)i * D + i] = 1.0; }
    if ((int)mu.size() != D || (int64_t)P.size() != (int64_t)D * D) throw std::invalid_argument("gauss: bad sizes");
  }
  std::string name() const override { return "gauss"; }
  int dim() const override { return D; }
  int num_out() const override { return D; }
  std::vector<std::string> out_names() const override {
    std::vector<std::string> r; for (int i = 0; i < D; ++i) r.push_back("x." + std::to_string(i + 1)); return r;
  }
  template <class T> T lp(const T* q, bool) const {
    T r = 0;
    for (int i = 0; i < D; ++i) {
      T s = 0;
      for (int j = 0; j < D; ++j) s += P[(size_t)i * D + j] * (q[j] - mu[j]);
      r += (q[i] - mu[i]) * s;
    }
    return -0.5 * r;
  }
  double log_prob(const double* q, bool jac) const override { return lp(q, jac); }
  cplx log_prob_c(const cplx* q, bool jac) const override { return lp(q, jac); }
  double log_prob_grad(const double* q, double* g, bool jac) const override {
    for (int i = 0; i < D; ++i) {
      double s = 0;
      for (int j = 0; j < D; ++j) s += P[(size_t)i * D + j] * (q[j] - mu[j]);
      g[i] = -s;
    }
    return lp(q, jac);
  }
  void write_array(const double* q, double* out) const override { for (int i = 0; i < D; ++i) out[i] = q[i]; }
};

// ------------------------------------------------------------------ funnel
// Neal's funnel (Stan User's Guide, "Reparameterization"): v ~ normal(0, 3),
// x_k ~ normal(0, exp(v / 2)), k = 1..K; no constraints.  Outputs v, the
// derived scale sigma = exp(v / 2), then x.  The model of the plugin example.
struct Funnel : Model {
  int K;
  explicit Funnel(const Data& d) {
    K = d.integer("K");
    if (K < 1) throw std::invalid_argument("funnel: K must be positive");
  }
  std::string name() const override { return "funnel"; }
  int dim() const override { return K + 1; }
  int num_out() const override { return K + 2; }
  std::vector<std::string> out_names() const override {
    std::vector<std::string> r = {"v", "sigma"};
    for (int k = 0; k < K; ++k) r.push_back("x." + std::to_string(k + 1));
    return r;
  }
  template <class T> T lp(const T* q, bool) const {
    using std::exp;
    const T v = q[0];
    T r = -v * v / 18.0;
    const T iv = exp(-v);
    for (int k = 0; k < K; ++k) r += -0.5 * q[1 + k] * q[1 + k] * iv - 0.5 * v;
    return r;
  }
  double log_prob(const double* q, bool jac) const override { return lp(q, jac); }
  cplx log_prob_c(const cplx* q, bool jac) const override { return lp(q, jac); }
  double log_prob_grad(const double* q, double* g, bool jac) const override {
    const double v = q[0], iv = std::exp(-v);
    double gv = -v / 9.0;
    for (int k = 0; k < K; ++k) {
      gv += 0.5 * q[1 + k] * q[1 + k] * iv - 0.5;
      g[1 + k] = -q[1 + k] * iv;
    }
    g[0] = gv;
    return lp(q, jac);
  }
  void write_array(const double* q, double* out) const override {
    out[0] = q[0];
    out[1] = std::exp(0.5 * q[0]);
    for (int k = 0; k < K; ++k) out[2 + k] = q[1 + k];
  }
};

inline std::unique_ptr<Model> make_model(const std::string& name, const Data& d) {
  if (name == "funnel") return std::unique_ptr<Model>(new Funnel(d));
  if (name == "bernoulli") return std::unique_ptr<Model>(new Bernoulli(d));
  if (name == "logistic_regression") return std::unique_ptr<Model>(new LogisticRegression(d));
  if (name == "eight_schools_cp") return std::unique_ptr<Model>(new EightSchoolsCP(d));
  if (name == "hier_normal_nc") return std::unique_ptr<Model>(new HierNormalNC(d));
  if (name == "hier_logistic_cp") return std::unique_ptr<Model>(new HierLogisticCP(d));
  if (name == "gauss") return std::unique_ptr<Model>(new Gauss(d));
  throw std::invalid_argument("unknown model: " + name);
}

}  // namespace oracle
