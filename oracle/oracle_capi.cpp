// ORACLE — test infrastructure only (tests/, smoke check, bench cpu_baseline).
// extern "C" surface of the FP64 CPU restatement, loaded by the tests via ctypes.
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "models.hpp"
#include "nuts.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;
}

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// fields: n arrays named names[i] with ndims[i] dims taken consecutively from dims.
void* oracle_model_create(const char* name, int n, const char** names, const int* ndims, const int64_t* dims,
                          const double** values) {
  try {
    Data d;
    int64_t off = 0;
    for (int i = 0; i < n; ++i) {
      Data::Field f;
      int64_t sz = 1;
      for (int k = 0; k < ndims[i]; ++k) { f.dims.push_back(dims[off + k]); sz *= dims[off + k]; }
      off += ndims[i];
      f.v.assign(values[i], values[i] + sz);
      d.f[names[i]] = f;
    }
    return make_model(name, d).release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void oracle_model_destroy(void* m) { delete static_cast<Model*>(m); }
int oracle_model_dim(void* m) { return static_cast<Model*>(m)->dim(); }
int oracle_model_num_out(void* m) { return static_cast<Model*>(m)->num_out(); }

double oracle_log_prob_grad(void* m, const double* q, int jacobian, double* grad) {
  return static_cast<Model*>(m)->log_prob_grad(q, grad, jacobian != 0);
}
double oracle_log_prob(void* m, const double* q, int jacobian) {
  return static_cast<Model*>(m)->log_prob(q, jacobian != 0);
}
// complex-step derivative of lp along coordinate k: Im lp(q + i h e_k) / h
double oracle_log_prob_cs(void* m, const double* q, int k, double h, int jacobian) {
  Model* mm = static_cast<Model*>(m);
  std::vector<cplx> qc(mm->dim());
  for (int i = 0; i < mm->dim(); ++i) qc[i] = cplx(q[i], i == k ? h : 0.0);
  return mm->log_prob_c(qc.data(), jacobian != 0).imag() / h;
}
void oracle_write_array(void* m, const double* q, double* out) { static_cast<Model*>(m)->write_array(q, out); }

void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  P4 c; for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  P4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}
double oracle_stream_uniform(uint64_t seed, uint32_t chain, uint32_t purpose, uint64_t idx) {
  return stream_uniform(seed, chain, purpose, idx);
}
void oracle_stream_normal_pair(uint64_t seed, uint32_t chain, uint64_t idx, double* out2) {
  stream_normal_pair(seed, chain, idx, &out2[0], &out2[1]);
}

// One leapfrog step (for closed-form and reversibility tests).
int oracle_leapfrog(void* m, int metric, const double* minv, double eps, double* q, double* p, double* g, double* V) {
  Model* mm = static_cast<Model*>(m);
  const int D = mm->dim();
  PhiloxDraws dummy(0, 0);
  NutsSampler s(mm, metric, &dummy);
  if (metric != UNIT_E) for (size_t i = 0; i < s.ham.minv.size(); ++i) s.ham.minv[i] = minv[i];
  PS z; z.q.assign(q, q + D); z.p.assign(p, p + D); z.g.assign(g, g + D); z.V = *V;
  s.evolve(z, eps);
  for (int i = 0; i < D; ++i) { q[i] = z.q[i]; p[i] = z.p[i]; g[i] = z.g[i]; }
  *V = z.V;
  return 0;
}

// One NUTS transition from q0 with injected draws (uniforms, normals consumed in
// order).  out_q gets the selected draw; info = {depth, n_leapfrog, divergent,
// accept_stat, energy, lp, uniforms_used, normals_used, eps_used}.
int oracle_transition_replay(void* m, int metric, const double* minv, double nom_eps, double jitter, int max_depth,
                             const double* q0, const double* u, int64_t nu, const double* z, int64_t nz,
                             double* out_q, double* info) {
  try {
    Model* mm = static_cast<Model*>(m);
    const int D = mm->dim();
    InjectedDraws draws(u, (size_t)nu, z, (size_t)nz);
    NutsSampler s(mm, metric, &draws);
    if (metric != UNIT_E) for (size_t i = 0; i < s.ham.minv.size(); ++i) s.ham.minv[i] = minv[i];
    s.nom_eps = nom_eps; s.jitter = jitter; s.max_depth = max_depth;
    vec q(q0, q0 + D);
    Sample smp = s.transition(q);
    for (int i = 0; i < D; ++i) out_q[i] = smp.q[i];
    info[0] = s.depth; info[1] = s.n_leapfrog; info[2] = s.divergent ? 1 : 0;
    info[3] = smp.accept_stat; info[4] = s.energy; info[5] = smp.lp;
    info[6] = (double)draws.iu; info[7] = (double)draws.iz; info[8] = s.eps;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// base_hmc::init_stepsize from q0 with injected normals (one momentum draw per
// probe).  out = {final nominal step size, normals used}.
int oracle_init_stepsize(void* m, int metric, const double* minv, double nom_eps, const double* q0, const double* z,
                         int64_t nz, double* out) {
  try {
    Model* mm = static_cast<Model*>(m);
    const int D = mm->dim();
    InjectedDraws draws(nullptr, 0, z, (size_t)nz);
    NutsSampler s(mm, metric, &draws);
    if (metric != UNIT_E) for (size_t i = 0; i < s.ham.minv.size(); ++i) s.ham.minv[i] = minv[i];
    s.nom_eps = nom_eps;
    s.z.q.assign(q0, q0 + D);
    s.ham.update_potential_gradient(s.z);
    s.init_stepsize();
    out[0] = s.nom_eps;
    out[1] = (double)draws.iz;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Dual averaging on a given accept-stat sequence (deterministic recurrence).
void oracle_dual_averaging(double mu, double delta, double gamma, double kappa, double t0, const double* accept,
                           int n, double* eps_out, double* x_bar_out) {
  StepsizeAdaptation sa;
  sa.mu = mu; sa.delta = delta; sa.gamma = gamma; sa.kappa = kappa; sa.t0 = t0;
  double e = 0;
  for (int i = 0; i < n; ++i) { sa.learn_stepsize(e, accept[i]); eps_out[i] = e; }
  *x_bar_out = sa.x_bar;
}

// Window schedule: out[i] = 1 if a metric update happens after warmup iteration i.
int oracle_window_schedule(unsigned nw, unsigned ib, unsigned tb, unsigned bw, int* collect, int* update) {
  WindowedAdaptation wa;
  int mode = wa.set_window_params(nw, ib, tb, bw);
  for (unsigned i = 0; i < nw; ++i) {
    collect[i] = wa.adaptation_window() ? 1 : 0;
    update[i] = 0;
    if (wa.end_adaptation_window()) { wa.compute_next_window(); update[i] = 1; }
    ++wa.counter;
  }
  return mode;
}

// Full run of `num_chains` chains (one std::thread per chain, up to n_threads at
// once).  cfg_d = {num_warmup, num_samples, thin, save_warmup, adapt, delta, gamma,
// kappa, t0, init_buffer, term_buffer, window, stepsize, jitter, max_depth,
// metric, init_radius}.  draws: [chain][row][col]; stepsize [chain];
// inv_metric [chain][D or D*D]; stats [chain] = {grad_evals, leapfrogs, seconds}.
int oracle_run(void* m, const double* cfg_d, uint64_t seed, int chain_id0, int num_chains, int n_threads,
               const double* init_q, const double* inv_metric_in, double* draws, double* stepsize,
               double* inv_metric, double* stats) {
  Model* mm = static_cast<Model*>(m);
  RunConfig c;
  c.num_warmup = (int)cfg_d[0]; c.num_samples = (int)cfg_d[1]; c.thin = (int)cfg_d[2];
  c.save_warmup = cfg_d[3] != 0; c.adapt = cfg_d[4] != 0; c.delta = cfg_d[5]; c.gamma = cfg_d[6];
  c.kappa = cfg_d[7]; c.t0 = cfg_d[8]; c.init_buffer = (int)cfg_d[9]; c.term_buffer = (int)cfg_d[10];
  c.window = (int)cfg_d[11]; c.stepsize = cfg_d[12]; c.jitter = cfg_d[13]; c.max_depth = (int)cfg_d[14];
  c.metric = (int)cfg_d[15]; c.init_radius = cfg_d[16];
  const int D = mm->dim();
  const size_t msz = c.metric == DENSE_E ? (size_t)D * D : (size_t)D;
  std::vector<RunResult> res(num_chains);
  std::vector<double> secs(num_chains);
  auto work = [&](int k) {
    auto t0 = std::chrono::steady_clock::now();
    PhiloxDraws rng(seed, (uint32_t)(chain_id0 + k));
    res[k] = run_chain(*mm, c, rng, init_q ? init_q + (size_t)k * D : nullptr,
                       inv_metric_in ? inv_metric_in + (size_t)k * msz : nullptr);
    secs[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  if (n_threads < 1) n_threads = 1;
  for (int base = 0; base < num_chains; base += n_threads) {
    std::vector<std::thread> th;
    for (int k = base; k < num_chains && k < base + n_threads; ++k) th.emplace_back(work, k);
    for (auto& t : th) t.join();
  }
  for (int k = 0; k < num_chains; ++k) {
    if (!res[k].error.empty()) { g_err = res[k].error; return 1; }
    std::memcpy(draws + (size_t)k * res[k].draws.size(), res[k].draws.data(), res[k].draws.size() * sizeof(double));
    stepsize[k] = res[k].stepsize;
    if (inv_metric) std::memcpy(inv_metric + (size_t)k * msz, res[k].inv_metric.data(), msz * sizeof(double));
    stats[3 * k + 0] = (double)res[k].grad_evals;
    stats[3 * k + 1] = (double)res[k].leapfrogs;
    stats[3 * k + 2] = secs[k];
  }
  return 0;
}

}  // extern "C"
