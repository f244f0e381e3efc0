// C-ABI implementation: data contexts, model registry and the batched NUTS
// driver (warmup phase, then sampling phase, every chain of every dataset in
// one batch on one GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cmdstan_amd.h"
#include "json.h"
#include "rdump.h"
#include "launch.h"
#include "flow.h"
#include "hier_logistic.h"
#include "glm_large.h"
#include "nuts_types.h"

using namespace nutsgpu;

namespace {

thread_local std::string g_err;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPC(x)                                                                                         \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess)                                                                               \
      throw HipError(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" +      \
                     std::to_string(__LINE__));                                                         \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() {}
  explicit DevBuf(size_t n_) { alloc(n_); }
  void alloc(size_t n_) {
    free();
    n = n_;
    if (n) HIPC(hipMalloc(&p, n * sizeof(T)));
  }
  void upload(const T* h, size_t cnt, hipStream_t st = 0) {
    if (cnt) HIPC(hipMemcpyAsync(p, h, cnt * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void upload(const std::vector<T>& h, hipStream_t st = 0) {
    if (h.size() > n) alloc(h.size());
    upload(h.data(), h.size(), st);
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// ------------------------------------------------------------------ data
struct Field {
  std::vector<int64_t> dims;
  std::vector<double> v;
  bool is_int = false;
};

}  // namespace

struct nuts_data {
  std::map<std::string, Field> f;
  const Field& get(const std::string& n) const {
    auto it = f.find(n);
    if (it == f.end()) throw std::invalid_argument("variable does not exist; variable name = " + n);
    return it->second;
  }
  bool has(const std::string& n) const { return f.count(n) != 0; }
  double scalar(const std::string& n) const {
    const Field& x = get(n);
    if (x.v.size() != 1) throw std::invalid_argument("variable " + n + " must be a scalar");
    return x.v[0];
  }
  double scalar_or(const std::string& n, double d) const { return has(n) ? scalar(n) : d; }
  int integer(const std::string& n) const {
    const double v = scalar(n);
    if (v != std::floor(v)) throw std::invalid_argument("variable " + n + " must be an integer");
    return (int)v;
  }
  const std::vector<double>& vec(const std::string& n, int64_t expect) const {
    const Field& x = get(n);
    if (expect >= 0 && (int64_t)x.v.size() != expect)
      throw std::invalid_argument("variable " + n + ": size mismatch; expecting " + std::to_string(expect) +
                                  " values, found " + std::to_string(x.v.size()));
    return x.v;
  }
};

namespace {

// -------------------------------------------------------------- models
struct Dataset {
  int D = 0, nout = 0;
  int n_glob = 0, group_size = 1;  // flow coupling masks: leading global dims, then groups
  int groups = 0;                  // hier_logistic_cp: G (the block itself lives on the device only)
  std::vector<std::string> names, unames;
  std::vector<double> block;
};

struct ModelImpl {
  std::string family;
  int kind = 0;
  std::vector<Dataset> ds;
  DevBuf<double> d_data;
  std::vector<int64_t> ds_off;
  // logistic regression batched paths (single dataset): glm for D <= 128 (FP64
  // or FP32/BF16 contraction), glm_large for larger D (FP32/BF16 contraction)
  bool glm = false, glm_large = false;
  int glmN = 0, glmD = 0;
  double glm_s2inv = 0;
  const double* d_X = nullptr;
  const double* d_y = nullptr;
  // autotuned GLM gradient per (chains, precision): path 0 = single-pass
  // kernel with blocks_per_cu resident blocks per CU, path 1 = two-pass kernels,
  // path 2 = single-pass BF16-split kernel (FP32-level accuracy on BF16 matrix cores)
  struct GlmTune { int path = 0, blocks_per_cu = 2; double ms = 0; };
  std::map<std::pair<int, int>, GlmTune> glm_tune;
  // caller-supplied model (nuts_model_create_plugin)
  bool plugin = false;
  nuts_model_plugin pl{};
  DevBuf<unsigned char> xlow;  // X converted to FP32 or BF16 (lazily)
  int xlow_prec = 0;
  DevBuf<unsigned short> xbf;  // X as BF16 terms for the BF16-split kernel (lazily)
  int xbf_prec = 0;
  const unsigned short* bf16_x(int precision, hipStream_t st) {
    if (xbf_prec != precision) {
      const int64_t n = (int64_t)glmN * glm_bf16_ldx(glmD);
      xbf.alloc((size_t)n * (precision == 1 ? 3 : 1));
      glm_bf16_prepare_x(d_X, glmN, glmD, precision, xbf.p, st);
      HIPC(hipStreamSynchronize(st));
      xbf_prec = precision;
    }
    return xbf.p;
  }
  const void* low_x(int precision, hipStream_t st) {
    if (xlow_prec != precision) {
      const int64_t n = (int64_t)glmN * glmD;
      xlow.alloc((size_t)n * (precision == 1 ? 4 : 2));
      glm_convert_x(d_X, xlow.p, n, precision, st);
      HIPC(hipStreamSynchronize(st));
      xlow_prec = precision;
    }
    return xlow.p;
  }
  // BF16 X transposed [Dp][ldn] for pass 2 of the BF16 large-D kernels (made
  // once from the BF16 X, kept beside it: both fit HBM many times over)
  DevBuf<unsigned short> xt;
  const unsigned short* bf16_xt(hipStream_t st) {
    if (!xt.p) {
      const GlLargePlan gp = gl_large_plan(glmN, 1, glmD, 0);
      const unsigned short* xb = static_cast<const unsigned short*>(low_x(2, st));
      xt.alloc(gp.xt_elems);
      gl_large_prepare_xt(xb, glmN, glmD, xt.p, gp, st);
      HIPC(hipStreamSynchronize(st));
    }
    return xt.p;
  }
};

// Scratch of the large-D two-pass gradient.  With X stored in BF16 (and D a
// multiple of 8) the products run on BF16 matrix cores with positions and
// residuals as three BF16 terms; CMDSTAN_AMD_GL_BF16=0 keeps them on FP32
// matrix cores.
struct GlScratch {
  DevBuf<float> qf, r;
  DevBuf<unsigned short> qb, rb;
  DevBuf<double> wg, wlp;
};
bool gl_bf16_enabled(int precision, int D) {
  if (precision != 2 || !gl_large_bf16_ok(D)) return false;
  const char* e = std::getenv("CMDSTAN_AMD_GL_BF16");
  return !e || std::atoi(e) != 0;
}
void gl_bind(ModelImpl& M, GlLargeArgs& a, const GlLargePlan& p, GlScratch& s, int precision, hipStream_t st) {
  s.wg.alloc(p.wg_doubles);
  s.wlp.alloc(p.wlp_doubles);
  a.y = M.d_y; a.N = M.glmN; a.D = M.glmD; a.s2inv = M.glm_s2inv;
  a.Wg = s.wg.p; a.Wlp = s.wlp.p; a.Dq = p.Dq; a.ldr = p.ldr;
  if (gl_bf16_enabled(precision, M.glmD)) {
    s.qb.alloc(p.qb_elems);
    s.rb.alloc(p.rb_elems);
    a.Qb = s.qb.p; a.Rb = s.rb.p; a.Xt = M.bf16_xt(st);
    a.Qf = nullptr; a.R = nullptr;
  } else {
    s.qf.alloc(p.q_floats);
    s.r.alloc(p.r_floats);
    a.Qf = s.qf.p; a.R = s.r.p;
    a.Qb = nullptr; a.Rb = nullptr; a.Xt = nullptr;
  }
}

// Register layout of the 16x16 f32 MFMA result (0: row 4*(l/16)+i, 1: row 4i + l/16).
constexpr int kF32Layout = 0;

std::vector<std::string> idx_names(const std::string& base, int n) {
  std::vector<std::string> r;
  for (int i = 0; i < n; ++i) r.push_back(base + "." + std::to_string(i + 1));
  return r;
}

void check_binary(const std::vector<double>& y, const std::string& n) {
  for (double v : y)
    if (v != 0 && v != 1) throw std::domain_error(n + " is " + std::to_string(v) + ", but must be in the interval [0, 1]");
}

Dataset make_dataset(const std::string& fam, const nuts_data& d, int& kind) {
  Dataset s;
  if (fam == "bernoulli") {
    kind = MK_BERNOULLI;
    const int N = d.integer("N");
    if (N < 0) throw std::domain_error("N is " + std::to_string(N) + ", but must be greater than or equal to 0");
    const auto& y = d.vec("y", N);
    check_binary(y, "y");
    double sy = 0;
    for (double v : y) sy += v;
    s.D = 1; s.nout = 1;
    s.names = {"theta"}; s.unames = {"theta"};
    s.block = {(double)N, sy};
  } else if (fam == "gauss") {
    kind = MK_GAUSS;
    const int D = d.integer("D");
    if (D < 1) throw std::domain_error("D must be positive");
    std::vector<double> mu = d.has("mu") ? d.vec("mu", D) : std::vector<double>(D, 0.0);
    std::vector<double> P;
    if (d.has("P")) P = d.vec("P", (int64_t)D * D);
    else { P.assign((size_t)D * D, 0.0); for (int i = 0; i < D; ++i) P[(size_t)i * D + i] = 1.0; }
    s.D = D; s.nout = D;
    s.names = idx_names("x", D); s.unames = s.names;
    s.block.push_back(D);
    s.block.insert(s.block.end(), mu.begin(), mu.end());
    s.block.insert(s.block.end(), P.begin(), P.end());
  } else if (fam == "eight_schools_cp" || fam == "hier_normal_nc") {
    const bool nc = fam == "hier_normal_nc";
    kind = nc ? MK_HIER_NORMAL_NC : MK_EIGHT_SCHOOLS;
    const int J = d.integer("J");
    if (J < 1) throw std::domain_error("J must be positive");
    const auto& y = d.vec("y", J);
    const auto& sg = d.vec("sigma", J);
    for (double v : sg) if (!(v > 0)) throw std::domain_error("sigma must be positive");
    const double mu_s = d.scalar_or("mu_scale", nc ? 5.0 : 0.0);
    const double tau_s = d.scalar_or("tau_scale", nc ? 5.0 : 0.0);
    s.D = J + 2;
    s.nout = nc ? 2 + 2 * J : 2 + J;
    s.names = {"mu", "tau"};
    s.unames = {"mu", "tau"};
    auto th = idx_names(nc ? "eta" : "theta", J);
    s.names.insert(s.names.end(), th.begin(), th.end());
    s.unames.insert(s.unames.end(), th.begin(), th.end());
    if (nc) { auto t2 = idx_names("theta", J); s.names.insert(s.names.end(), t2.begin(), t2.end()); }
    s.n_glob = 2;
    s.block = {(double)J, mu_s, tau_s};
    s.block.insert(s.block.end(), y.begin(), y.end());
    s.block.insert(s.block.end(), sg.begin(), sg.end());
  } else if (fam == "hier_logistic_cp") {
    kind = MK_HIER_LOGISTIC;
    const int N = d.integer("N"), K = d.integer("K"), G = d.integer("G");
    if (K < 1 || K > 16) throw std::domain_error("K must be in [1, 16]");
    if (G < 1) throw std::domain_error("G must be positive");
    const auto& X = d.vec("X", (int64_t)N * K);
    const auto& y = d.vec("y", N);
    const auto& gg = d.vec("group", N);
    check_binary(y, "y");
    std::vector<int> grp(N);
    for (int n = 0; n < N; ++n) {
      grp[n] = (int)gg[n] - 1;
      if (grp[n] < 0 || grp[n] >= G) throw std::domain_error("group out of range");
    }
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return grp[a] < grp[b]; });
    std::vector<double> gstart(G + 1, 0.0);
    for (int n = 0; n < N; ++n) gstart[grp[n] + 1] += 1;
    for (int g = 0; g < G; ++g) gstart[g + 1] += gstart[g];
    s.D = 2 * K + G * K;
    s.nout = s.D;
    s.names = idx_names("mu", K);
    auto t = idx_names("tau", K);
    s.names.insert(s.names.end(), t.begin(), t.end());
    s.unames = s.names;
    for (int k = 0; k < K; ++k)
      for (int g = 0; g < G; ++g) s.names.push_back("beta." + std::to_string(g + 1) + "." + std::to_string(k + 1));
    for (int g = 0; g < G; ++g)
      for (int k = 0; k < K; ++k) s.unames.push_back("beta." + std::to_string(g + 1) + "." + std::to_string(k + 1));
    s.n_glob = 2 * K;
    s.group_size = K;
    s.groups = G;
    s.block = {(double)N, (double)K, (double)G, d.scalar_or("mu_scale", 5.0)};
    s.block.insert(s.block.end(), gstart.begin(), gstart.end());
    for (int n : order) for (int k = 0; k < K; ++k) s.block.push_back(X[(size_t)n * K + k]);
    for (int n : order) s.block.push_back(y[n]);
  } else if (fam == "logistic_regression") {
    kind = MK_LOGISTIC;
    const int N = d.integer("N"), D = d.integer("D");
    if (N < 0 || D < 1) throw std::domain_error("bad N or D");
    const auto& X = d.vec("X", (int64_t)N * D);
    const auto& y = d.vec("y", N);
    check_binary(y, "y");
    const double sc = d.scalar_or("prior_scale", 2.5);
    s.D = D; s.nout = D;
    s.names = idx_names("beta", D); s.unames = s.names;
    s.block.reserve(3 + (size_t)N * D + N);
    s.block = {(double)N, (double)D, 1.0 / (sc * sc)};
    s.block.insert(s.block.end(), X.begin(), X.end());
    s.block.insert(s.block.end(), y.begin(), y.end());
  } else {
    throw std::invalid_argument("unknown model family: " + fam);
  }
  return s;
}

int num_cus() {
  static int n = 0;
  if (n == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) != hipSuccess) return 256;
    n = pr.multiProcessorCount;
  }
  return n;
}

// ------------------------------------------------------------- windows
// stan::mcmc::windowed_adaptation::set_window_params + restart (see the
// device logic in nuts_device.h); returns the initial per-chain window state.
struct WindowInit {
  uint32_t num_warmup = 0, init_buffer = 0, term_buffer = 0, base_window = 0;
  uint32_t counter = 0, size = 0, next = 0;
};
WindowInit window_init(unsigned nw, unsigned ib, unsigned tb, unsigned bw) {
  WindowInit w;
  auto restart = [&]() { w.counter = 0; w.size = w.base_window; w.next = w.init_buffer + w.size - 1; };
  restart();  // constructor state
  if (nw < 20) return w;
  if (ib + bw + tb > nw) {
    w.num_warmup = nw;
    w.init_buffer = (uint32_t)(0.15 * nw);
    w.term_buffer = (uint32_t)(0.1 * nw);
    w.base_window = nw - (w.init_buffer + w.term_buffer);
    return w;  // no restart: the schedule keeps its constructor state
  }
  w.num_warmup = nw; w.init_buffer = ib; w.term_buffer = tb; w.base_window = bw;
  restart();
  return w;
}

// ---------------------------------------------------------------- runs
struct RunSpec {
  nuts_config cfg;
  const double* init = nullptr;        // [C][D]
  const double* inv_metric = nullptr;  // [C][D]
  const double* inj_u = nullptr;
  int64_t n_u = 0;
  const double* inj_z = nullptr;
  int64_t n_z = 0;
  int first_dataset = 0, num_datasets = -1;
  bool constrain = true;
  FlowNet* flow = nullptr;                      // sample z with x = T(z) (single dataset, lockstep)
  const double* stepsize_per_chain = nullptr;   // [C] resume with these step sizes (no adaptation)
  bool diagnostics = false;                     // record q, p, g of every saved draw
  bool energy_error = false;                    // record each saved row's max |H - H0|
  int block_mode = -1;                          // lockstep state machine: -1 by dimension, 0 wave, 1 block per chain
  int fused_block = -1;                         // fused kernel, block per chain: -1 once few chains remain, 0 never, 1 always
};

struct RunResult {
  std::vector<double> draws;  // constrained (or unconstrained if !constrain) rows, chains back to back
  std::vector<int64_t> draw_off;
  std::vector<int> cols;
  std::vector<ChainState> st;
  std::vector<double> minv;  // compact [C][D]
  double warm_s = 0, samp_s = 0;
  int64_t leaf_steps = 0;
  double grad_s = 0;
  int64_t grad_timed = 0;
  long long grad_chains = 0;  // chains evaluated, summed over GLM gradient launches
  int tuned_path = -1, tuned_bpc = 0;
  int threads_per_chain = 0;                    // lockstep state machine: 64 or 256 (0 fused)
  double wave_launch_s = 0, block_launch_s = 0;  // timed means when the mode was tuned
  std::vector<double> z_last;  // [C][D] unconstrained q of each chain's last saved row (before any flow map)
  int rows = 0;
  std::vector<double> diag;    // diagnostics: chains back to back, rows x (7 + 3 D)
  std::vector<double> eerr;    // energy error: [C][rows]
};

int output_rows(const nuts_config& c) {
  int r = 0;
  if (c.save_warmup) r += (c.num_warmup + c.thin - 1) / c.thin;
  r += (c.num_samples + c.thin - 1) / c.thin;
  return r;
}

void validate(const nuts_config& c) {
  if (c.num_samples < 0) throw std::invalid_argument("num_samples must be >= 0");
  if (c.num_warmup < 0) throw std::invalid_argument("num_warmup must be >= 0");
  if (c.thin < 1) throw std::invalid_argument("thin must be > 0");
  if (c.adapt_engaged && c.num_warmup == 0)
    throw std::invalid_argument("The number of warmup samples (num_warmup) must be greater than zero if adaptation "
                                "is enabled.");
  if (!(c.delta >= 0.1 && c.delta < 1.0)) throw std::invalid_argument("delta must be in [0.1, 1)");
  if (!(c.gamma > 0)) throw std::invalid_argument("gamma must be positive");
  if (!(c.kappa > 0)) throw std::invalid_argument("kappa must be positive");
  if (!(c.t0 > 0)) throw std::invalid_argument("t0 must be positive");
  if (!(c.stepsize > 0)) throw std::invalid_argument("stepsize must be positive");
  if (!(c.stepsize_jitter >= 0 && c.stepsize_jitter <= 1)) throw std::invalid_argument("stepsize_jitter must be in [0, 1]");
  if (c.max_depth < 1) throw std::invalid_argument("max_depth must be > 0");
  if (c.max_depth > NUTS_MAX_DEPTH)
    throw std::invalid_argument("max_depth above " + std::to_string(NUTS_MAX_DEPTH) + " is not supported");
  if (c.metric < 0 || c.metric > 2) throw std::invalid_argument("metric must be unit_e, diag_e or dense_e");

  if (c.num_chains < 1) throw std::invalid_argument("num_chains must be >= 1");
  if (c.init_radius < 0) throw std::invalid_argument("init radius must be >= 0");
  if (c.precision < 0 || c.precision > 2) throw std::invalid_argument("precision must be 0 (fp64), 1 (fp32) or 2 (bf16)");
  if (c.flow < 0 || c.flow > 2) throw std::invalid_argument("flow must be 0, 1 or 2");
  if (c.flow) {
    if (c.flow_layers < 1 || c.flow_hidden < 1 || c.flow_train_steps < 0 || c.flow_batch < 1 || !(c.flow_lr > 0) ||
        c.flow_warmup < 1 || c.flow_pilot < 4)
      throw std::invalid_argument("flow: layers, hidden, batch, lr, flow_warmup must be positive and flow_pilot >= 4");
    if (!c.adapt_engaged) throw std::invalid_argument("flow: adaptation must be engaged");
    if (c.save_warmup) throw std::invalid_argument("flow: save_warmup is not supported");
  }
}

// iterations completed by the slowest chain (one block)
__global__ void min_iter_kernel(const ChainState* st, int C, int* out) {
  __shared__ int m[256];
  int v = 0x7fffffff;
  for (int c = threadIdx.x; c < C; c += blockDim.x) v = min(v, st[c].iter);
  m[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) m[threadIdx.x] = min(m[threadIdx.x], m[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = m[0];
}

__global__ void resume_kernel(ChainState* st, int C) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C && st[c].phase == PH_PAUSED) st[c].phase = PH_TRANS_BEGIN;
}

// Times the candidate GLM gradient implementations on the model's own data with
// every chain active (positions at zero) and keeps the fastest per
// (chains, precision).  Candidates: the single-pass kernel at 1-4 resident
// blocks per CU, and for FP32/BF16 the two-pass kernels.
ModelImpl::GlmTune glm_autotune(ModelImpl& M, int C, int precision, hipStream_t st) {
  const auto key = std::make_pair(C, precision);
  auto it = M.glm_tune.find(key);
  if (it != M.glm_tune.end()) return it->second;
  const int N = M.glmN, D = M.glmD, Dp = (D + 15) / 16 * 16;
  DevBuf<double> Q((size_t)C * Dp), G((size_t)C * Dp), V(C);
  HIPC(hipMemsetAsync(Q.p, 0, (size_t)C * Dp * sizeof(double), st));
  const void* xl = precision ? M.low_x(precision, st) : nullptr;
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  auto time_it = [&](auto&& launch) {
    launch();
    HIPC(hipEventRecord(e0, st));
    for (int r = 0; r < 3; ++r) launch();
    HIPC(hipEventRecord(e1, st));
    HIPC(hipEventSynchronize(e1));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    return ms / 3.0;
  };
  ModelImpl::GlmTune best;
  best.ms = 1e30;
  for (int bpc = 1; bpc <= 4; ++bpc) {
    const GlmPlan gp = glm_mfma_plan(N, C, D, num_cus(), bpc);
    DevBuf<double> work(gp.work_doubles);
    GlmArgs ga{M.d_X, M.d_y, N, D, C, gp.rows_per_chunk, M.glm_s2inv, Q.p, Dp, G.p, V.p, work.p};
    const double ms = time_it([&]() {
      if (precision == 0) glm_grad_mfma_f64(ga, gp, st);
      else glm_grad_mfma_f32(ga, gp, xl, precision, kF32Layout, st);
    });
    if (ms < best.ms) { best.ms = ms; best.path = 0; best.blocks_per_cu = bpc; }
  }
  if (precision != 0) {
    const unsigned short* xb = M.bf16_x(precision, st);
    for (int bpc = 1; bpc <= 4; ++bpc) {
      const GlmPlan gp = glm_mfma_plan(N, C, D, num_cus(), bpc);
      DevBuf<double> work(gp.work_doubles);
      GlmArgs ga{M.d_X, M.d_y, N, D, C, gp.rows_per_chunk, M.glm_s2inv, Q.p, Dp, G.p, V.p, work.p};
      const double ms = time_it([&]() { glm_grad_mfma_bf16(ga, gp, xb, precision, st); });
      if (ms < best.ms) { best.ms = ms; best.path = 2; best.blocks_per_cu = bpc; }
    }
    const GlLargePlan gp = gl_large_plan(N, C, D, num_cus());
    GlScratch gs;
    GlLargeArgs a;
    gl_bind(M, a, gp, gs, precision, st);
    a.C = C; a.Q = Q.p; a.ldq = Dp; a.G = G.p; a.V = V.p;
    const double ms = time_it([&]() { gl_large_grad(a, gp, xl, precision, st); });
    if (ms < best.ms) { best.ms = ms; best.path = 1; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPC(hipStreamSynchronize(st));
  // CMDSTAN_AMD_GLM_PATH=0|1|2 forces a reduced-precision path (tests compare them)
  if (const char* fp = std::getenv("CMDSTAN_AMD_GLM_PATH"))
    if (precision != 0) {
      const int want = std::atoi(fp);
      if (want >= 0 && want <= 2 && want != best.path) { best.path = want; best.blocks_per_cu = 2; }
    }
  M.glm_tune[key] = best;
  return best;
}

RunResult run_batch(ModelImpl& M, const RunSpec& R) {
  const nuts_config& cfg = R.cfg;
  validate(cfg);
  const int nds = R.num_datasets < 0 ? (int)M.ds.size() - R.first_dataset : R.num_datasets;
  const int C = nds * cfg.num_chains;
  const int rows = output_rows(cfg);
  std::vector<int> D(C), Dp(C), nout(C);
  std::vector<int64_t> off(C), moff(C), doff(C), dsoff(C), ooff(C);
  std::vector<uint32_t> uid(C);
  int64_t tot = 0, dtot = 0, otot = 0;
  const int msp = cfg.max_saved_params;
  auto nsave_of = [msp](int d) { return msp < 0 ? d : std::min(d, msp); };
  for (int c = 0; c < C; ++c) {
    const int m = R.first_dataset + c / cfg.num_chains;
    D[c] = M.ds[m].D;
    Dp[c] = (D[c] + 15) / 16 * 16;
    nout[c] = M.ds[m].nout;
    off[c] = tot;
    moff[c] = 0;
    tot += Dp[c];
    doff[c] = dtot;
    dtot += (int64_t)rows * (7 + nsave_of(D[c]));
    ooff[c] = otot;
    otot += (int64_t)rows * (7 + nout[c]);
    dsoff[c] = M.ds_off[m];
    uid[c] = (uint32_t)(cfg.id + c);
  }
  const int maxd = cfg.max_depth;
  const int nslots = maxd + 2;
  const bool dense = cfg.metric == NUTS_METRIC_DENSE_E;
  const int NV = PV_FIXED + 3 * maxd + 2 * nslots + (dense ? 2 * maxd : 0) + (R.diagnostics ? nslots : 0);
  if (R.diagnostics && R.flow) throw std::invalid_argument("diagnostic rows are not available with the flow");
  int64_t mtot = 0;
  if (dense) {
    for (int c = 0; c < C; ++c) {
      if (D[c] > 1024) throw std::invalid_argument("dense_e supports D <= 1024");
      moff[c] = mtot;
      mtot += (int64_t)D[c] * D[c];
    }
  }

  hipStream_t st;
  HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{st};
  // the flow works on the sampler's stream for this run
  struct FlowBind {
    FlowNet* f; hipStream_t old;
    ~FlowBind() { if (f) f->bind_stream(old); }
  } fb{R.flow, R.flow ? R.flow->bind_stream(st) : nullptr};

  DevBuf<int> dD(C), dDp(C), dneed(C), dstop(1);
  DevBuf<int64_t> doffb(C), dmoff(C), ddoff(C), ddsoff(C), dooff(C);
  DevBuf<uint32_t> duid(C);
  DevBuf<double> Q(tot), G(tot), V(C), pool((size_t)tot * NV), minv(tot), draws(std::max<int64_t>(dtot, 1));
  DevBuf<ChainState> dst(C);
  DevBuf<double> uinit, du, dz;
  std::vector<int64_t> diag_off(C);
  int64_t diag_tot = 0;
  for (int c = 0; c < C; ++c) {
    diag_off[c] = diag_tot;
    diag_tot += (int64_t)rows * (7 + 3 * D[c]);
  }
  DevBuf<double> ddiag;
  DevBuf<int64_t> ddiag_off;
  if (R.diagnostics) {
    ddiag.alloc(std::max<int64_t>(diag_tot, 1));
    ddiag_off.alloc(C);
    ddiag_off.upload(diag_off, st);
  }
  dD.upload(D, st); dDp.upload(Dp, st); doffb.upload(off, st); dmoff.upload(moff, st); ddoff.upload(doff, st);
  ddsoff.upload(dsoff, st); dooff.upload(ooff, st); duid.upload(uid, st);
  HIPC(hipMemsetAsync(Q.p, 0, tot * sizeof(double), st));
  HIPC(hipMemsetAsync(G.p, 0, tot * sizeof(double), st));
  HIPC(hipMemsetAsync(V.p, 0, C * sizeof(double), st));
  HIPC(hipMemsetAsync(pool.p, 0, (size_t)tot * NV * sizeof(double), st));
  HIPC(hipMemsetAsync(dneed.p, 0, C * sizeof(int), st));
  HIPC(hipMemsetAsync(dstop.p, 0, sizeof(int), st));
  std::vector<double> hm(tot, 1.0);  // diagonal inverse metric as uploaded
  {
    if (R.inv_metric && cfg.metric == NUTS_METRIC_DIAG_E) {
      for (int c = 0; c < C; ++c)
        for (int i = 0; i < D[c]; ++i) {
          const double v = R.inv_metric[(size_t)c * D[0] + i];
          if (!(v > 0) || !std::isfinite(v)) throw std::invalid_argument("inverse metric must be positive and finite");
          hm[off[c] + i] = v;
        }
    }
    minv.upload(hm, st);
    HIPC(hipStreamSynchronize(st));
  }
  DevBuf<double> dMb, dLb, dWb;
  if (dense) {
    // column-major (i, j) at [j*D + i]; the metric file holds a symmetric matrix
    std::vector<double> hm(mtot, 0.0);
    for (int c = 0; c < C; ++c) {
      const int d = D[c];
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) {
          double v = (i == j) ? 1.0 : 0.0;
          if (R.inv_metric) {
            v = R.inv_metric[(size_t)c * d * d + (size_t)i * d + j];
            if (!std::isfinite(v)) throw std::invalid_argument("inverse metric must be finite");
          }
          hm[moff[c] + (int64_t)j * d + i] = v;
        }
    }
    dMb.upload(hm, st);
    dLb.alloc(mtot);
    // the identity start metric is its own Cholesky factor: no factorisation
    if (!R.inv_metric) dLb.upload(hm.data(), hm.size(), st);
    HIPC(hipStreamSynchronize(st));
    dWb.alloc(mtot);
    HIPC(hipMemsetAsync(dWb.p, 0, mtot * sizeof(double), st));
  }
  if (R.init) {
    std::vector<double> hi(tot, 0.0);
    for (int c = 0; c < C; ++c)
      for (int i = 0; i < D[c]; ++i) hi[off[c] + i] = R.init[(size_t)c * D[0] + i];
    uinit.upload(hi, st);
    HIPC(hipStreamSynchronize(st));
  }
  if (R.inj_u) { du.alloc((size_t)C * std::max<int64_t>(R.n_u, 1)); du.upload(R.inj_u, (size_t)C * R.n_u, st); }
  if (R.inj_z) { dz.alloc((size_t)C * std::max<int64_t>(R.n_z, 1)); dz.upload(R.inj_z, (size_t)C * R.n_z, st); }

  const WindowInit w = window_init(cfg.num_warmup, cfg.init_buffer, cfg.term_buffer, cfg.window);
  std::vector<ChainState> hs(C);
  for (int c = 0; c < C; ++c) {
    ChainState& s = hs[c];
    std::memset(&s, 0, sizeof(s));
    s.phase = PH_INIT_TRY;
    s.sample_slot = 0;
    s.free_mask = ((nslots >= 32) ? 0xFFFFFFFFu : ((1u << nslots) - 1u)) & ~1u;
    s.nom_eps = R.stepsize_per_chain ? R.stepsize_per_chain[c] : cfg.stepsize;
    s.eps = s.nom_eps;
    s.adapt_flag = cfg.adapt_engaged ? 1 : 0;
    s.sa_mu = std::log(10 * cfg.stepsize);
    s.w_counter = w.counter; s.w_size = w.size; s.w_next = w.next;
    for (int k = 0; k < NUTS_MAX_DEPTH; ++k) s.L_prop[k] = -1;
  }
  dst.upload(hs, st);

  NutsParams P;
  std::memset(&P, 0, sizeof(P));
  P.num_warmup = cfg.num_warmup; P.num_samples = cfg.num_samples; P.thin = cfg.thin;
  P.save_warmup = cfg.save_warmup; P.adapt = cfg.adapt_engaged; P.max_depth = cfg.max_depth;
  P.max_init_tries = (R.init || cfg.init_radius == 0) ? 1 : 100;
  P.iter_target = cfg.num_warmup;
  P.delta = cfg.delta; P.gamma = cfg.gamma; P.kappa = cfg.kappa; P.t0 = cfg.t0;
  P.jitter = cfg.stepsize_jitter; P.stepsize = cfg.stepsize; P.init_radius = cfg.init_radius;
  P.max_deltaH = 1000;
  P.w_num_warmup = w.num_warmup; P.w_init_buffer = w.init_buffer; P.w_term_buffer = w.term_buffer;
  P.w_base_window = w.base_window; P.w_counter0 = w.counter; P.w_size0 = w.size; P.w_next0 = w.next;
  P.metric_adapt = 1;
  P.seed = cfg.seed;
  P.nslots = nslots; P.maxd = maxd; P.NV = NV;
  P.fused_budget = 256;
  P.save_params = msp;
  P.diag = R.diagnostics ? 1 : 0;
  P.out_rows = rows;

  NutsBuffers B;
  std::memset(&B, 0, sizeof(B));
  B.C = C; B.D = dD.p; B.Dp = dDp.p; B.off = doffb.p; B.moff = dmoff.p; B.uid = duid.p;
  B.Q = Q.p; B.G = G.p; B.V = V.p; B.pool = pool.p; B.st = dst.p; B.minv = minv.p;
  B.draws = draws.p; B.draw_off = ddoff.p; B.user_init = R.init ? uinit.p : nullptr;
  B.n_stopped = dstop.p; B.need = dneed.p;
  B.inj_u = R.inj_u ? du.p : nullptr; B.inj_z = R.inj_z ? dz.p : nullptr;
  B.n_inj_u = R.n_u; B.n_inj_z = R.n_z;
  B.dM = dMb.p; B.dL = dLb.p; B.dW = dWb.p;
  B.diag = ddiag.p; B.diag_off = ddiag_off.p;
  DevBuf<double> deerr;
  if (R.energy_error && rows > 0) {
    deerr.alloc((size_t)C * rows);
    B.eerr = deerr.p;
  }
  if (dense && R.inv_metric) {
    DevBuf<int> nfail(1);
    HIPC(hipMemsetAsync(nfail.p, 0, sizeof(int), st));
    launch_dense_factor(B, nfail.p, st);
    int nf = 0;
    HIPC(hipMemcpyAsync(&nf, nfail.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (nf) throw std::invalid_argument("inverse metric is not positive definite");
  }

  ModelDataDev md{M.d_data.p, ddsoff.p, 1.0};

  bool fused = cfg.fused == 1;
  // FP32/BF16 contraction of a small-D model through the two-pass large-D kernels
  // (selected by the autotuner, or forced with CMDSTAN_AMD_GLM_LARGE=0/1)
  const char* gle = std::getenv("CMDSTAN_AMD_GLM_LARGE");
  ModelImpl::GlmTune tune;
  int res_tuned_path = -1, res_tuned_bpc = 0;
  // the batched GLM gradients also serve the flow (at the rows x = T(z))
  const bool glm_batch = M.glm && nds == 1 && (cfg.fused != 1 || R.flow);
  if (glm_batch) tune = glm_autotune(M, C, cfg.precision, st);
  const bool small_as_large =
      M.glm && cfg.precision != 0 && (gle ? std::atoi(gle) != 0 : tune.path == 1);
  const bool glm_ok = glm_batch && !small_as_large;
  if (glm_batch) {
    res_tuned_path = small_as_large ? 1 : tune.path;
    res_tuned_bpc = tune.blocks_per_cu;
  }
  // hierarchical logistic: batched group kernel (lane = chain) over one dataset
  const bool hl_ok = M.kind == MK_HIER_LOGISTIC && nds == 1;
  // large-D logistic regression: two-pass FP32 matrix-core gradient (X FP32 or BF16)
  const bool gl_ok = (M.glm_large || small_as_large) && nds == 1 && cfg.precision != 0 && (cfg.fused != 1 || R.flow);
  if (M.plugin && cfg.fused == 1) throw std::invalid_argument("plugin models run on the lockstep kernels (fused=0)");
  if (cfg.fused < 0) fused = !glm_ok && !hl_ok && !gl_ok && !M.plugin;
  if (R.flow) {
    if (nds != 1 || R.flow->dim() != D[0] || R.flow->max_rows() < C)
      throw std::invalid_argument("flow: one dataset whose dimension matches the flow, at most max_rows chains");
    fused = false;
  }
  if (!fused && M.kind == MK_LOGISTIC && !glm_ok) fused = false;  // per-chain gradient kernel in lockstep
  // FP32 leapfrog: the fused per-chain kernel keeps q, p, g, the checkpoints and
  // the metric in FP32 (arithmetic, H, tree weights and U-turn sums in FP64)
  const bool vec32 = fused && cfg.precision == 1 && !M.glm && !M.glm_large;
  if (vec32 && (dense || !model_supports_vec32(M.kind)))
    throw std::invalid_argument("precision=fp32 on the per-chain kernel needs unit_e or diag_e and a model with FP32 vectors");
  if (fused && cfg.precision == 2 && !M.glm && !M.glm_large)
    throw std::invalid_argument("precision=bf16 applies to the logistic regression design matrix only");
  if (vec32) {
    // same element offsets, FP32 values in the first half of each buffer
    std::vector<float> hf(hm.begin(), hm.end());
    HIPC(hipMemcpyAsync(minv.p, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
  }

  DevBuf<double> work;
  DevBuf<int> act_idx, n_act;
  const void* xlow = nullptr;
  const unsigned short* xbf = nullptr;  // BF16 terms of X (path 2)
  GlmArgs ga;
  GlmPlan gplan;
  // large chains: one 256-thread block per chain for the state machine
  int maxD = 0;
  for (int c = 0; c < C; ++c) maxD = std::max(maxD, D[c]);
  const char* bpc = std::getenv("CMDSTAN_AMD_BLOCK_PER_CHAIN");
  bool block_per_chain =
      !fused && (R.block_mode >= 0 ? R.block_mode == 1 : (bpc ? std::atoi(bpc) != 0 : maxD >= 512));
  // unless the caller or the environment fixes it, the first check intervals
  // of the run time one mode each, and the other mode replaces the default by
  // dimension only when it is more than 20% faster.  The modes order their
  // reductions differently (draws agree to rounding, not bit for bit), so the
  // margin keeps near-ties, whose timing could go either way, on the default:
  // a seed then reproduces its draws
  const bool tune_threads = !fused && R.block_mode < 0 && !bpc;
  const bool default_block = block_per_chain;
  // fused kernel: once at most two chains per CU are still running (the tails
  // of a phase, where the chains left are the largest and deepest), one block
  // per chain for models whose gradient spreads over a block.  The switch
  // depends only on how many chains have stopped, so a seed reproduces its
  // draws; CMDSTAN_AMD_FUSED_BLOCK=0 keeps one wave per chain throughout.
  const char* fbe = std::getenv("CMDSTAN_AMD_FUSED_BLOCK");
  const bool fused_block_ok = fused && !M.plugin && model_supports_fused_block(M.kind, cfg.metric) &&
                              R.fused_block != 0 && !(fbe && std::atoi(fbe) == 0);
  if (R.fused_block == 1 && !fused_block_ok)
    throw std::invalid_argument("this model has no block-per-chain fused kernel");
  const char* fbc = std::getenv("CMDSTAN_AMD_FUSED_BLOCK_CHAINS");  // threshold override (chains still running)
  const int fused_block_chains = fbc ? std::atoi(fbc) : 2 * num_cus();
  DevBuf<double> hlwork;
  HlArgs hla{};
  if (hl_ok && !fused) {
    const Dataset& hd = M.ds[R.first_dataset];
    hlwork.alloc(hl_work_doubles(hd.groups, hd.group_size, C));
    act_idx.alloc(C);
    n_act.alloc(1);
    hla.block = M.d_data.p + M.ds_off[R.first_dataset];
    hla.idx = act_idx.p; hla.n_dev = n_act.p; hla.n_rows = C; hla.work = hlwork.p; hla.jac = 1.0;
    if (R.flow) {  // packed rows x = T(z) in, packed gradient out
      hla.idx_in = false; hla.idx_out = false;
    } else {
      hla.Q = Q.p; hla.ldq = Dp[0]; hla.idx_in = true; hla.idx_out = true;
      hla.Gout = G.p; hla.ldg = Dp[0]; hla.Vout = V.p;
    }
  }
  const int hl_groups = hl_ok ? M.ds[R.first_dataset].groups : 0;
  DevBuf<double> fX, fGX, fVX, fLD;
  if (R.flow) {
    if (!act_idx.p) {
      act_idx.alloc(C);
      n_act.alloc(1);
    }
    fX.alloc((size_t)C * D[0]); fGX.alloc((size_t)C * D[0]); fVX.alloc(C); fLD.alloc(C);
    if (hl_ok) {
      hla.Q = fX.p; hla.ldq = D[0]; hla.Gout = fGX.p; hla.ldg = D[0]; hla.Vout = fVX.p;
    }
  }
  // flow: the GLM gradients read the packed rows x = T(z) and write packed
  // rows; the identity list maps position r to row r
  DevBuf<int> ident;
  if (R.flow && (glm_ok || gl_ok)) {
    std::vector<int> h(C);
    std::iota(h.begin(), h.end(), 0);
    ident.upload(h, st);
  }
  if (!fused && M.plugin && !act_idx.p) {
    act_idx.alloc(C);
    n_act.alloc(1);
  }
  GlLargeArgs gla;
  GlLargePlan glp;
  GlScratch gls;
  if (!fused && gl_ok) {
    act_idx.alloc(C);
    n_act.alloc(1);
    glp = gl_large_plan(M.glmN, C, M.glmD, num_cus());
    gl_bind(M, gla, glp, gls, cfg.precision, st);
    gla.C = C; gla.Q = Q.p; gla.ldq = Dp[0]; gla.G = G.p; gla.V = V.p; gla.idx = act_idx.p; gla.n_active = n_act.p;
    if (R.flow) {
      gla.Q = fX.p; gla.ldq = D[0]; gla.G = fGX.p; gla.V = fVX.p; gla.idx = ident.p;
    }
    xlow = M.low_x(cfg.precision, st);
  }
  if (!fused && glm_ok) {
    act_idx.alloc(C);
    n_act.alloc(1);
    gplan = glm_mfma_plan(M.glmN, C, M.glmD, num_cus(), tune.blocks_per_cu);
    work.alloc(gplan.work_doubles);
    ga.X = M.d_X; ga.y = M.d_y; ga.N = M.glmN; ga.D = M.glmD; ga.C = C; ga.rows_per_chunk = gplan.rows_per_chunk;
    ga.s2inv = M.glm_s2inv; ga.Q = Q.p; ga.ldq = Dp[0]; ga.G = G.p; ga.V = V.p; ga.work = work.p;
    ga.idx = act_idx.p; ga.n_active = n_act.p;
    if (R.flow) {
      ga.Q = fX.p; ga.ldq = D[0]; ga.G = fGX.p; ga.V = fVX.p; ga.idx = ident.p;
    }
    if (cfg.precision != 0) {
      if (tune.path == 2) xbf = M.bf16_x(cfg.precision, st);
      else xlow = M.low_x(cfg.precision, st);
    }
  }
  // the GLM gradient of the chosen kernel (FP64, FP32 or BF16-split)
  auto glm_grad = [&]() {
    const int rcg = cfg.precision == 0 ? glm_grad_mfma_f64(ga, gplan, st)
                    : xbf ? glm_grad_mfma_bf16(ga, gplan, xbf, cfg.precision, st)
                          : glm_grad_mfma_f32(ga, gplan, xlow, cfg.precision, kF32Layout, st);
    if (rcg) throw std::runtime_error("GLM kernel: unsupported D");
  };

  RunResult res;
  res.rows = rows;
  res.tuned_path = res_tuned_path;
  res.tuned_bpc = res_tuned_bpc;
  int stopped = 0;
  // one event pair per launch of a check interval brackets the gradient's
  // contraction kernel (the whole gradient kernel off the GLM path)
  constexpr int kCheck = 16;
  struct EvPairs {
    hipEvent_t a[kCheck] = {}, b[kCheck] = {};
    ~EvPairs() {
      for (int k = 0; k < kCheck; ++k) {
        if (a[k]) (void)hipEventDestroy(a[k]);
        if (b[k]) (void)hipEventDestroy(b[k]);
      }
    }
  } ev;
  EvPairs evs;  // around the state-machine launches of the tuning intervals
  // four timed intervals in the order wave, block, block, wave: each mode's
  // figure is the mean of its two interval medians, so a launch cost that
  // changes steadily over the first intervals (the step size and the trees
  // grow through early warmup) weighs on both modes alike
  int tune_left = tune_threads ? 4 : 0;
  double tune_s[2] = {0, 0};
  DevBuf<long long> tally;
  if (!fused) {
    for (int k = 0; k < kCheck; ++k) {
      HIPC(hipEventCreate(&ev.a[k]));
      HIPC(hipEventCreate(&ev.b[k]));
      if (tune_threads) {
        HIPC(hipEventCreate(&evs.a[k]));
        HIPC(hipEventCreate(&evs.b[k]));
      }
    }
    tally.alloc(1);
    HIPC(hipMemsetAsync(tally.p, 0, sizeof(long long), st));
  }
  const char* lim = std::getenv("CMDSTAN_AMD_MAX_SECONDS");
  const double max_seconds = lim ? std::atof(lim) : 3600.0;
  const auto t_begin = std::chrono::steady_clock::now();
  DevBuf<int> miter(1);
  int min_it = 0;
  int next_report = cfg.refresh > 0 ? cfg.refresh : 0;
  // One check interval of lockstep work: kCheck x (advance, gradient).  With
  // events, each gradient's contraction (or whole gradient) is bracketed.
  auto enqueue_bundle = [&](bool with_events, bool time_advance = false) {
    // active chains are tallied over the timed launches only (the roofline pairs them with their times)
    long long* tal = with_events ? tally.p : nullptr;
    for (int k = 0; k < kCheck; ++k) {
      if (time_advance) HIPC(hipEventRecord(evs.a[k], st));
      if (block_per_chain) launch_advance_external_block(cfg.metric, P, B, st);
      else launch_advance_external(cfg.metric, P, B, st);
      if (time_advance) HIPC(hipEventRecord(evs.b[k], st));
      if (R.flow) {
        // pulled-back gradient: x = T(z) for the active chains, model at x, VJP to z
        launch_compact_active(dneed.p, C, act_idx.p, n_act.p, tal, st);
        if (with_events) HIPC(hipEventRecord(ev.a[k], st));
        R.flow->forward(Q.p, Dp[0], act_idx.p, n_act.p, C, fX.p, fLD.p);
        if (hl_ok) {
          hl_grad_launch(hla, hl_groups, st);
        } else if (glm_ok) {
          ga.n_active = n_act.p;
          glm_grad();
        } else if (gl_ok) {
          gl_large_grad(gla, glp, xlow, cfg.precision, st);
        } else if (M.plugin) {
          if (M.pl.log_prob_grad(M.pl.user, fX.p, D[0], nullptr, n_act.p, C, 1, fGX.p, fVX.p, st))
            throw std::runtime_error("plugin log_prob_grad failed");
        } else {
          launch_model_grad_rows(M.kind, act_idx.p, n_act.p, C, D[0], fX.p, fGX.p, fVX.p, md, st);
        }
        R.flow->pullback(fGX.p, fVX.p, fLD.p, act_idx.p, n_act.p, C, G.p, Dp[0], V.p);
        if (with_events) HIPC(hipEventRecord(ev.b[k], st));
      } else if (glm_ok) {
        launch_compact_active(dneed.p, C, act_idx.p, n_act.p, tal, st);
        ga.ev_begin = with_events ? ev.a[k] : nullptr;
        ga.ev_end = with_events ? ev.b[k] : nullptr;
        glm_grad();
      } else if (gl_ok) {
        launch_compact_active(dneed.p, C, act_idx.p, n_act.p, tal, st);
        gla.ev_begin = with_events ? ev.a[k] : nullptr;
        gla.ev_end = with_events ? ev.b[k] : nullptr;
        gl_large_grad(gla, glp, xlow, cfg.precision, st);
      } else if (hl_ok) {
        launch_compact_active(dneed.p, C, act_idx.p, n_act.p, tal, st);
        if (with_events) HIPC(hipEventRecord(ev.a[k], st));
        hl_grad_launch(hla, hl_groups, st);
        if (with_events) HIPC(hipEventRecord(ev.b[k], st));
      } else if (M.plugin) {
        launch_compact_active(dneed.p, C, act_idx.p, n_act.p, tal, st);
        if (with_events) HIPC(hipEventRecord(ev.a[k], st));
        if (M.pl.log_prob_grad(M.pl.user, Q.p, Dp[0], act_idx.p, n_act.p, C, 1, G.p, V.p, st))
          throw std::runtime_error("plugin log_prob_grad failed");
        if (with_events) HIPC(hipEventRecord(ev.b[k], st));
      } else {
        if (with_events) HIPC(hipEventRecord(ev.a[k], st));
        launch_model_grad(M.kind, C, dD.p, doffb.p, Q.p, G.p, V.p, dneed.p, md, st);
        if (with_events) HIPC(hipEventRecord(ev.b[k], st));
      }
    }
  };
  // Lockstep launch sequences are replayed from a captured HIP graph (one per
  // phase: the phase's parameters are baked in); default on, except for
  // plugin models (their gradient callback is the caller's code).  One bundle
  // in kTimeEvery is launched eagerly with events around the gradient's
  // contraction kernel (the whole gradient off the GLM path): the kernel
  // timing comes from those, outside the captured graph.
  const char* genv = std::getenv("CMDSTAN_AMD_GRAPH");
  const bool use_graph = !fused && (genv ? std::atoi(genv) != 0 : !M.plugin);
  constexpr int kTimeEvery = 8;
  int bundle_no = 0;
  struct GraphHold {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    void reset() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
      x = nullptr; g = nullptr;
    }
    ~GraphHold() { reset(); }
  } graph;
  auto run_phase = [&]() {
    graph.reset();
    for (;;) {
      if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count() > max_seconds)
        throw std::runtime_error("sampler did not finish within CMDSTAN_AMD_MAX_SECONDS");
      if (cfg.interrupt && cfg.interrupt(cfg.interrupt_user)) throw std::runtime_error("sampling interrupted");
      bool timed = false;
      int tuning = -1;  // mode timed in this interval (0 wave, 1 block)
      if (fused) {
        const bool blk = fused_block_ok && (R.fused_block == 1 || C - stopped <= fused_block_chains);
        if (!(blk && launch_advance_fused_block(M.kind, cfg.metric, vec32, P, B, md, st)))
          launch_advance_fused(M.kind, cfg.metric, vec32, P, B, md, st);
      } else if (tune_left > 0) {
        tuning = (tune_left == 4 || tune_left == 1) ? 0 : 1;
        block_per_chain = tuning == 1;
        enqueue_bundle(true, true);
        res.leaf_steps += kCheck;
        timed = true;
      } else if (use_graph && bundle_no++ % kTimeEvery != 0) {
        if (!graph.x) {
          HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          enqueue_bundle(false);
          HIPC(hipStreamEndCapture(st, &graph.g));
          HIPC(hipGraphInstantiate(&graph.x, graph.g, nullptr, nullptr, 0));
        }
        HIPC(hipGraphLaunch(graph.x, st));
        res.leaf_steps += kCheck;
      } else {
        enqueue_bundle(true);
        res.leaf_steps += kCheck;
        timed = true;
      }
      HIPC(hipGetLastError());
      HIPC(hipMemcpyAsync(&stopped, dstop.p, sizeof(int), hipMemcpyDeviceToHost, st));
      if (cfg.refresh > 0) {
        hipLaunchKernelGGL(min_iter_kernel, dim3(1), dim3(256), 0, st, dst.p, C, miter.p);
        HIPC(hipMemcpyAsync(&min_it, miter.p, sizeof(int), hipMemcpyDeviceToHost, st));
      }
      HIPC(hipStreamSynchronize(st));
      if (cfg.refresh > 0) {
        // generate_transitions' progress line, for the slowest chain
        const int total = cfg.num_warmup + cfg.num_samples;
        while (next_report <= min_it && next_report <= total) {
          std::fprintf(stderr, "Iteration: %*d / %d [%3d%%]  (%s)\n", (int)std::to_string(total).size(), next_report,
                       total, (int)(100.0 * next_report / std::max(1, total)),
                       next_report <= cfg.num_warmup ? "Warmup" : "Sampling");
          std::fflush(stderr);
          next_report += cfg.refresh;
        }
      }
      if (timed) {
        for (int k = 0; k < kCheck; ++k) {
          float ms = 0;
          HIPC(hipEventElapsedTime(&ms, ev.a[k], ev.b[k]));
          res.grad_s += ms * 1e-3;
          res.grad_timed++;
        }
      }
      if (tuning >= 0) {
        // median launch of the interval, its first launch left out (it can
        // carry the one-time cost of the kernel's first use)
        float t[kCheck - 1];
        for (int k = 1; k < kCheck; ++k) HIPC(hipEventElapsedTime(&t[k - 1], evs.a[k], evs.b[k]));
        std::nth_element(t, t + (kCheck - 1) / 2, t + kCheck - 1);
        tune_s[tuning] += 0.5 * t[(kCheck - 1) / 2] * 1e-3;
        if (--tune_left == 0) {
          block_per_chain = default_block ? !(tune_s[0] < 0.8 * tune_s[1]) : (tune_s[1] < 0.8 * tune_s[0]);
          res.wave_launch_s = tune_s[0];
          res.block_launch_s = tune_s[1];
        }
      }
      if (stopped >= C) break;
    }
  };

  HIPC(hipStreamSynchronize(st));
  auto t0 = std::chrono::steady_clock::now();
  run_phase();
  if (R.flow) R.flow->check_barriers();
  auto t1 = std::chrono::steady_clock::now();
  // resume paused chains for the sampling phase
  {
    std::vector<ChainState> cur(C);
    HIPC(hipMemcpyAsync(cur.data(), dst.p, C * sizeof(ChainState), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    int done = 0;
    for (auto& s : cur) if (s.phase == PH_DONE || s.phase == PH_FAILED) ++done;
    HIPC(hipMemcpyAsync(dstop.p, &done, sizeof(int), hipMemcpyHostToDevice, st));
    stopped = done;  // chains still to run in the sampling phase: C - done
  }
  hipLaunchKernelGGL(resume_kernel, dim3((C + 255) / 256), dim3(256), 0, st, dst.p, C);
  P.iter_target = cfg.num_warmup + cfg.num_samples;
  HIPC(hipStreamSynchronize(st));
  auto t2 = std::chrono::steady_clock::now();
  run_phase();
  if (R.flow) R.flow->check_barriers();
  auto t3 = std::chrono::steady_clock::now();
  res.warm_s = std::chrono::duration<double>(t1 - t0).count();
  res.samp_s = std::chrono::duration<double>(t3 - t2).count();
  res.threads_per_chain = fused ? 0 : (block_per_chain ? 256 : 64);
  if (tally.p) {
    HIPC(hipMemcpyAsync(&res.grad_chains, tally.p, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
  }

  res.st.resize(C);
  HIPC(hipMemcpyAsync(res.st.data(), dst.p, C * sizeof(ChainState), hipMemcpyDeviceToHost, st));
  if (dense) {
    std::vector<double> hm(mtot);
    HIPC(hipMemcpyAsync(hm.data(), dMb.p, mtot * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    for (int c = 0; c < C; ++c)
      for (int i = 0; i < D[c]; ++i)
        for (int j = 0; j < D[c]; ++j) res.minv.push_back(hm[moff[c] + (int64_t)j * D[c] + i]);
  } else if (vec32) {
    std::vector<float> hf(tot);
    HIPC(hipMemcpyAsync(hf.data(), minv.p, tot * sizeof(float), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    for (int c = 0; c < C; ++c) for (int i = 0; i < D[c]; ++i) res.minv.push_back(hf[off[c] + i]);
  } else {
    std::vector<double> hmo(tot);
    HIPC(hipMemcpyAsync(hmo.data(), minv.p, tot * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    for (int c = 0; c < C; ++c) for (int i = 0; i < D[c]; ++i) res.minv.push_back(hmo[off[c] + i]);
  }
  if (rows > 0 && (msp < 0 || msp >= D[0])) {
    bool uni = true;
    for (int c = 1; c < C; ++c) uni = uni && D[c] == D[0];
    if (uni) {
      // last saved row of each chain, unconstrained (the state a later run resumes from)
      const int ld = 7 + nsave_of(D[0]);
      std::vector<double> last((size_t)C * ld);
      for (int c = 0; c < C; ++c)
        HIPC(hipMemcpyAsync(last.data() + (size_t)c * ld, draws.p + doff[c] + (int64_t)(rows - 1) * ld,
                            ld * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPC(hipStreamSynchronize(st));
      res.z_last.resize((size_t)C * D[0]);
      for (int c = 0; c < C; ++c)
        for (int i = 0; i < D[0]; ++i) res.z_last[(size_t)c * D[0] + i] = last[(size_t)c * ld + 7 + i];
    }
  }
  if (R.flow) {
    // saved rows hold z: replace it by x = T(z) and lp__ by log p(x) = lp(z) - log|det dT/dz|
    if (msp >= 0 && msp < D[0]) throw std::invalid_argument("flow: max_saved_params must keep every dimension");
    const int64_t nrows = (int64_t)C * rows, ld = 7 + D[0];
    const int R0 = R.flow->max_rows();
    for (int64_t k0 = 0; k0 < nrows; k0 += R0) {
      const int n = (int)std::min<int64_t>(R0, nrows - k0);
      R.flow->map_rows(draws.p + k0 * ld, ld, n);
    }
  }
  if (R.constrain && msp < 0 && M.plugin) {
    // write_array on the host: unconstrained rows down, constrained outputs by callback
    std::vector<double> un(dtot);
    HIPC(hipMemcpyAsync(un.data(), draws.p, dtot * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    res.draws.assign(otot, 0.0);
    std::vector<double> q((size_t)rows * D[0]), o((size_t)rows * nout[0]);
    for (int c = 0; c < C; ++c) {
      const double* src = un.data() + doff[c];
      double* dst = res.draws.data() + ooff[c];
      for (int r = 0; r < rows; ++r)
        for (int i = 0; i < D[c]; ++i) q[(size_t)r * D[c] + i] = src[(int64_t)r * (7 + D[c]) + 7 + i];
      if (M.pl.write_array) {
        if (M.pl.write_array(M.pl.user, q.data(), rows, o.data())) throw std::runtime_error("plugin write_array failed");
      } else {
        o = q;
      }
      for (int r = 0; r < rows; ++r) {
        for (int j = 0; j < 7; ++j) dst[(int64_t)r * (7 + nout[c]) + j] = src[(int64_t)r * (7 + D[c]) + j];
        for (int j = 0; j < nout[c]; ++j) dst[(int64_t)r * (7 + nout[c]) + 7 + j] = o[(size_t)r * nout[c] + j];
      }
    }
    res.draw_off = ooff;
    for (int c = 0; c < C; ++c) res.cols.push_back(7 + nout[c]);
  } else if (R.constrain && msp < 0) {
    DevBuf<double> out(std::max<int64_t>(otot, 1));
    DevBuf<int> dn(C);
    dn.upload(nout, st);
    launch_write_array(M.kind, C, rows, dD.p, ddoff.p, dooff.p, dn.p, draws.p, out.p, md, st);
    res.draws.resize(otot);
    HIPC(hipMemcpyAsync(res.draws.data(), out.p, otot * sizeof(double), hipMemcpyDeviceToHost, st));
    res.draw_off = ooff;
    for (int c = 0; c < C; ++c) res.cols.push_back(7 + nout[c]);
  } else {
    res.draws.resize(dtot);
    HIPC(hipMemcpyAsync(res.draws.data(), draws.p, dtot * sizeof(double), hipMemcpyDeviceToHost, st));
    res.draw_off = doff;
    for (int c = 0; c < C; ++c) res.cols.push_back(7 + nsave_of(D[c]));
  }
  if (R.diagnostics) {
    res.diag.resize(diag_tot);
    HIPC(hipMemcpyAsync(res.diag.data(), ddiag.p, diag_tot * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  if (B.eerr) {
    res.eerr.resize((size_t)C * rows);
    HIPC(hipMemcpyAsync(res.eerr.data(), deerr.p, res.eerr.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  HIPC(hipStreamSynchronize(st));
  HIPC(hipGetLastError());
  return res;
}

// min bulk-ESS over the D unconstrained coordinates of rows [r0, r1) of every
// chain, divided by the leapfrog steps those rows took
double pilot_ess_per_grad(const RunResult& r, int r0, int r1, int D, double* rhat_max = nullptr) {
  const int C = (int)r.st.size();
  const int n = r1 - r0;
  if (n < 4) return 0.0;
  std::vector<double> dr((size_t)D * C * n);
  double grads = 0;
  for (int c = 0; c < C; ++c) {
    const double* base = r.draws.data() + r.draw_off[c];
    const int ld = r.cols[c];
    for (int k = 0; k < n; ++k) {
      const double* row = base + (int64_t)(r0 + k) * ld;
      grads += row[4];
      for (int p = 0; p < D; ++p) dr[((size_t)p * C + c) * n + k] = row[7 + p];
    }
  }
  std::vector<double> st((size_t)D * 10);
  if (nuts_summary(dr.data(), n, C, D, 16, st.data()) != NUTS_OK) throw std::runtime_error("pilot summary failed");
  double mn = std::numeric_limits<double>::infinity();
  double rh = 0.0;
  for (int p = 0; p < D; ++p) {
    const double e = st[(size_t)p * 10 + 7];
    if (std::isfinite(e)) mn = std::min(mn, e);
    const double rr = st[(size_t)p * 10 + 9];
    if (std::isfinite(rr)) rh = std::max(rh, rr);
  }
  if (rhat_max) *rhat_max = rh;
  if (!std::isfinite(mn) || grads <= 0) return 0.0;
  return mn / grads;
}

bool all_done(const RunResult& r) {
  for (const auto& s : r.st) if (s.phase != PH_DONE) return false;
  return true;
}

struct FlowReport {
  int used = 0;
  double train_s = 0, loss = 0, egp = 0, egf = 0, esp = 0, esf = 0, rhp = 0, rhf = 0, pre_s = 0;
  std::vector<int64_t> nlf, ngrad;  // per chain, over every phase
  std::vector<int> rej_lp, rej_grad;  // initial values rejected in the plain phase
};

// Plain warmup + pilot, flow training, flow warmup + pilot, then the better
// sampler continues for num_samples iterations (no adaptation).
RunResult run_with_flow(ModelImpl& M, const RunSpec& R0, FlowReport& rep) {
  const nuts_config& cfg = R0.cfg;
  if (M.ds.size() != 1) throw std::invalid_argument("flow: the model must have one dataset");
  const Dataset& ds = M.ds[0];
  const int D = ds.D, C = cfg.num_chains, P = cfg.flow_pilot;
  if (D < 2) throw std::invalid_argument("flow: the model needs at least 2 dimensions");
  if (R0.inj_u || R0.inj_z || R0.stepsize_per_chain) throw std::invalid_argument("flow: unsupported run options");
  auto add_counts = [&](const RunResult& r) {
    if (rep.nlf.empty()) { rep.nlf.assign(C, 0); rep.ngrad.assign(C, 0); }
    for (int c = 0; c < C; ++c) { rep.nlf[c] += r.st[c].n_leapfrog_total; rep.ngrad[c] += r.st[c].n_grad; }
  };
  // phase A: plain NUTS, adaptive warmup + pilot, every row saved unconstrained
  RunSpec A = R0;
  A.cfg.flow = 0; A.cfg.save_warmup = 1; A.cfg.thin = 1; A.cfg.num_samples = P; A.cfg.max_saved_params = -1;
  A.constrain = false;
  A.energy_error = false;
  RunResult ra = run_batch(M, A);
  add_counts(ra);
  for (int c = 0; c < C; ++c) {
    rep.rej_lp.push_back(ra.st[c].init_rej_lp);
    rep.rej_grad.push_back(ra.st[c].init_rej_grad);
  }
  if (!all_done(ra)) return ra;
  const int rowsA = cfg.num_warmup + P;
  rep.egp = pilot_ess_per_grad(ra, cfg.num_warmup, rowsA, D, &rep.rhp);
  // training set: second half of warmup and the pilot
  const int first = cfg.num_warmup / 2;
  const int64_t ntr = (int64_t)C * (rowsA - first);
  std::vector<double> tr((size_t)ntr * D);
  std::vector<double> mean(D, 0.0), m2(D, 0.0);
  int64_t k = 0;
  for (int c = 0; c < C; ++c) {
    const double* base = ra.draws.data() + ra.draw_off[c];
    for (int r = first; r < rowsA; ++r, ++k) {
      const double* row = base + (int64_t)r * ra.cols[c] + 7;
      for (int i = 0; i < D; ++i) {
        tr[(size_t)k * D + i] = row[i];
        const double d = row[i] - mean[i];
        mean[i] += d / (double)(k + 1);
        m2[i] += d * (row[i] - mean[i]);
      }
    }
  }
  std::vector<double> sd(D);
  for (int i = 0; i < D; ++i) sd[i] = std::max(std::sqrt(m2[i] / std::max<int64_t>(1, ntr - 1)), 1e-8);
  FlowSpec spec = flow_default_spec(D, cfg.flow_layers, cfg.flow_hidden, ds.n_glob, ds.group_size);
  FlowNet net(spec, std::max(C, cfg.flow_batch));
  net.init_params(cfg.seed * 0x9E3779B97F4A7C15ull + 17);
  net.set_standardization(mean.data(), sd.data());
  {
    DevBuf<double> dtr((size_t)ntr * D);
    dtr.upload(tr, net.own_stream());
    HIPC(hipStreamSynchronize(net.own_stream()));
    const auto t0 = std::chrono::steady_clock::now();
    rep.loss = net.train(dtr.p, ntr, cfg.flow_train_steps, cfg.flow_batch, (float)cfg.flow_lr,
                         cfg.seed + 0x5851F42D4C957F2Dull, std::min(20, std::max(1, cfg.flow_train_steps)));
    rep.train_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  // phase B: flow NUTS from z0 = T^{-1}(last pilot draw), adaptive warmup + pilot
  std::vector<double> z0((size_t)C * D);
  {
    DevBuf<double> dx((size_t)C * D), dz((size_t)C * D);
    dx.upload(ra.z_last, net.own_stream());
    net.inverse(dx.p, C, dz.p, nullptr);
    HIPC(hipMemcpyAsync(z0.data(), dz.p, z0.size() * sizeof(double), hipMemcpyDeviceToHost, net.own_stream()));
    HIPC(hipStreamSynchronize(net.own_stream()));
  }
  RunSpec Bs = R0;
  Bs.cfg.flow = 0; Bs.cfg.num_warmup = cfg.flow_warmup; Bs.cfg.num_samples = P; Bs.cfg.save_warmup = 0;
  Bs.cfg.thin = 1; Bs.cfg.max_saved_params = -1;
  Bs.init = z0.data(); Bs.inv_metric = nullptr; Bs.flow = &net; Bs.constrain = false;
  Bs.energy_error = false;
  RunResult rb = run_batch(M, Bs);
  add_counts(rb);
  const bool flow_ok = all_done(rb);
  rep.egf = flow_ok ? pilot_ess_per_grad(rb, 0, P, D, &rep.rhf) : 0.0;
  // the flow must raise min-ESS per gradient, and pay for its own cost: the
  // pilot's min-ESS per second must rise too (gradients through the flow are
  // dearer than plain ones)
  auto pilot_grads = [&](const RunResult& r, int r0, int r1) {
    double gsum = 0;
    for (int c = 0; c < C; ++c) {
      const double* base = r.draws.data() + r.draw_off[c];
      for (int k = r0; k < r1; ++k) gsum += base[(int64_t)k * r.cols[c] + 4];
    }
    return gsum;
  };
  rep.esp = ra.samp_s > 0 ? rep.egp * pilot_grads(ra, cfg.num_warmup, rowsA) / ra.samp_s : 0.0;
  rep.esf = flow_ok && rb.samp_s > 0 ? rep.egf * pilot_grads(rb, 0, P) / rb.samp_s : 0.0;
  // ... unless plain NUTS is visibly not mixing (pilot R-hat above 1.05) and the
  // flow mixes better: then the speed of unconverged plain draws is no argument
  const bool plain_stuck = rep.rhp > 1.05 && rep.rhf < rep.rhp;
  rep.used = flow_ok && (cfg.flow == 2 || (rep.egf > rep.egp && (rep.esf > rep.esp || plain_stuck)));
  rep.pre_s = ra.warm_s + ra.samp_s + rep.train_s + rb.warm_s + rb.samp_s;
  // phase C: the better sampler continues without adaptation
  const RunResult& src = rep.used ? rb : ra;
  std::vector<double> eps(C);
  for (int c = 0; c < C; ++c) eps[c] = src.st[c].nom_eps;
  RunSpec Cs = R0;
  Cs.cfg.flow = 0; Cs.cfg.num_warmup = 0; Cs.cfg.adapt_engaged = 0; Cs.cfg.save_warmup = 0;
  Cs.init = src.z_last.data(); Cs.inv_metric = src.minv.data(); Cs.stepsize_per_chain = eps.data();
  Cs.flow = rep.used ? &net : nullptr;
  RunResult rc = run_batch(M, Cs);
  add_counts(rc);
  return rc;
}

const char* failure_text(int f) {
  switch (f) {
    case FAIL_INIT: return "Initialization failed.";
    case FAIL_STEPSIZE_IMPROPER: return "Posterior is improper. Please check your model.";
    case FAIL_STEPSIZE_ZERO: return "No acceptably small step size could be found. Perhaps the posterior is not continuous?";
    case FAIL_METRIC_OVERFLOW: return "Numerical overflow in metric adaptation.";
    case FAIL_METRIC_NOT_PD: return "Adapted dense metric is not positive definite.";
    default: return "";
  }
}

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return NUTS_NOT_OK;
  }
}

}  // namespace

struct nuts_model {
  ModelImpl impl;
};

extern "C" {

const char* nuts_last_error(void) { return g_err.c_str(); }
const char* nuts_version(void) { return "cmdstan_amd 0.1 (gfx950)"; }

nuts_data* nuts_data_create(void) { return new nuts_data(); }
void nuts_data_destroy(nuts_data* d) { delete d; }

int nuts_data_add(nuts_data* d, const char* name, int ndim, const int64_t* dims, const double* values, int is_int) {
  return guarded([&]() {
    Field f;
    int64_t n = 1;
    for (int i = 0; i < ndim; ++i) { f.dims.push_back(dims[i]); n *= dims[i]; }
    f.v.assign(values, values + n);
    f.is_int = is_int != 0;
    d->f[name] = std::move(f);
    return NUTS_OK;
  });
}

int nuts_data_parse_json(nuts_data* d, const char* text) {
  return guarded([&]() {
    std::string t(text);
    JsonParser p(t);
    JVal v = p.parse();
    if (v.kind != JVal::OBJ) throw std::invalid_argument("JSON data must be an object");
    for (auto& kv : v.obj) {
      Field f;
      bool all_int = true;
      json_flatten(kv.second, f.dims, f.v, all_int, kv.first);
      f.is_int = all_int;
      d->f[kv.first] = std::move(f);
    }
    return NUTS_OK;
  });
}

int nuts_data_parse_rdump(nuts_data* d, const char* text) {
  return guarded([&]() {
    std::string t(text);
    RDumpParser p(t);
    for (auto& v : p.parse()) {
      Field f;
      f.dims = v.dims;
      f.v = std::move(v.vals);
      f.is_int = v.is_int;
      d->f[v.name] = std::move(f);
    }
    return NUTS_OK;
  });
}

int nuts_data_get(const nuts_data* d, const char* name, int* ndim, int64_t* dims, const double** values) {
  return guarded([&]() {
    const Field& f = d->get(name);
    *ndim = (int)f.dims.size();
    for (size_t i = 0; i < f.dims.size() && i < 8; ++i) dims[i] = f.dims[i];
    *values = f.v.data();
    return NUTS_OK;
  });
}

nuts_model* nuts_model_create(const char* family, const nuts_data* const* datasets, int num_datasets) {
  try {
    if (num_datasets < 1) throw std::invalid_argument("num_datasets must be >= 1");
    std::unique_ptr<nuts_model> m(new nuts_model());
    ModelImpl& M = m->impl;
    M.family = family;
    std::vector<double> packed;
    for (int i = 0; i < num_datasets; ++i) {
      int kind = 0;
      M.ds.push_back(make_dataset(family, *datasets[i], kind));
      M.kind = kind;
      M.ds_off.push_back((int64_t)packed.size());
      packed.insert(packed.end(), M.ds.back().block.begin(), M.ds.back().block.end());
      M.ds.back().block.clear();
      M.ds.back().block.shrink_to_fit();
    }
    M.d_data.alloc(std::max<size_t>(packed.size(), 1));
    M.d_data.upload(packed.data(), packed.size());
    HIPC(hipDeviceSynchronize());
    if (M.kind == MK_LOGISTIC && num_datasets == 1) {
      M.glm = M.ds[0].D <= 128;
      M.glm_large = !M.glm;
      M.glmN = (int)packed[0];
      M.glmD = (int)packed[1];
      M.glm_s2inv = packed[2];
      M.d_X = M.d_data.p + 3;
      M.d_y = M.d_data.p + 3 + (size_t)M.glmN * M.glmD;
    }
    return m.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

nuts_model* nuts_model_create_plugin(const nuts_model_plugin* p) {
  try {
    if (!p || p->dim < 1 || p->num_outputs < 1 || !p->log_prob_grad || !p->param_names || !p->unconstrained_names)
      throw std::invalid_argument("plugin: dim, num_outputs, names and log_prob_grad are required");
    if (!p->write_array && p->num_outputs != p->dim)
      throw std::invalid_argument("plugin: without write_array the outputs are the unconstrained values");
    std::unique_ptr<nuts_model> m(new nuts_model());
    ModelImpl& M = m->impl;
    M.family = "plugin";
    M.kind = MK_PLUGIN;
    M.plugin = true;
    M.pl = *p;
    Dataset d;
    d.D = p->dim;
    d.nout = p->num_outputs;
    for (int i = 0; i < p->num_outputs; ++i) {
      if (!p->param_names[i]) throw std::invalid_argument("plugin: null parameter name");
      d.names.push_back(p->param_names[i]);
    }
    for (int i = 0; i < p->dim; ++i) {
      if (!p->unconstrained_names[i]) throw std::invalid_argument("plugin: null unconstrained name");
      d.unames.push_back(p->unconstrained_names[i]);
    }
    M.pl.param_names = nullptr;  // copied above; the caller may free its arrays
    M.pl.unconstrained_names = nullptr;
    M.ds.push_back(std::move(d));
    M.ds_off.push_back(0);
    M.d_data.alloc(1);
    return m.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void nuts_model_destroy(nuts_model* m) { delete m; }
int nuts_model_num_datasets(const nuts_model* m) { return (int)m->impl.ds.size(); }

int nuts_model_dims(const nuts_model* m, int dataset, int* num_unconstrained, int* num_outputs) {
  return guarded([&]() {
    if (dataset < 0 || dataset >= (int)m->impl.ds.size()) throw std::out_of_range("dataset index");
    if (num_unconstrained) *num_unconstrained = m->impl.ds[dataset].D;
    if (num_outputs) *num_outputs = m->impl.ds[dataset].nout;
    return NUTS_OK;
  });
}

static int copy_name(const std::string& s, char* buf, int buflen) {
  if ((int)s.size() + 1 > buflen) throw std::length_error("name buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return NUTS_OK;
}

int nuts_model_param_name(const nuts_model* m, int dataset, int i, char* buf, int buflen) {
  return guarded([&]() { return copy_name(m->impl.ds.at(dataset).names.at(i), buf, buflen); });
}
int nuts_model_unconstrained_name(const nuts_model* m, int dataset, int i, char* buf, int buflen) {
  return guarded([&]() { return copy_name(m->impl.ds.at(dataset).unames.at(i), buf, buflen); });
}

static int log_prob_grad_impl(nuts_model* m, int dataset, const double* q, int n_points, int jacobian,
                              int precision, double* lp, double* grad) {
  return guarded([&]() {
    ModelImpl& M = m->impl;
    if (dataset < 0 || dataset >= (int)M.ds.size()) throw std::out_of_range("dataset index");
    if (n_points < 1) return NUTS_OK;
    const int D = M.ds[dataset].D;
    const int Dp = (D + 15) / 16 * 16;
    const int C = n_points;
    std::vector<int> Dv(C, D);
    std::vector<int64_t> off(C), dso(C, M.ds_off[dataset]);
    std::vector<double> hq((size_t)C * Dp, 0.0);
    for (int c = 0; c < C; ++c) {
      off[c] = (int64_t)c * Dp;
      for (int i = 0; i < D; ++i) hq[(size_t)c * Dp + i] = q[(size_t)c * D + i];
    }
    DevBuf<int> dD(C);
    DevBuf<int64_t> doff(C), ddso(C);
    DevBuf<double> Q(hq.size()), G(hq.size()), V(C);
    dD.upload(Dv); doff.upload(off); ddso.upload(dso); Q.upload(hq);
    ModelDataDev md{M.d_data.p, ddso.p, jacobian ? 1.0 : 0.0};
    if (precision != 0 && !((M.glm || M.glm_large) && dataset == 0))
      throw std::invalid_argument("reduced precision is available for the batched logistic regression only");
    if (M.plugin) {
      if (M.pl.log_prob_grad(M.pl.user, Q.p, Dp, nullptr, nullptr, C, jacobian ? 1 : 0, G.p, V.p, nullptr))
        throw std::runtime_error("plugin log_prob_grad failed");
      HIPC(hipDeviceSynchronize());
    } else if (M.glm_large && dataset == 0 && precision != 0) {
      const GlLargePlan gp = gl_large_plan(M.glmN, C, D, num_cus());
      GlScratch gs;
      GlLargeArgs a;
      gl_bind(M, a, gp, gs, precision, 0);
      a.C = C; a.Q = Q.p; a.ldq = Dp; a.G = G.p; a.V = V.p;
      gl_large_grad(a, gp, M.low_x(precision, 0), precision, 0);
      HIPC(hipDeviceSynchronize());
    } else if (M.glm && dataset == 0) {
      // reduced precision: the kernel the sampler's autotuner picks for this many points
      const ModelImpl::GlmTune tn = precision != 0 ? glm_autotune(M, C, precision, 0) : ModelImpl::GlmTune{};
      if (tn.path == 1) {
        const GlLargePlan gp = gl_large_plan(M.glmN, C, D, num_cus());
        GlScratch gs;
        GlLargeArgs a;
        gl_bind(M, a, gp, gs, precision, 0);
        a.C = C; a.Q = Q.p; a.ldq = Dp; a.G = G.p; a.V = V.p;
        gl_large_grad(a, gp, M.low_x(precision, 0), precision, 0);
      } else {
        const GlmPlan gp = glm_mfma_plan(M.glmN, C, D, num_cus());
        DevBuf<double> work(gp.work_doubles);
        GlmArgs ga{M.d_X, M.d_y, M.glmN, D, C, gp.rows_per_chunk, M.glm_s2inv, Q.p, Dp, G.p, V.p, work.p};
        const int rcg = precision == 0 ? glm_grad_mfma_f64(ga, gp, 0)
                        : tn.path == 2 ? glm_grad_mfma_bf16(ga, gp, M.bf16_x(precision, 0), precision, 0)
                                       : glm_grad_mfma_f32(ga, gp, M.low_x(precision, 0), precision, kF32Layout, 0);
        if (rcg) throw std::runtime_error("GLM kernel: unsupported D");
      }
      HIPC(hipDeviceSynchronize());
    } else if (M.kind == MK_HIER_LOGISTIC) {
      // the sampler's batched group kernel (lane = point)
      const Dataset& hd = M.ds[dataset];
      DevBuf<double> work(hl_work_doubles(hd.groups, hd.group_size, C));
      HlArgs a{};
      a.block = M.d_data.p + M.ds_off[dataset];
      a.Q = Q.p; a.ldq = Dp; a.idx = nullptr; a.idx_in = false; a.idx_out = false; a.n_dev = nullptr; a.n_rows = C;
      a.Gout = G.p; a.ldg = Dp; a.Vout = V.p; a.work = work.p; a.jac = jacobian ? 1.0 : 0.0;
      hl_grad_launch(a, hd.groups, 0);
      HIPC(hipDeviceSynchronize());
    } else {
      launch_model_grad(M.kind, C, dD.p, doff.p, Q.p, G.p, V.p, nullptr, md, 0);
    }
    HIPC(hipGetLastError());
    std::vector<double> hg(hq.size()), hv(C);
    HIPC(hipMemcpy(hg.data(), G.p, hg.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hv.data(), V.p, C * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < C; ++c) {
      lp[c] = -hv[c];
      for (int i = 0; i < D; ++i) grad[(size_t)c * D + i] = -hg[(size_t)c * Dp + i];
    }
    return NUTS_OK;
  });
}

// stan::model::model_base::transform_inits for the registered families: the
// named constrained values of `inits` to the unconstrained vector (D values);
// parameters absent from `inits` are NaN (drawn at random by initialize).
int nuts_model_transform_inits(nuts_model* m, int dataset, const nuts_data* inits, double* out) {
  return guarded([&]() {
    ModelImpl& M = m->impl;
    if (dataset < 0 || dataset >= (int)M.ds.size()) throw std::out_of_range("dataset index");
    const Dataset& ds = M.ds[dataset];
    const int D = ds.D;
    for (int i = 0; i < D; ++i) out[i] = std::numeric_limits<double>::quiet_NaN();
    if (M.plugin) {
      if (!M.pl.transform_inits) throw std::invalid_argument("init: this plugin model has no transform_inits");
      if (M.pl.transform_inits(M.pl.user, inits, out)) throw std::invalid_argument("init: plugin transform_inits failed");
      return NUTS_OK;
    }
    auto take = [&](const std::string& name, int off, int n, int kind) {
      // kind 0 identity, 1 lower=0 (log), 2 (0,1) (logit)
      if (!inits->has(name)) return;
      const Field& f = inits->get(name);
      if ((int)f.v.size() != n)
        throw std::invalid_argument("init: " + name + " has " + std::to_string(f.v.size()) + " values, expected " +
                                    std::to_string(n));
      for (int i = 0; i < n; ++i) {
        const double v = f.v[i];
        if (!std::isfinite(v)) throw std::invalid_argument("init: " + name + " must be finite");
        if (kind == 1) {
          if (!(v > 0)) throw std::domain_error("init: " + name + " is " + std::to_string(v) + ", but must be > 0");
          out[off + i] = std::log(v);
        } else if (kind == 2) {
          if (!(v > 0 && v < 1))
            throw std::domain_error("init: " + name + " is " + std::to_string(v) + ", but must be in (0, 1)");
          out[off + i] = std::log(v / (1 - v));
        } else {
          out[off + i] = v;
        }
      }
    };
    switch (M.kind) {
      case MK_BERNOULLI: take("theta", 0, 1, 2); break;
      case MK_GAUSS: take("x", 0, D, 0); break;
      case MK_LOGISTIC: take("beta", 0, D, 0); break;
      case MK_EIGHT_SCHOOLS: take("mu", 0, 1, 0); take("tau", 1, 1, 1); take("theta", 2, D - 2, 0); break;
      case MK_HIER_NORMAL_NC: take("mu", 0, 1, 0); take("tau", 1, 1, 1); take("eta", 2, D - 2, 0); break;
      case MK_HIER_LOGISTIC: {
        const int K = ds.group_size;
        take("mu", 0, K, 0);
        take("tau", K, K, 1);
        take("beta", 2 * K, D - 2 * K, 0);  // beta[G][K], row-major
        break;
      }
      default: throw std::invalid_argument("init: unknown model family");
    }
    return NUTS_OK;
  });
}

int nuts_log_prob_grad(nuts_model* m, int dataset, const double* q, int n_points, int jacobian, double* lp,
                       double* grad) {
  return log_prob_grad_impl(m, dataset, q, n_points, jacobian, 0, lp, grad);
}

int nuts_log_prob_grad_precision(nuts_model* m, int dataset, const double* q, int n_points, int jacobian,
                                 int precision, double* lp, double* grad) {
  return log_prob_grad_impl(m, dataset, q, n_points, jacobian, precision, lp, grad);
}

void nuts_config_default(nuts_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->num_samples = 1000; c->num_warmup = 1000; c->save_warmup = 0; c->thin = 1;
  c->adapt_engaged = 1; c->gamma = 0.05; c->delta = 0.8; c->kappa = 0.75; c->t0 = 10;
  c->init_buffer = 75; c->term_buffer = 50; c->window = 25;
  c->max_depth = 10; c->metric = NUTS_METRIC_DIAG_E; c->stepsize = 1; c->stepsize_jitter = 0;
  c->num_chains = 1; c->id = 1; c->seed = 0; c->init_radius = 2; c->fused = -1; c->precision = 0;
  c->max_saved_params = -1;
  c->flow = 0; c->flow_layers = 8; c->flow_hidden = 128; c->flow_train_steps = 1000; c->flow_batch = 256;
  c->flow_lr = 1e-3; c->flow_warmup = 300; c->flow_pilot = 100;
  c->refresh = 0;
}

int nuts_config_validate(const nuts_config* cfg) {
  return guarded([&]() { validate(*cfg); return NUTS_OK; });
}

int nuts_output_rows(const nuts_config* cfg) { return output_rows(*cfg); }

int nuts_sample(nuts_model* m, const nuts_config* cfg, const double* init, const double* inv_metric,
                nuts_output* out) {
  return guarded([&]() {
    RunSpec R;
    R.cfg = *cfg;
    R.init = init;
    R.inv_metric = inv_metric;
    R.diagnostics = out && out->diagnostics;
    R.energy_error = out && out->energy_error;
    if (R.diagnostics && cfg->flow) throw std::invalid_argument("diagnostic rows are not available with the flow");
    if (init || inv_metric) {
      // per-chain arrays index with a uniform stride: require one dimension
      for (auto& d : m->impl.ds)
        if (d.D != m->impl.ds[0].D) throw std::invalid_argument("init/inv_metric arrays need equal dimensions");
    }
    FlowReport rep;
    RunResult r = cfg->flow ? run_with_flow(m->impl, R, rep) : run_batch(m->impl, R);
    const int C = (int)r.st.size();
    if (out) {
      if (out->draws) std::memcpy(out->draws, r.draws.data(), r.draws.size() * sizeof(double));
      int64_t moff = 0;
      for (int c = 0; c < C; ++c) {
        const ChainState& s = r.st[c];
        if (out->stepsize) out->stepsize[c] = s.nom_eps;
        if (out->n_leapfrog) out->n_leapfrog[c] = cfg->flow ? rep.nlf[c] : s.n_leapfrog_total;
        if (out->grad_evals) out->grad_evals[c] = cfg->flow ? rep.ngrad[c] : s.n_grad;
        if (out->status) out->status[c] = s.phase == PH_DONE ? 0 : (s.failure ? s.failure : -1);
      }
      if (out->inv_metric) std::memcpy(out->inv_metric, r.minv.data(), r.minv.size() * sizeof(double));
      if (out->diagnostics) std::memcpy(out->diagnostics, r.diag.data(), r.diag.size() * sizeof(double));
      if (out->energy_error) std::memcpy(out->energy_error, r.eerr.data(), r.eerr.size() * sizeof(double));
      if (out->init_rejections)
        for (int c = 0; c < C; ++c) {
          out->init_rejections[2 * c] = cfg->flow ? rep.rej_lp[c] : r.st[c].init_rej_lp;
          out->init_rejections[2 * c + 1] = cfg->flow ? rep.rej_grad[c] : r.st[c].init_rej_grad;
        }
      (void)moff;
      out->warmup_seconds = r.warm_s;
      out->sampling_seconds = r.samp_s;
      out->leaf_steps = r.leaf_steps;
      out->grad_seconds_per_launch = r.grad_timed ? r.grad_s / r.grad_timed : 0.0;
      out->grad_launches_timed = r.grad_timed;
      out->grad_chains_per_launch = r.grad_timed ? (double)r.grad_chains / r.grad_timed : 0.0;
      out->tuned_path = r.tuned_path;
      out->tuned_blocks_per_cu = r.tuned_bpc;
      out->tuned_threads_per_chain = r.threads_per_chain;
      out->tuned_wave_launch_seconds = r.wave_launch_s;
      out->tuned_block_launch_seconds = r.block_launch_s;
      if (cfg->flow) {
        out->warmup_seconds = rep.pre_s;  // plain warmup + pilot, training, flow warmup + pilot
        out->flow_used = rep.used;
        out->flow_train_seconds = rep.train_s;
        out->flow_loss = rep.loss;
        out->pilot_ess_per_grad_plain = rep.egp;
        out->pilot_ess_per_grad_flow = rep.egf;
        out->pilot_ess_per_sec_plain = rep.esp;
        out->pilot_ess_per_sec_flow = rep.esf;
        out->pilot_rhat_plain = rep.rhp;
        out->pilot_rhat_flow = rep.rhf;
      }
    }
    for (int c = 0; c < C; ++c)
      if (r.st[c].phase != PH_DONE) {
        g_err = std::string("chain ") + std::to_string(cfg->id + c) + ": " + failure_text(r.st[c].failure);
        return NUTS_NOT_OK;
      }
    return NUTS_OK;
  });
}

int nuts_transition_replay(nuts_model* m, int dataset, int num_chains, int metric, double stepsize, double jitter,
                           int max_depth, const double* inv_metric, const double* q0, const double* u, int64_t n_u,
                           const double* z, int64_t n_z, double* out_q, double* info) {
  return nuts_transition_replay_precision(m, dataset, num_chains, metric, stepsize, jitter, max_depth, inv_metric, q0,
                                          u, n_u, z, n_z, 0, out_q, info);
}

int nuts_transition_replay_precision(nuts_model* m, int dataset, int num_chains, int metric, double stepsize,
                                     double jitter, int max_depth, const double* inv_metric, const double* q0,
                                     const double* u, int64_t n_u, const double* z, int64_t n_z, int precision,
                                     double* out_q, double* info) {
  return nuts_transition_replay_driver(m, dataset, num_chains, metric, stepsize, jitter, max_depth, inv_metric, q0, u,
                                       n_u, z, n_z, precision, NUTS_DRIVER_FUSED, out_q, info);
}

int nuts_transition_replay_driver(nuts_model* m, int dataset, int num_chains, int metric, double stepsize,
                                  double jitter, int max_depth, const double* inv_metric, const double* q0,
                                  const double* u, int64_t n_u, const double* z, int64_t n_z, int precision,
                                  int driver, double* out_q, double* info) {
  return guarded([&]() {
    if (driver < NUTS_DRIVER_FUSED || driver > NUTS_DRIVER_FUSED_BLOCK)
      throw std::invalid_argument(
          "driver must be 0 (fused), 1 (lockstep, wave per chain), 2 (lockstep, block) or 3 (fused, block)");
    RunSpec R;
    nuts_config_default(&R.cfg);
    R.cfg.precision = precision;
    validate(R.cfg);
    R.cfg.num_warmup = 0; R.cfg.num_samples = 1; R.cfg.adapt_engaged = 0;
    R.cfg.metric = metric; R.cfg.stepsize = stepsize; R.cfg.stepsize_jitter = jitter;
    R.cfg.max_depth = max_depth; R.cfg.num_chains = num_chains;
    const bool fused_drv = driver == NUTS_DRIVER_FUSED || driver == NUTS_DRIVER_FUSED_BLOCK;
    R.cfg.fused = fused_drv ? 1 : 0;
    R.block_mode = fused_drv ? -1 : (driver == NUTS_DRIVER_LOCKSTEP_BLOCK ? 1 : 0);
    R.fused_block = driver == NUTS_DRIVER_FUSED_BLOCK ? 1 : (driver == NUTS_DRIVER_FUSED ? 0 : -1);
    R.init = q0; R.inv_metric = inv_metric;
    R.inj_u = u; R.n_u = n_u; R.inj_z = z; R.n_z = n_z;
    R.first_dataset = dataset; R.num_datasets = 1;
    R.constrain = false;
    RunResult r = run_batch(m->impl, R);
    const int D = m->impl.ds[dataset].D;
    for (int c = 0; c < num_chains; ++c) {
      const ChainState& s = r.st[c];
      if (s.phase != PH_DONE) throw std::runtime_error(std::string("replay failed: ") + failure_text(s.failure));
      const double* row = r.draws.data() + r.draw_off[c];
      for (int i = 0; i < D; ++i) out_q[(size_t)c * D + i] = row[7 + i];
      double* inf = info + (size_t)c * 9;
      inf[0] = row[3]; inf[1] = row[4]; inf[2] = row[5]; inf[3] = row[1]; inf[4] = row[6]; inf[5] = row[0];
      inf[6] = (double)s.uctr; inf[7] = (double)s.nctr; inf[8] = row[2];
    }
    return NUTS_OK;
  });
}

// -------------------------------------------- allocation-free log_prob_grad
// nuts_log_prob_grad_noalloc keeps its scratch (the potentials, the per-point
// dimension / offset tables and the gradient kernels' partial sums) in this
// file-scope device array: one copy per GPU, held until the process ends.
// Host code never passes the array's name to a kernel: it passes the per-device
// address that hipGetSymbolAddress returns.
#define NUTS_NOALLOC_SCRATCH_BYTES (48ull << 20)
#define NUTS_NOALLOC_MAX_POINTS 4096
static __device__ __align__(256) unsigned char scratch[NUTS_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;  // device -> address of its copy of `scratch`

// address of the current device's copy of `scratch` (looked up once per device)
unsigned char* scratch_for_current_device() {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw HipError(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  return static_cast<unsigned char*>(p);
}

// per-point tables of the per-chain model kernels: dimension, row offset of
// the point (rows of D doubles), dataset block offset
__global__ void noalloc_tables(int* Dv, int64_t* off, int64_t* dso, int n, int D, int64_t ds_off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Dv[c] = D;
  off[c] = (int64_t)c * D;
  dso[c] = ds_off;
}

// V = -lp and G = -grad lp (in place in grad) to lp and grad lp
__global__ void noalloc_finish(const double* V, double* lp, double* grad, int n, int D) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) lp[t] = -V[t];
  if (t < (int64_t)n * D) grad[t] = -grad[t];
}

}  // namespace

extern "C" int nuts_log_prob_grad_noalloc(nuts_model* m, int dataset, const double* d_q, int n_points, int jacobian,
                                          double* d_lp, double* d_grad, void* stream) {
  return guarded([&]() {
    ModelImpl& M = m->impl;
    if (dataset < 0 || dataset >= (int)M.ds.size()) throw std::out_of_range("dataset index");
    if (n_points < 0 || n_points > NUTS_NOALLOC_MAX_POINTS)
      throw std::invalid_argument("nuts_log_prob_grad_noalloc: n_points must be in [0, " +
                                  std::to_string(NUTS_NOALLOC_MAX_POINTS) + "]");
    if (n_points == 0) return NUTS_OK;
    if (!d_q || !d_lp || !d_grad) throw std::invalid_argument("nuts_log_prob_grad_noalloc: null device buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int D = M.ds[dataset].D, n = n_points;
    unsigned char* base = scratch_for_current_device();
    // scratch regions, 256-byte aligned: V [n], D [n], offsets [n], dataset offsets [n], kernel partials
    size_t at = 0;
    auto region = [&](size_t bytes) {
      unsigned char* p = base + at;
      at += (bytes + 255) & ~size_t(255);
      return p;
    };
    double* V = reinterpret_cast<double*>(region(sizeof(double) * n));
    int* Dv = reinterpret_cast<int*>(region(sizeof(int) * n));
    int64_t* off = reinterpret_cast<int64_t*>(region(sizeof(int64_t) * n));
    int64_t* dso = reinterpret_cast<int64_t*>(region(sizeof(int64_t) * n));
    double* work = reinterpret_cast<double*>(base + at);
    const size_t work_cap = (NUTS_NOALLOC_SCRATCH_BYTES - at) / sizeof(double);
    auto need = [&](size_t doubles) {
      if (doubles > work_cap)
        throw std::invalid_argument("nuts_log_prob_grad_noalloc: " + std::to_string(n) + " points need " +
                                    std::to_string(doubles * 8) + " bytes of scratch, more than the " +
                                    std::to_string(work_cap * 8) + " available; use fewer points");
    };
    if (M.plugin) {
      if (M.pl.log_prob_grad(M.pl.user, d_q, D, nullptr, nullptr, n, jacobian ? 1 : 0, d_grad, V, st))
        throw std::runtime_error("plugin log_prob_grad failed");
    } else if (M.glm && dataset == 0) {
      // the plan of nuts_log_prob_grad, so both give the same sums
      const GlmPlan gp = glm_mfma_plan(M.glmN, n, D, num_cus());
      need(gp.work_doubles);
      GlmArgs ga{M.d_X, M.d_y, M.glmN, D, n, gp.rows_per_chunk, M.glm_s2inv, d_q, D, d_grad, V, work};
      if (glm_grad_mfma_f64(ga, gp, st)) throw std::runtime_error("GLM kernel: unsupported D");
    } else if (M.kind == MK_HIER_LOGISTIC) {
      const Dataset& hd = M.ds[dataset];
      need(hl_work_doubles(hd.groups, hd.group_size, n));
      HlArgs a{};
      a.block = M.d_data.p + M.ds_off[dataset];
      a.Q = d_q; a.ldq = D; a.idx = nullptr; a.idx_in = false; a.idx_out = false; a.n_dev = nullptr; a.n_rows = n;
      a.Gout = d_grad; a.ldg = D; a.Vout = V; a.work = work; a.jac = jacobian ? 1.0 : 0.0;
      hl_grad_launch(a, hd.groups, st);
    } else {
      hipLaunchKernelGGL(noalloc_tables, dim3((n + 255) / 256), dim3(256), 0, st, Dv, off, dso, n, D,
                         M.ds_off[dataset]);
      ModelDataDev md{M.d_data.p, dso, jacobian ? 1.0 : 0.0};
      launch_model_grad(M.kind, n, Dv, off, d_q, d_grad, V, nullptr, md, st);
    }
    const int64_t tot = std::max<int64_t>((int64_t)n * D, n);
    hipLaunchKernelGGL(noalloc_finish, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, V, d_lp, d_grad, n, D);
    HIPC(hipGetLastError());
    return NUTS_OK;
  });
}

// ------------------------------------------------------------------ flow
struct nuts_flow {
  std::unique_ptr<FlowNet> net;
};

nuts_flow* nuts_flow_create(int D, int n_layers, int hidden, int n_glob, int group_size, int max_rows,
                            uint64_t seed) {
  try {
    if (n_glob < 0 || group_size < 1 || n_glob >= D) throw std::invalid_argument("flow: bad n_glob / group_size");
    auto f = std::make_unique<nuts_flow>();
    f->net = std::make_unique<FlowNet>(flow_default_spec(D, n_layers, hidden, n_glob, group_size), max_rows);
    f->net->init_params(seed);
    return f.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void nuts_flow_destroy(nuts_flow* f) { delete f; }

int64_t nuts_flow_num_params(const nuts_flow* f) { return f ? f->net->num_params() : -1; }

int nuts_flow_get_params(const nuts_flow* f, float* out) {
  return guarded([&]() { f->net->get_params(out); return NUTS_OK; });
}

int nuts_flow_set_params(nuts_flow* f, const float* in) {
  return guarded([&]() { f->net->set_params(in); return NUTS_OK; });
}

int nuts_flow_set_standardization(nuts_flow* f, const double* loc, const double* scale) {
  return guarded([&]() { f->net->set_standardization(loc, scale); return NUTS_OK; });
}

static void flow_rows_check(const nuts_flow* f, int n) {
  if (!f) throw std::invalid_argument("flow: null handle");
  if (n < 1 || n > f->net->max_rows()) throw std::invalid_argument("flow: n must be in [1, max_rows]");
}

int nuts_flow_forward(nuts_flow* f, const double* z, int n, double* x, double* logdet) {
  return guarded([&]() {
    flow_rows_check(f, n);
    const int D = f->net->dim();
    hipStream_t st = f->net->own_stream();
    DevBuf<double> dz((size_t)n * D), dx((size_t)n * D), dl(n);
    dz.upload(z, (size_t)n * D, st);
    f->net->forward(dz.p, D, nullptr, nullptr, n, dx.p, dl.p);
    HIPC(hipMemcpyAsync(x, dx.p, (size_t)n * D * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(logdet, dl.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    f->net->check_barriers();
    return NUTS_OK;
  });
}

int nuts_flow_inverse(nuts_flow* f, const double* x, int n, double* z, double* logdet) {
  return guarded([&]() {
    flow_rows_check(f, n);
    const int D = f->net->dim();
    hipStream_t st = f->net->own_stream();
    DevBuf<double> dz((size_t)n * D), dx((size_t)n * D), dl(n);
    dx.upload(x, (size_t)n * D, st);
    f->net->inverse(dx.p, n, dz.p, dl.p);
    HIPC(hipMemcpyAsync(z, dz.p, (size_t)n * D * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(logdet, dl.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return NUTS_OK;
  });
}

int nuts_flow_pullback_grad(nuts_flow* f, const double* z, const double* gx, int n, double* gz) {
  return guarded([&]() {
    flow_rows_check(f, n);
    const int D = f->net->dim();
    hipStream_t st = f->net->own_stream();
    DevBuf<double> dz((size_t)n * D), dx((size_t)n * D), dgx((size_t)n * D), dgz((size_t)n * D), dl(n), dvx(n), dv(n);
    dz.upload(z, (size_t)n * D, st);
    dgx.upload(gx, (size_t)n * D, st);
    HIPC(hipMemsetAsync(dvx.p, 0, n * sizeof(double), st));
    f->net->forward(dz.p, D, nullptr, nullptr, n, dx.p, dl.p);
    f->net->pullback(dgx.p, dvx.p, dl.p, nullptr, nullptr, n, dgz.p, D, dv.p);
    HIPC(hipMemcpyAsync(gz, dgz.p, (size_t)n * D * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    f->net->check_barriers();
    return NUTS_OK;
  });
}

int nuts_flow_loss_grad(nuts_flow* f, const double* x, int n, double* loss, float* grad) {
  return guarded([&]() {
    flow_rows_check(f, n);
    const int D = f->net->dim();
    hipStream_t st = f->net->own_stream();
    DevBuf<double> dx((size_t)n * D);
    dx.upload(x, (size_t)n * D, st);
    const double l = f->net->loss_grad(dx.p, n, grad);
    if (loss) *loss = l;
    return NUTS_OK;
  });
}

int nuts_flow_train(nuts_flow* f, const double* x, int64_t n, int steps, int batch, double lr, uint64_t seed,
                    double* final_loss) {
  return guarded([&]() {
    if (!f || n < 1 || steps < 0 || batch < 1 || batch > f->net->max_rows() || !(lr > 0))
      throw std::invalid_argument("flow: bad training arguments");
    const int D = f->net->dim();
    hipStream_t st = f->net->own_stream();
    DevBuf<double> dx((size_t)n * D);
    dx.upload(x, (size_t)n * D, st);
    HIPC(hipStreamSynchronize(st));
    const double l = f->net->train(dx.p, n, steps, batch, (float)lr, seed, std::min(20, std::max(1, steps)));
    if (final_loss) *final_loss = l;
    return NUTS_OK;
  });
}

}  // extern "C"
