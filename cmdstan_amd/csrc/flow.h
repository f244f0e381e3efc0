// Affine-coupling normalizing flow on the GPU: the learned reparameterisation
// NUTS samples through.  x = loc + scale * f_{L-1}(... f_0(z)); layer l keeps
// the dimensions a_l and maps b_l as v_b = u_b * exp(s) + t with
// (s_raw, t) = MLP(u_a) (two tanh hidden layers of width H),
// s = 3 tanh(s_raw / 3).  FP32 weights and activations; the contractions run
// on FP32 matrix cores; log-densities and log-determinants are summed in FP64.
//
// Published algorithms: RealNVP coupling (Dinh et al. 2017), Adam (Kingma &
// Ba 2015), NUTS on the pulled-back density (Hoffman et al. 2019).  CmdStan
// has no counterpart.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace nutsgpu {

struct ModelDataDev;
struct FpArgs;  // arguments of the persistent sampling kernels (flow.hip)

struct FlowSpec {
  int D = 0, L = 0, H = 0;
  std::vector<std::vector<int>> ia, ib;  // per layer: conditioner / transformed dimensions
};

// Coupling masks of a model with n_glob leading global dimensions and groups
// of group_size dimensions after them (the same rule as the CPU oracle).
FlowSpec flow_default_spec(int D, int L, int H, int n_glob, int group_size);

class FlowNet {
 public:
  FlowNet(const FlowSpec& spec, int max_rows);  // works on a stream of its own until bind_stream()
  ~FlowNet();
  FlowNet(const FlowNet&) = delete;
  FlowNet& operator=(const FlowNet&) = delete;

  int dim() const { return spec_.D; }
  int max_rows() const { return R_; }
  int64_t num_params() const { return n_params_; }
  const FlowSpec& spec() const { return spec_; }
  // Later calls run on s (e.g. the sampler's stream); returns the previous stream.
  hipStream_t bind_stream(hipStream_t s) { hipStream_t o = st_; st_ = s; return o; }
  hipStream_t own_stream() const { return own_st_; }

  // host <-> device parameters (FP32, layout: per layer W1 [Da][H], b1, W2 [H][H], b2, W3 [H][2Db], b3)
  void set_params(const float* host);
  void get_params(float* host) const;
  void init_params(uint64_t seed);  // Glorot-uniform W1, W2; zero W3, b3 (identity flow)
  void set_standardization(const double* loc_host, const double* scale_host);

  // Rows r < *n_dev (n_dev == nullptr: n_rows) of z taken from Q (row idx[r],
  // or r when idx == nullptr, stride ldq): X [r][D] = T(z), logdet[r].  The
  // activations stay for pullback().
  void forward(const double* Q, int64_t ldq, const int* idx, const int* n_dev, int n_rows, double* X,
               double* logdet);
  // After forward(): grad_z of V(T(z)) - log|det| from GX = grad V at X;
  // writes G[idx[r]] (stride ldg) and V[idx[r]] = VX[r] - logdet[r].
  void pullback(const double* GX, const double* VX, const double* logdet, const int* idx, const int* n_dev,
                int n_rows, double* G, int64_t ldg, double* V);
  // In place on n_rows sampler rows (stride ld, z in columns 7..7+D): z -> x = T(z),
  // column 0 (lp of z) -> lp(z) - log|det dT/dz|.
  void map_rows(double* rows, int64_t ld, int n_rows);
  // Throws if a grid barrier of the persistent sampling kernels timed out
  // since the last check (their results are then not to be used); synchronizes
  // the stream when those kernels have run.
  void check_barriers();
  // z = T^{-1}(x) for n_rows rows of X (stride D), n_rows <= max_rows.
  void inverse(const double* X, int n_rows, double* Z, double* logdet);

  // Adam training on n_draws device rows of unconstrained draws (stride D):
  // maximum likelihood of the draws under the flow.  Returns the mean loss
  // over the last `report` steps' batches.
  double train(const double* draws, int64_t n_draws, int steps, int batch, float lr, uint64_t seed, int report);
  // Loss and parameter gradient of one batch of rows (tests).
  double loss_grad(const double* X, int n_rows, float* grad_host);

 private:
  struct Layer {
    int Da, Db;
    int* ia; int* ib;          // device index lists
    int64_t off[6];            // W1 b1 W2 b2 W3 b3 offsets in the flat parameter vector
  };
  void mlp(int l, const float* U, int ldu, const int* ia, const int* n_dev, int rows);
  void gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C,
            int64_t ldc, int M, int N, int K, const int* Mdev, int epi, const float* bias, const float* aux,
            int64_t ldaux, const int* aidx_k = nullptr, const int* aidx_m = nullptr, const int* cidx_n = nullptr);
  bool fused_ok(int n_rows) const;  // the fused per-layer kernels fit (H, partial workspace)
  bool persistent_ok(int n_rows) const;  // the persistent forward / pullback kernels apply
  void persistent_args(FpArgs& a, const int* idx, const int* n_dev, int n_rows) const;
  void inverse_pass(int rows);     // U_[L] -> U_[0], keeping the training tape
  void row_losses(int rows);       // rowbuf_ = per-row loss after inverse_pass
  void backward_params(int rows);  // parameter gradient of the mean loss into grad_

  FlowSpec spec_;
  int R_, Dmax_a_ = 0, Dmax_b_ = 0;
  int64_t n_params_ = 0;
  hipStream_t st_ = nullptr, own_st_ = nullptr;
  std::vector<Layer> layers_;
  float* params_ = nullptr; float* grad_ = nullptr; float* adam_m_ = nullptr; float* adam_v_ = nullptr;
  double* loc_ = nullptr; double* scale_ = nullptr; double log_scale_sum_ = 0;
  float* U_ = nullptr;     // (L+1) x R x D layer inputs/outputs
  float* H1_ = nullptr; float* H2_ = nullptr;  // L x R x H
  float* O_ = nullptr;     // R x 2 Dmax_b
  float* S_ = nullptr;     // L x R x Dmax_b
  float* GO_ = nullptr;    // R x 2 Dmax_b
  float* GP2_ = nullptr; float* GP1_ = nullptr;  // R x H
  float* GC_ = nullptr;    // R x D
  float* part_ = nullptr; int64_t part_n_ = 0;  // split-K partials
  double* rowbuf_ = nullptr; double* rowbuf2_ = nullptr;  // R
  double* xrows_ = nullptr;   // R x D (map_rows)
  int64_t* pick_ = nullptr;   // R minibatch rows
  int* db_dev_ = nullptr;     // L
  double* ld_part_ = nullptr; // L x ceil(Dmax_b / 64) x R: per-tile row sums of s (fused forward)
  unsigned* bar_ = nullptr;   // grid-barrier counters of the persistent kernels (arrivals, departures)
  int* bar_err_ = nullptr;    // set on the device when a grid barrier times out
  int n_cu_ = 0;              // blocks of a persistent launch (one per CU)
  uint64_t spin_ticks_ = 0;   // wall-clock ticks a grid barrier may wait (2 s)
  bool ran_persistent_ = false;
  int64_t adam_t_ = 0;
};

// Model gradient at packed rows X (one wave per row): VX[r], GX row r for the
// chain idx[r] (its dataset); rows r < *n_dev.
void launch_model_grad_rows(int kind, const int* idx, const int* n_dev, int n_rows, int D, const double* X,
                            double* GX, double* VX, const ModelDataDev& md, hipStream_t st);

}  // namespace nutsgpu
