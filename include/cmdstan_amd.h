/* cmdstan_amd — batched NUTS on AMD Instinct MI355X (gfx950) behind a C ABI.
 *
 * The entry points replace, for `method=sample algorithm=hmc engine=nuts`, the
 * calls CmdStan's command layer makes into the (absent) stan services library:
 *
 *   nuts_data_*          stan::io::var_context built from `data file=`
 *                        (src/cmdstan/command_helper.hpp:116-133 get_var_context)
 *   nuts_model_create    new_model(var_context&, seed, ostream*)
 *                        (src/cmdstan/command.hpp:71-72, 170-171)
 *   nuts_sample          stan::services::sample::hmc_nuts_{unit,diag,dense}_e[_adapt]
 *                        (src/cmdstan/command.hpp:603-697; diag_e + adapt at :681-688)
 *   nuts_log_prob_grad   services_log_prob_grad -> stan::model::log_prob_grad<true, jacobian>
 *                        (src/cmdstan/command_helper.hpp:569-600)
 *   nuts_last_error      the message CmdStan prints before returning
 *                        return_codes::NOT_OK (src/cmdstan/main.cpp:4-15)
 *
 * Conventions: plain pointers to caller-owned host memory; device memory is
 * owned by the library.  Every function returns NUTS_OK (0) or NUTS_NOT_OK (1)
 * and leaves a message for nuts_last_error().  One handle per host thread.
 */
#ifndef CMDSTAN_AMD_H
#define CMDSTAN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NUTS_OK 0
#define NUTS_NOT_OK 1

#define NUTS_METRIC_UNIT_E 0
#define NUTS_METRIC_DIAG_E 1
#define NUTS_METRIC_DENSE_E 2

typedef struct nuts_data nuts_data;
typedef struct nuts_model nuts_model;

/* ---------------------------------------------------------------- data */
/* A named-array context: stan::io::var_context (command_helper.hpp:116-133). */
nuts_data* nuts_data_create(void);
void nuts_data_destroy(nuts_data* d);
/* Adds a real (is_int=0) or integer (is_int=1) array, row-major, ndim 0 = scalar. */
int nuts_data_add(nuts_data* d, const char* name, int ndim, const int64_t* dims, const double* values, int is_int);
/* Parses CmdStan JSON data text (json_data_handler semantics: numbers, nested
 * arrays, "inf"/"-inf"/"NaN" strings). */
int nuts_data_parse_json(nuts_data* d, const char* json_text);
/* Parses R dump text, the other format of get_var_context (command_helper.hpp:
 * 116-133, stan::io::dump): name <- value with numbers, c(...), a:b and
 * structure(..., .Dim = c(...)); arrays are stored row-major like JSON ones. */
int nuts_data_parse_rdump(nuts_data* d, const char* rdump_text);
int nuts_data_get(const nuts_data* d, const char* name, int* ndim, int64_t* dims /* >= 8 */, const double** values);

/* --------------------------------------------------------------- model */
/* Creates a model of a registered family for one or more datasets (a batch of
 * models of uneven size sampled in one launch).  Data are copied to HBM here.
 * Families: bernoulli, logistic_regression, eight_schools_cp, hier_normal_nc,
 * hier_logistic_cp, gauss. */
nuts_model* nuts_model_create(const char* family, const nuts_data* const* datasets, int num_datasets);
void nuts_model_destroy(nuts_model* m);
int nuts_model_num_datasets(const nuts_model* m);
/* Unconstrained dimension and number of constrained outputs of a dataset. */
int nuts_model_dims(const nuts_model* m, int dataset, int* num_unconstrained, int* num_outputs);
/* Name of output column i (constrained parameters in Stan CSV order, e.g. "theta.1"). */
int nuts_model_param_name(const nuts_model* m, int dataset, int i, char* buf, int buflen);
/* Name of unconstrained coordinate i (log_prob output "g_<name>"). */
int nuts_model_unconstrained_name(const nuts_model* m, int dataset, int i, char* buf, int buflen);

/* ------------------------------------------------------- model plugins */
/* A model outside the registered families, supplied by the caller in the role
 * of stanc's generated model_base (created by new_model, command.hpp:71-72,
 * 170-171; its log_prob, write_array, transform_inits and *_param_names are
 * what command_helper.hpp:253-404 and the services call): host callbacks for
 * the names and transforms, and a log-density gradient the caller builds for
 * gfx950 and enqueues on the sampler's stream.  The sampler drives plugin
 * models with its lockstep kernels (the gradient is a separate launch per
 * step of the state machine, batched over every chain that needs one). */
typedef struct nuts_model_plugin {
  int dim;                                /* unconstrained dimension D >= 1 */
  int num_outputs;                        /* constrained outputs of a draw */
  const char* const* param_names;         /* [num_outputs], Stan CSV names ("theta", "beta.1") */
  const char* const* unconstrained_names; /* [dim] (log_prob columns g_<name>) */
  void* user;                             /* passed back to every callback */
  /* V = -log p and G = -grad log p, log p in Stan's propto convention plus the
   * log-Jacobian of the constraining transforms when jacobian = 1, for rows
   * r < n, n = *d_n when d_n is not NULL (a device count, <= n_rows), else
   * n_rows.  Point r is row idx[r] of Q (row r when idx is NULL), stride ldq;
   * its results go to V[idx[r]] and row idx[r] of G (stride ldq).  Q, G, V,
   * idx and d_n are device pointers.  Enqueue the work on `stream` (a
   * hipStream_t) without synchronising or allocating; return 0, or nonzero to
   * report an error.  Required. */
  int (*log_prob_grad)(void* user, const double* Q, int64_t ldq, const int* idx, const int* d_n, int n_rows,
                       int jacobian, double* G, double* V, void* stream);
  /* write_array: constrained outputs out [n][num_outputs] of the unconstrained
   * host rows q [n][dim].  NULL: the outputs are the unconstrained values. */
  int (*write_array)(void* user, const double* q, int n, double* out);
  /* transform_inits: the named constrained values in `inits` to the
   * unconstrained out[dim], NaN for parameters absent from `inits`.  NULL:
   * init files cannot be used with this model. */
  int (*transform_inits)(void* user, const nuts_data* inits, double* out);
} nuts_model_plugin;

/* Creates a model from a plugin (one dataset; the plugin holds its own data).
 * The names are copied; the callbacks and `user` must outlive the model. */
nuts_model* nuts_model_create_plugin(const nuts_model_plugin* plugin);

/* ------------------------------------------------------------ log_prob */
/* lp and its gradient at n_points unconstrained points of one dataset, computed
 * on the GPU in FP64 (propto=true; jacobian=0/1).  q, grad: [n_points][D]. */
int nuts_log_prob_grad(nuts_model* m, int dataset, const double* q, int n_points, int jacobian, double* lp,
                       double* grad);

/* The same at the reduced precisions of the batched GLM path (precision 1:
 * FP32 contraction, X in FP32; 2: X stored in BF16), for error statements
 * against FP64.  No counterpart in the reference. */
int nuts_log_prob_grad_precision(nuts_model* m, int dataset, const double* q, int n_points, int jacobian,
                                 int precision, double* lp, double* grad);

/* nuts_log_prob_grad over caller-owned DEVICE buffers, allocating nothing:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  d_q [n_points][D], d_lp [n_points], d_grad [n_points][D]
 * are device pointers; the work is enqueued on `stream` (a hipStream_t, NULL
 * for the null stream) and the call returns without waiting for it.  Scratch
 * is a fixed 48 MiB device array per GPU, held until the process ends;
 * n_points is at most 4096 (fewer when a model's partial sums need more
 * scratch: the call then fails with a message).  Calls that overlap on one
 * GPU share that scratch and race: serialise them on one stream. */
int nuts_log_prob_grad_noalloc(nuts_model* m, int dataset, const double* d_q, int n_points, int jacobian,
                               double* d_lp, double* d_grad, void* stream);

/* -------------------------------------------------------------- sample */
/* Argument tree of `sample ... algorithm=hmc engine=nuts` with CmdStan defaults
 * (arg_sample.hpp:18-29, arg_adapt.hpp:19-37, arg_hmc.hpp:19-27,
 * arg_nuts.hpp:15-16, arg_metric.hpp:17-22, arg_id.hpp, arg_seed.hpp, arg_init.hpp). */
typedef struct nuts_config {
  int num_samples;        /* 1000 */
  int num_warmup;         /* 1000 */
  int save_warmup;        /* 0 */
  int thin;               /* 1 */
  int adapt_engaged;      /* 1 */
  double gamma;           /* 0.05 */
  double delta;           /* 0.8 */
  double kappa;           /* 0.75 */
  double t0;              /* 10 */
  int init_buffer;        /* 75 */
  int term_buffer;        /* 50 */
  int window;             /* 25 */
  int max_depth;          /* 10 */
  int metric;             /* NUTS_METRIC_DIAG_E */
  double stepsize;        /* 1 */
  double stepsize_jitter; /* 0 */
  int num_chains;         /* chains per dataset, 1 */
  int id;                 /* 1: chain k uses RNG stream id + k */
  uint64_t seed;          /* random seed */
  double init_radius;     /* 2 */
  int fused;              /* -1 auto, 0 lockstep (batched gradient kernels), 1 fused per-chain */
  int precision;          /* batched GLM contraction: 0 FP64 (default), 1 FP32 with X in FP32,
                             2 FP32 with X stored in BF16; on the fused per-chain kernel (other
                             models, unit_e/diag_e) 1 keeps q, p, g, checkpoints and the metric in
                             FP32 (the FP32 leapfrog); sums, H, U-turn and tree weights stay FP64 */
  int max_saved_params;   /* -1 (default): draws hold the constrained parameters; k >= 0: draws hold
                             the first k unconstrained coordinates only (bounded output for very
                             large batches) */
  /* Learned reparameterisation (no counterpart in CmdStan): flow = 1 runs plain
   * adaptive warmup plus flow_pilot draws, trains an affine-coupling flow on the
   * second half of warmup and the pilot draws, runs flow_warmup adaptive
   * iterations and flow_pilot draws of NUTS on the pulled-back density, and then
   * samples num_samples iterations with the flow only if its pilot raised min
   * bulk-ESS per gradient evaluation over the plain pilot and either raised
   * min bulk-ESS per second too or mixed where plain NUTS did not (pilot max
   * R-hat above 1.05 for plain NUTS and lower for the flow); plain NUTS is the
   * fallback.  One dataset; at least 2 dimensions; lockstep kernels. */
  int flow;               /* 0 (default) off, 1 on with the plain-NUTS fallback, 2 on without it */
  int flow_layers;        /* 8 coupling layers */
  int flow_hidden;        /* 128: width of the two tanh hidden layers */
  int flow_train_steps;   /* 1000 Adam steps */
  int flow_batch;         /* 256 draws per step */
  double flow_lr;         /* 1e-3 */
  int flow_warmup;        /* 300 adaptive iterations in the flow's coordinates */
  int flow_pilot;         /* 100 pilot draws per chain for each of the two samplers */
  int refresh;            /* 0 (default): silent; k > 0: progress line on stderr each time the
                             slowest chain passes a multiple of k iterations (CmdStan's refresh=) */
  /* callbacks::interrupt (command.hpp:186): when not NULL, called on the host
   * between check intervals of the sampler (every 16 lockstep steps, or every
   * fused launch); a nonzero return stops the run with NUTS_NOT_OK and the
   * message "sampling interrupted". */
  int (*interrupt)(void* user);
  void* interrupt_user;
} nuts_config;

void nuts_config_default(nuts_config* cfg);
/* Validates like CmdStan's command layer: e.g. adapt engaged with num_warmup=0
 * is an error (command.hpp:567-574), delta in [0.1, 1), jitter in [0, 1]. */
int nuts_config_validate(const nuts_config* cfg);
/* Rows of each chain's draws block: saved warmup + sampling iterations. */
int nuts_output_rows(const nuts_config* cfg);

/* Caller-allocated outputs (any pointer may be NULL).  Chains are ordered by
 * dataset, num_chains per dataset; chain k of dataset m is chain m*num_chains+k.
 *   draws:     chains back to back, each rows x (7 + num_outputs of its dataset):
 *              lp__, accept_stat__, stepsize__, treedepth__, n_leapfrog__,
 *              divergent__, energy__, then the constrained parameters
 *   stepsize:  [chains] adapted (or given) step size
 *   inv_metric:[chains] blocks of D (diag_e/unit_e) or D*D (dense_e)
 *   n_leapfrog:[chains] total leapfrog steps; grad_evals: [chains] all gradient
 *              evaluations (initialisation and step-size search included)
 *   status:    [chains] 0 ok, else a failure code (1 init, 2 improper posterior,
 *              3 zero step size, 4 metric overflow, 5 dense metric not positive definite) */
typedef struct nuts_output {
  double* draws;
  double* stepsize;
  double* inv_metric;
  int64_t* n_leapfrog;
  int64_t* grad_evals;
  int* status;
  double warmup_seconds;   /* filled by nuts_sample: GPU wall time of the warmup phase */
  double sampling_seconds; /* ... and of the sampling phase */
  int64_t leaf_steps;      /* lockstep gradient launches (0 in fused mode) */
  double grad_seconds_per_launch; /* lockstep: mean duration of the batched gradient's
                                     contraction kernel (the whole gradient off the GLM
                                     path) over the timed launches, from HIP events on
                                     the sampler's stream (0 in fused mode); with graph
                                     replay one bundle of 16 launches in 8 is launched
                                     eagerly and timed, spread over warmup and sampling */
  int64_t grad_launches_timed;
  double grad_chains_per_launch;  /* lockstep: mean number of active chains over the same
                                     timed launches (0 in fused mode) */
  /* flow = 1 only */
  int flow_used;                  /* 1: the draws come from NUTS on the pulled-back density */
  double flow_train_seconds;      /* GPU training time */
  double flow_loss;               /* mean negative log-likelihood of the last training batches */
  double pilot_ess_per_grad_plain;
  double pilot_ess_per_grad_flow;
  double pilot_ess_per_sec_plain; /* the flow is used only if it raises both, or */
  double pilot_ess_per_sec_flow;
  double pilot_rhat_plain;        /* if plain NUTS does not mix (max R-hat > 1.05) */
  double pilot_rhat_flow;         /* and the flow mixes better */
  /* batched GLM gradient chosen by the autotuner (lockstep logistic regression):
     path 0 single-pass kernel with tuned_blocks_per_cu blocks per CU, 1 two-pass
     kernels; -1 when not applicable */
  int tuned_path;
  int tuned_blocks_per_cu;
  /* diagnostic_file rows (mcmc_writer::write_diagnostic_params): when not NULL,
   * chains back to back, each rows x (7 + 3 D): the seven sampler columns, then
   * the unconstrained draw q, its momentum p and the potential gradient
   * g = -grad log p (columns named <name>, p_<name>, g_<name>).  Not with flow. */
  double* diagnostics;
  /* per-transition energy error (no counterpart in CmdStan): when not NULL,
   * [chains][rows], for each saved row the largest |H - H0| over the leaves of
   * the trajectory that produced it (H0: the Hamiltonian at the transition's
   * start; +inf for a divergent leaf).  With the flow: the final sampling phase. */
  double* energy_error;
  /* services::util::initialize: when not NULL, [chains][2], the initial values
   * rejected because the log density [0] or its gradient [1] was not finite
   * (the "Rejecting initial value" messages of command_test.cpp:151-250). */
  int* init_rejections;
  /* lockstep state machine (no counterpart in CmdStan): threads per chain the
   * run settled on -- 64 (one wave) or 256 (one block), timed on the first
   * four check intervals (wave, block, block, wave) when neither the caller
   * nor CMDSTAN_AMD_BLOCK_PER_CHAIN fixes it -- and the time of one
   * state-machine launch in each mode, the mean of its two intervals' medians
   * (0 when not timed; fused mode reports 0 threads) */
  int tuned_threads_per_chain;
  double tuned_wave_launch_seconds;
  double tuned_block_launch_seconds;
} nuts_output;

/* model_base::transform_inits (the `init=` file path of command_helper.hpp:144-245):
 * the named constrained values in `inits` to the unconstrained vector out[D];
 * parameters absent from `inits` are NaN, which nuts_sample draws uniformly in
 * (-R, R) like services::util::initialize. */
int nuts_model_transform_inits(nuts_model* m, int dataset, const nuts_data* inits, double* out);

/* Runs warmup then sampling for every chain of every dataset in one batch.
 * init: unconstrained inits [chains][D] or NULL (uniform(-R, R), R = init_radius);
 *       NaN entries are drawn the same way;
 * inv_metric: per-chain inverse metric [chains][D or D*D] or NULL (identity),
 * the `metric_file=` input. */
int nuts_sample(nuts_model* m, const nuts_config* cfg, const double* init, const double* inv_metric,
                nuts_output* out);

/* One NUTS transition per chain from q0 with injected uniforms and normals
 * (consumed in the order of the upstream recursion).  No counterpart in the
 * reference: this is the parity surface for the sampler arithmetic.
 * q0: [chains][D]; u: [chains][n_u]; z: [chains][n_z]; inv_metric [chains][D];
 * out_q: [chains][D]; info: [chains][9] = depth, n_leapfrog, divergent,
 * accept_stat, energy, lp, uniforms used, normals used, step size. */
int nuts_transition_replay(nuts_model* m, int dataset, int num_chains, int metric, double stepsize, double jitter,
                           int max_depth, const double* inv_metric, const double* q0, const double* u, int64_t n_u,
                           const double* z, int64_t n_z, double* out_q, double* info);
/* The same with the sampler's vectors at a reduced precision (precision 1: the
 * FP32 leapfrog of the per-chain kernel; H, tree weights and U-turn sums stay
 * FP64).  No counterpart in the reference. */
int nuts_transition_replay_precision(nuts_model* m, int dataset, int num_chains, int metric, double stepsize,
                                     double jitter, int max_depth, const double* inv_metric, const double* q0,
                                     const double* u, int64_t n_u, const double* z, int64_t n_z, int precision,
                                     double* out_q, double* info);

/* The same with the driver of the state machine chosen: driver 0 the fused
 * per-chain kernel (gradient inside the state machine), 1 the lockstep kernels
 * with one wavefront per chain and the model's batched gradient kernel (the
 * FP64 matrix-core GEMM for logistic regression, the group kernel for the
 * hierarchical logistic model), 2 the lockstep kernels with one 256-thread
 * block per chain, 3 the fused kernel with one 256-thread block per chain
 * (models whose gradient spreads over a block; unit/diag).  Sampling uses
 * driver 3 by itself for the last chains of a phase.  No counterpart in the
 * reference. */
#define NUTS_DRIVER_FUSED 0
#define NUTS_DRIVER_LOCKSTEP_WAVE 1
#define NUTS_DRIVER_LOCKSTEP_BLOCK 2
#define NUTS_DRIVER_FUSED_BLOCK 3
int nuts_transition_replay_driver(nuts_model* m, int dataset, int num_chains, int metric, double stepsize,
                                  double jitter, int max_depth, const double* inv_metric, const double* q0,
                                  const double* u, int64_t n_u, const double* z, int64_t n_z, int precision,
                                  int driver, double* out_q, double* info);

/* ------------------------------------------------------------ summary */
/* stansummary statistics (src/cmdstan/stansummary_helper.hpp:397-428 get_stats)
 * for n_params parameters; draws[p][chain][draw].  out[p][10] = Mean, MCSE,
 * StdDev, MAD, 5%, 50%, 95%, ESS_bulk, ESS_tail, R_hat.  Host code. */
int nuts_summary(const double* draws, int n_draws, int n_chains, int n_params, int n_threads, double* out);

/* ------------------------------------------------------ learned flow */
/* Affine-coupling normalizing flow used by nuts_config.flow = 1, exposed for
 * its parity tests.  No counterpart in the reference.  x = loc + scale *
 * f_{L-1}(...f_0(z)); layer l keeps dims a_l and maps b_l by
 * v_b = u_b exp(s) + t, (s_raw, t) = MLP(u_a) with two tanh layers of width
 * `hidden`, s = 3 tanh(s_raw / 3).  Masks: n_glob leading global dims, then
 * groups of group_size dims (group effects conditioned on the globals).
 * Parameters are FP32; all row buffers are host [n][D] FP64 with n <= max_rows. */
typedef struct nuts_flow nuts_flow;
nuts_flow* nuts_flow_create(int D, int n_layers, int hidden, int n_glob, int group_size, int max_rows,
                            uint64_t seed);
void nuts_flow_destroy(nuts_flow* f);
int64_t nuts_flow_num_params(const nuts_flow* f);
/* layout per layer: W1 [Da][hidden], b1, W2 [hidden][hidden], b2, W3 [hidden][2 Db], b3 */
int nuts_flow_get_params(const nuts_flow* f, float* out);
int nuts_flow_set_params(nuts_flow* f, const float* in);
int nuts_flow_set_standardization(nuts_flow* f, const double* loc, const double* scale);
/* x = T(z), logdet = log|det dT/dz| */
int nuts_flow_forward(nuts_flow* f, const double* z, int n, double* x, double* logdet);
/* z = T^{-1}(x), logdet = log|det dT^{-1}/dx| */
int nuts_flow_inverse(nuts_flow* f, const double* x, int n, double* z, double* logdet);
/* gz = d/dz [V(T(z)) - log|det dT/dz|] given gx = grad V at x = T(z) */
int nuts_flow_pullback_grad(nuts_flow* f, const double* z, const double* gx, int n, double* gz);
/* mean over rows of 0.5 |T^{-1}(x)|^2 - log|det dT^{-1}/dx| and its parameter gradient */
int nuts_flow_loss_grad(nuts_flow* f, const double* x, int n, double* loss, float* grad);
/* Adam on minibatches of `batch` rows drawn from the n rows of x */
int nuts_flow_train(nuts_flow* f, const double* x, int64_t n, int steps, int batch, double lr, uint64_t seed,
                    double* final_loss);

const char* nuts_last_error(void);
const char* nuts_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CMDSTAN_AMD_H */
