/*
 * yannq_amd C ABI: the variational Monte Carlo step of yannq (Metropolis
 * sampling of an RBM, local energies, the log-derivative matrix O and the
 * stochastic-reconfiguration solve) on one AMD Instinct MI355X (gfx950).
 *
 * The reference exposes this path as compile-time C++ concepts, not an FFI.
 * Each entry point below replaces one member of those concepts; the citation
 * names the reference member (paths under the yannq source tree).  The
 * header-only C++ wrappers in include/yannq_amd/ rebuild the concepts
 * (Machine, Sampler, SRMat, Runner) on top of these calls.
 *
 * Conventions
 *   - Every function returns a yq_status; 0 is success.  yq_last_error()
 *     gives the message of the last failure on that context.  Nothing throws.
 *   - Host arrays are caller-owned and copied in or out.  Device buffers are
 *     owned by the context.  One host thread per context; all work is
 *     ordered on the context's HIP stream and calls return after it finished.
 *   - Real parameters are doubles in the layout [W (m x n, column-major) |
 *     a (n) | b (m)] of Machines/RBM.hpp:271-281 (W block only without bias).
 *   - A complex machine (YQ_DTYPE_C128) uses the same layout with complex
 *     entries stored as interleaved (re, im) doubles, as std::complex<double>
 *     arrays are; so are its theta, local energies, O, S, F, column mean and
 *     the vectors of yq_sr_apply / yq_sr_solve (2 doubles per entry).
 *   - Spins are int8 (+1 / -1), one row of n per configuration.
 */
#ifndef YANNQ_AMD_H
#define YANNQ_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YQ_API __attribute__((visibility("default")))

typedef struct yq_ctx yq_ctx;

typedef enum {
  YQ_OK = 0,
  YQ_ERR_ARG = 1,          /* invalid argument or size */
  YQ_ERR_STATE = 2,        /* call out of order (e.g. no samples yet) */
  YQ_ERR_HIP = 3,          /* HIP / rocBLAS / rocSOLVER failure */
  YQ_ERR_NAN = 4,          /* NaN or Inf in parameters or results */
  YQ_ERR_NOT_SPD = 5,      /* Cholesky of S + shift I failed */
  YQ_ERR_NOT_CONVERGED = 6,/* iterative solve hit max_iter */
  YQ_ERR_OOM = 7,          /* device allocation failed */
  YQ_ERR_UNSUPPORTED = 8   /* combination not implemented */
} yq_status;

typedef enum { YQ_TFI = 0, YQ_XXZ = 1, YQ_J1J2 = 2 } yq_ham_kind;
/* machine scalar type: RBM<double> or RBM<std::complex<double>> */
typedef enum { YQ_DTYPE_F64 = 1, YQ_DTYPE_C128 = 2 } yq_dtype;
typedef enum { YQ_LOCAL = 0, YQ_EXCHANGE = 1 } yq_sweeper_kind;
/* YQ_FP64: every operation in double.  YQ_MIXED: sampling and local energies
 * in float with theta refreshed from double once per sweep; the CG products
 * as split-bf16 MFMA with float accumulation; residual, refinement and update
 * in double. */
typedef enum { YQ_FP64 = 0, YQ_MIXED = 1 } yq_precision;
/* YQ_SOLVE_EXPLICIT: S = O_c^T O_c / Ns and a Cholesky solve; with YQ_MIXED
 *   (real machines, unweighted samples) S comes from split-bf16 MFMA with
 *   float accumulation and the FP32 Cholesky factor drives an FP64
 *   iterative refinement against the FP64 operator.
 * YQ_SOLVE_CG: matrix-free conjugate gradients on O_c (materialised O in
 *   FP64, the reference's apply), no S.
 * YQ_SOLVE_CG_FACTORED: matrix-free CG on the factored O (sigma and
 *   tanh(theta) only, O never stored); with YQ_MIXED the inner products run
 *   on split-bf16 MFMA and an FP64 iterative refinement drives the residual. */
typedef enum {
  YQ_SOLVE_EXPLICIT = 0,
  YQ_SOLVE_CG = 1,
  YQ_SOLVE_CG_FACTORED = 2
} yq_solver;
typedef enum {
  YQ_SR_F = 0,      /* energy gradient F (dim) */
  YQ_SR_OMEAN = 1,  /* column mean of O (dim) */
  YQ_SR_O = 2,      /* centred O (Ns x dim, row-major); explicit/CG only */
  YQ_SR_S = 3,      /* S (dim x dim, column-major, full); explicit only */
  YQ_SR_ELOC = 4    /* local energies (Ns) */
} yq_sr_item;

typedef struct {
  int kind;       /* yq_ham_kind */
  int n;          /* sites, periodic chain */
  double J;       /* TFI: J; XXZ: J; J1J2: J1 */
  double h;       /* TFI: coefficient of the transverse term */
  double delta;   /* XXZ: anisotropy */
  double J2;      /* J1J2: next-nearest coupling */
  int sign_rule;  /* XXZ / J1J2: 1 applies the Marshall sign rule */
} yq_ham_desc;

/* ---- context ---------------------------------------------------------- */
/* No reference counterpart (the reference has no device); owns the stream. */
YQ_API int yq_ctx_create(int device, yq_ctx** out);
YQ_API int yq_ctx_destroy(yq_ctx* ctx);
YQ_API const char* yq_last_error(const yq_ctx* ctx);

/* ---- Machine: Machines/RBM.hpp ---------------------------------------- */
/* RBM(n, m, useBias), all parameters zero: RBM.hpp:50-56.  dtype is a
 * yq_dtype; 2 <= n <= 4096 and 1 <= m <= 8192.  Complex machines run every
 * solver: with YQ_MIXED sampling and E_loc run in float, the explicit S is
 * formed in split-bf16 and solved by FP64-refined Cholesky, and the dense and
 * factored CG stay FP64 (the factored complex CG composes four real factored
 * products per complex one and has no Jacobi option).  Parallel tempering and
 * the many-machine calls need m <= 512. */
YQ_API int yq_rbm_create(yq_ctx* ctx, int dtype, int n, int m, int use_bias);
/* getN / getM / getDim: RBM.hpp:115-130 */
YQ_API int yq_rbm_dims(const yq_ctx* ctx, int* n, int* m, int* dim);
/* setParams / getParams: RBM.hpp:271-291 */
YQ_API int yq_rbm_set_params(yq_ctx* ctx, const double* params);
YQ_API int yq_rbm_get_params(yq_ctx* ctx, double* params);
/* updateParams (params += delta): RBM.hpp:261-269 */
YQ_API int yq_rbm_update_params(yq_ctx* ctx, const double* delta);
/* initializeRandom(re, sigma): RBM.hpp:301-322 (complex: :327-349, re and
 * im of each entry from consecutive draws), normals drawn in the order a, b,
 * W from a Philox stream keyed by seed (values differ from the reference's
 * std::normal_distribution by design). */
YQ_API int yq_rbm_init_random(yq_ctx* ctx, uint64_t seed, double sigma);
/* hasNaN: RBM.hpp:293-296 */
YQ_API int yq_rbm_has_nan(yq_ctx* ctx, int* out);

/* ---- Checkpoints: Serializers/SerializeRBM.hpp + SerializeEigen.hpp ----- */
/* The reference's weight files: cereal BinaryArchive of RBM<T>
 * (SerializeRBM.hpp:21-79, Eigen matrices as SerializeEigen.hpp:10-35),
 * written by RunRBM.hpp:155-165 as w%04d.dat through a std::unique_ptr
 * (pointer_wrapper = 1: one leading "valid" byte) and read back by
 * AbstractRunner.hpp:83-94; pointer_wrapper = 0 is the bare object form
 * (TestSerializeRBM.cpp's second case).  The file does not name its scalar
 * type, so the caller gives it: */
typedef enum {
  YQ_CKPT_F32 = 1,  /* RBM<float> */
  YQ_CKPT_F64 = 2,  /* RBM<double> */
  YQ_CKPT_C64 = 3,  /* RBM<std::complex<float>> */
  YQ_CKPT_C128 = 4  /* RBM<std::complex<double>> */
} yq_ckpt_elem;
/* Bytes of the encoding (0 for invalid arguments).  Host only. */
YQ_API size_t yq_ckpt_rbm_bytes(int elem, int n, int m, int use_bias, int pointer_wrapper);
/* Encode parameters given in this ABI's layout (doubles; complex as (re, im)
 * pairs; float files round to nearest).  Host only. */
YQ_API int yq_ckpt_rbm_encode(int elem, int n, int m, int use_bias, const double* params,
                              int pointer_wrapper, unsigned char* out, size_t size);
/* Decode: n, m and use_bias always; params too when not null (params_len
 * doubles available).  YQ_ERR_ARG when the bytes are not exactly such an
 * encoding (wrong type, form, sizes or length).  Host only. */
YQ_API int yq_ckpt_rbm_decode(int elem, const unsigned char* buf, size_t size,
                              int pointer_wrapper, int* n, int* m, int* use_bias,
                              double* params, size_t params_len);
/* Write the context's machine to a file (elem must match real / complex). */
YQ_API int yq_rbm_save(yq_ctx* ctx, const char* path, int elem, int pointer_wrapper);
/* Read a file and make it the context's machine (yq_rbm_create with the
 * file's n, m, useBias, then its parameters). */
YQ_API int yq_rbm_load(yq_ctx* ctx, const char* path, int elem, int pointer_wrapper);

/* ---- Hamiltonian: Hamiltonians/{TFIsing,XXZ,J1J2}.hpp ------------------ */
/* A POD descriptor replaces the E_loc functor operator()(const State&):
 * TFIsing.hpp:28-39, XXZ.hpp:37-49, J1J2.hpp:35-55. */
YQ_API int yq_ham_set(yq_ctx* ctx, const yq_ham_desc* desc);

/* ---- Sampler: Samplers/SamplerMT.hpp + Local/SwapSweeper --------------- */
/* SamplerMT(qs, nTmps, nChainsPer, sweeper): SamplerMT.hpp:44-60, with the
 * sweeper's nSweep (LocalSweeper.hpp:19, SwapSweeper.hpp:15) as
 * sweep_factor.  seed keys the Philox streams of every chain.  n_temps 1, or
 * even and <= 16 (real and complex machines): parallel tempering with nTmps x nChainsPer
 * chain states, state g * n_temps + t at beta_t = (n_temps - t) / n_temps,
 * replica swaps after every sweep (mixChains, SamplerMT.hpp:120-153) and one
 * sample per chain group per sweep from the beta = 1 slot; state and counter
 * arrays of the calls below then cover all n_temps x n_chains_per states. */
YQ_API int yq_sampler_config(yq_ctx* ctx, int sweeper, int n_temps,
                             int n_chains_per, int sweep_factor, int precision,
                             uint64_t seed);
/* Threads whose partial sums form one chain's log-ratio for a machine with
 * m hidden units: 64 up to m = 512, else 64 times the smallest power of two
 * that leaves at most 8 hidden units per thread.  Hidden unit j is summed by
 * thread j % lanes in increasing j, then the threads are combined by the tree
 * with strides lanes/2, ..., 1 (what a CPU replay of the accept decisions
 * needs).  No counterpart in the reference, whose reduction is sequential. */
YQ_API int yq_chain_lanes(int m);
/* randomizeSigma() / randomizeSigma(nup): SamplerMT.hpp:81-98; nup < 0
 * draws unconstrained spins. */
YQ_API int yq_sampler_randomize(yq_ctx* ctx, int nup);
/* randomize(randomizer) with caller-chosen configurations (chains x n):
 * SamplerMT.hpp:103-110. */
YQ_API int yq_sampler_set_sigma(yq_ctx* ctx, const int8_t* sigma);
/* Chain state: spins (chains x n), per-chain proposal counters and accepted
 * proposal counts. */
YQ_API int yq_sampler_get_state(yq_ctx* ctx, int8_t* sigma, uint64_t* ctr,
                                uint64_t* acc);
/* sample(nSweeps, nDiscard): SamplerMT.hpp:180-204.  Samples stay on the
 * device; *n_samples = n_chains_per * n_sweeps, ordered sweep-major. */
YQ_API int yq_sample(yq_ctx* ctx, int n_sweeps, int n_discard, int* n_samples);
/* State::logRatio(k), logRatio(k, l) and logRatio(vector) (States/
 * RBMState.hpp:48-63, 78-93, 108-126) for ns configurations at once: out[s]
 * = log psi(sigma_s with the sites listed in row s of sites flipped) -
 * log psi(sigma_s), where row s holds up to k site indices (list order;
 * entries < 0 are skipped, so rows may list fewer sites).  sigma is ns x n
 * (+1 / -1), sites ns x k; out holds ns doubles, or ns (re, im) pairs for a
 * complex machine (the imaginary part is the sum of principal logs, defined
 * modulo 2 pi as the reference's).  FP64; independent of any batch. */
YQ_API int yq_rbm_log_ratio(yq_ctx* ctx, int ns, const int8_t* sigma, int k,
                            const int* sites, double* out);
/* Copy the current batch out: spins (Ns x n) and theta (Ns x m, doubles). */
YQ_API int yq_samples_get(yq_ctx* ctx, int8_t* sigma, double* theta);
/* Replace the batch by caller configurations (Ns x n); theta = W sigma + b is
 * computed on the device (makeData, RBM.hpp:362-365). precision selects the
 * arithmetic of the following yq_eloc / yq_sr_build. */
YQ_API int yq_samples_set(yq_ctx* ctx, int ns, const int8_t* sigma,
                          int precision);

/* ---- Exact sums: GroundState/SRMatExact.hpp + Samplers/ExactSampler.hpp -- */
/* The batch becomes every configuration of a basis, in the reference's order
 * -- Basis/BasisFull.hpp (nup < 0: the integers 0 .. 2^n - 1) or
 * Basis/BasisJz.hpp (the integers with nup set bits, increasing) -- as spins
 * sigma_i = 1 - 2 bit_i (toSigma, Utilities/Utility.hpp:91-99), weighted by
 * |psi|^2 / sum |psi|^2 over the basis (getProbs, RBM.hpp:443-451).  The
 * following yq_eloc returns the exact energy and variance and yq_sr_build
 * (explicit or dense CG solver, FP64) the exact O mean, F and S of
 * SRMatExact::constructExact / corrMat / energyGrad (SRMatExact.hpp:97-149).
 * n <= 24; *basis_size = number of configurations. */
YQ_API int yq_exact_batch(yq_ctx* ctx, int nup, int* basis_size);
/* The weights of the current exact batch (basis_size doubles). */
YQ_API int yq_exact_probs(yq_ctx* ctx, double* probs);
/* ExactSampler::sample (ExactSampler.hpp:49-77): n_samples configurations
 * drawn independently from |psi|^2 over the basis by inverse CDF (one Philox
 * uniform per draw, keyed by seed), as an ordinary unweighted batch;
 * indices (n_samples ints, may be null) receives the basis positions. */
YQ_API int yq_exact_sample(yq_ctx* ctx, int nup, int n_samples, uint64_t seed, int* indices);

/* ---- Energy: Observables/Energy.hpp + GroundState/utils.hpp:108-126 ---- */
/* E_loc of every sample (eloc may be null), mean = Re <E_loc>,
 * var = <Re E_loc^2> - mean^2: Energy.hpp:53-58 (finIter). */
YQ_API int yq_eloc(yq_ctx* ctx, double* eloc, double* mean, double* var);

/* ---- SR: GroundState/SRMat.hpp ----------------------------------------- */
/* constructFromSamples: O rows, column mean, centring, F = O_c^T E / Ns
 * (SRMat.hpp:89-103); with YQ_SOLVE_EXPLICIT also S (corrMat, :142-149).
 * Computes E_loc first if yq_eloc was not called on this batch. */
YQ_API int yq_sr_build(yq_ctx* ctx, int solver, int precision);
/* fisherAvg of RunRBM.hpp:143-144, 194-203 (ObsAvg<Matrix>, :13-36) on the
 * device: after an explicit yq_sr_build, avg = beta2 avg + (1 - beta2) S and
 * S := avg / (1 - beta2^t) for the t-th call since the last reset (reset =
 * 1, or a change of shape, starts from avg = 0, t = 0), so the following
 * yq_sr_get(YQ_SR_S) / yq_sr_apply / yq_sr_solve use the averaged matrix.
 * The average lives in the context across builds (dim x dim doubles). */
YQ_API int yq_sr_fisher_average(yq_ctx* ctx, double beta2, int reset);
/* Copy one of the built items to the host. */
YQ_API int yq_sr_get(yq_ctx* ctx, int item, double* out);
/* apply: O_c^T (O_c v) / Ns + shift v, SRMat.hpp:170-189. */
YQ_API int yq_sr_apply(yq_ctx* ctx, double shift, const double* v,
                       double* out);
/* Solve (S + shift I) x = rhs (rhs = F when null): solveExact
 * (SRMat.hpp:231-237, RunRBM.hpp:194-203) for the explicit solver,
 * solveCG (SRMat.hpp:197-223) otherwise.  resid = |(S+shift)x - rhs|/|rhs|,
 * iters = CG iterations (0 for an FP64 Cholesky, the refinement rounds for
 * a mixed explicit solve).  tol is the CG relative residual (for YQ_MIXED:
 * the outer FP64 refinement target, 1e-12 when tol <= 0 on the mixed
 * explicit solve). */
YQ_API int yq_sr_solve(yq_ctx* ctx, double shift, double tol, int max_iter,
                       const double* rhs, double* x, double* resid,
                       int* iters);

/* Options of the iterative solvers (no reference counterpart: the reference
 * runs Eigen's CG with IdentityPreconditioner from x = 0, which is the
 * default here).  precond 1: Jacobi preconditioner 1 / (diag S + shift),
 * diag S = Var(O_k) from the batch.  warm_start 1: start the next solve from
 * the previous solution.  Both leave the solved system and its tolerance
 * unchanged. */
YQ_API int yq_sr_set_options(yq_ctx* ctx, int precond, int warm_start);
/* Launch mode of the mixed-precision CG (no reference counterpart): 1, the
 * default, replays chunks of 8 iterations as a captured HIP graph that stops
 * itself on the device once converged; an even number 2..256 does the same
 * with chunks of that many iterations (the host polls the convergence flag
 * once per chunk); 0 issues the same kernels one by one (polling every 8
 * iterations).  All modes perform identical arithmetic and return identical
 * results. */
YQ_API int yq_set_graph_mode(yq_ctx* ctx, int use_graphs);
/* Tile-shape autotuning of the mixed-precision products (no reference
 * counterpart): 1 times every (forward column group, backward chunk) pair on
 * the first operands of each shape and keeps the fastest for the process;
 * 0, the default, uses fixed shapes.  The choice changes the order of the
 * float sums inside the products, so results agree to float rounding. */
YQ_API int yq_set_autotune(yq_ctx* ctx, int on);
/* Fix the tile shapes instead (reproducible sums across processes):
 * col_tiles in {2, 4, 8}, chunk in {128, 256, 512}; 0, 0 releases them. */
YQ_API int yq_set_tiles(yq_ctx* ctx, int col_tiles, int chunk);

/* ---- Runner: Runners/RunRBM.hpp:151-220 --------------------------------- */
typedef struct {
  int n_sweeps;      /* sample(nSweeps, .) */
  int n_discard;     /* sample(., nDiscard) */
  int randomize;     /* re-randomise chains before sampling (RunRBM.hpp:166) */
  int nup;           /* constraint for randomize, < 0 for none */
  int solver;        /* yq_solver */
  int precision;     /* yq_precision */
  double lambda;     /* diagonal shift for this step */
  double cg_tol;     /* CG / refinement relative residual */
  int cg_max_iter;
  double lr;         /* SGD step eta (already max(alpha/t^p, 1e-4)) */
  double grad_clip;  /* 0 = off (RunRBM.hpp:205-209) */
  int precond;       /* CG preconditioner, see yq_sr_set_options */
  int warm_start;    /* CG warm start, see yq_sr_set_options */
} yq_step_cfg;
typedef struct {
  double energy;     /* srm.eloc() */
  double energy_var;
  double update_norm;/* |v| before clipping */
  double cg_err;     /* |apply(v) - F| / |F| */
  int cg_iters;
  int n_samples;
  double sample_ms;  /* sampling time (device) */
  double solve_ms;   /* E_loc + SR build + solve + update (device) */
} yq_step_out;
/* One epoch body: [randomize], sample, constructFromSamples, solve, SGD
 * update, all on the device. */
YQ_API int yq_sr_step(yq_ctx* ctx, const yq_step_cfg* cfg, yq_step_out* out);

/* ---- Many machines in one context (no reference counterpart) -------------
 * A set of independent RBMs (each with its own Hamiltonian, chains and Philox
 * seed), as one would run many RunRBM jobs side by side.  Sampling is one
 * persistent launch over every (machine, chain) pair, ordered by cost so the
 * waves stay balanced across uneven sizes; local energies are one launch per
 * pass for all machines.  With the mixed factored solver (and no warm
 * start) the SR solves are grouped: every CG iteration is one forward
 * product, one backward product and one CG step launch for all machines,
 * each machine's blocks running its own tile shapes, with the FP64
 * refinement of each machine on its own stream; otherwise the solves run
 * concurrently, one host worker thread and one HIP stream per machine.  Each
 * machine behaves exactly like a single-machine context with the same
 * parameters and seed (the same chains, samples, energies and, grouped or
 * not, the same SR solution bit for bit). */
typedef struct {
  int n, m, use_bias;  /* RBM(n, m, useBias), real double parameters */
  int chains;          /* nChainsPer */
  int sweep_factor;    /* the sweeper's nSweep */
  uint64_t seed;       /* Philox key of this machine's chains */
  yq_ham_desc ham;
} yq_multi_desc;
/* Create count machines; all share the sweeper kind and sampling precision. */
YQ_API int yq_multi_create(yq_ctx* ctx, int count, const yq_multi_desc* descs, int sweeper,
                           int precision);
YQ_API int yq_multi_set_params(yq_ctx* ctx, int idx, const double* params);
YQ_API int yq_multi_get_params(yq_ctx* ctx, int idx, double* params);
/* initializeRandom for every machine; machine i is keyed by seed + i */
YQ_API int yq_multi_init_random(yq_ctx* ctx, uint64_t seed, double sigma);
/* sample(nSweeps, nDiscard) for every machine in one launch */
YQ_API int yq_multi_sample(yq_ctx* ctx, int n_sweeps, int n_discard);
YQ_API int yq_multi_samples_get(yq_ctx* ctx, int idx, int8_t* sigma, double* theta);
YQ_API int yq_multi_sampler_state(yq_ctx* ctx, int idx, int8_t* sigma, uint64_t* ctr,
                                  uint64_t* acc);
/* E_loc of every machine's batch; mean[i], var[i] (either may be null) */
YQ_API int yq_multi_eloc(yq_ctx* ctx, double* mean, double* var);
/* One epoch body (see yq_sr_step) for every machine: out[count]; sampling
 * and E_loc as above, then the per-machine SR on up to n_threads host
 * threads (0: one per machine, at most 16). */
YQ_API int yq_multi_sr_step(yq_ctx* ctx, const yq_step_cfg* cfg, yq_step_out* out,
                            int n_threads);
/* Grouped SR solves in yq_multi_sr_step on (1, the default) or off (0: one
 * CG per machine on its own stream).  In the grouped step every machine's
 * solve_ms is the wall time of the whole SR phase (build, solve, update). */
YQ_API int yq_multi_set_grouped(yq_ctx* ctx, int on);

/* ---- FeedForward machine: Machines/FeedForward.hpp --------------------- */
/* The network of the C4 configuration: Conv1D(1 -> channels, kernel width,
 * periodic, stride 1; layers/Conv1D.hpp) -> LnCosh -> FullyConnected(
 * channels*n -> hidden; layers/FullyConnected.hpp) -> Tanh -> Mean
 * (layers/Mean.hpp), all without bias (the layer defaults).  dtype 0 = float,
 * 1 = double.  Parameters (doubles on the host) in FeedForward's order:
 * [conv kernel, index ki + oc*kernel | FC weight W(i, o) at i + o*channels*n]. */
typedef struct {
  int n, channels, kernel, hidden;
  int dtype;
} yq_ff_desc;
/* FeedForward() + addLayer(...) x5: FeedForward.hpp:76-82 */
YQ_API int yq_ff_create(yq_ctx* ctx, const yq_ff_desc* desc);

/* A general layer stack (FeedForward.hpp:76-82, any sequence of addLayer
 * calls), one yq_ff_layer per addLayer:
 *  - YQ_LAYER_CONV1D  Conv1D(a = inChannels, b = outChannels, c = kernelSize,
 *    d = stride, use_bias) (layers/Conv1D.hpp:56-64, forward :144-171,
 *    backprop :173-225): periodic, input and output channel-major;
 *  - YQ_LAYER_FC      FullyConnected(a = inputDim, b = outputDim, use_bias)
 *    (layers/FullyConnected.hpp:57-58, forward :150-157, backprop :160-179);
 *  - YQ_LAYER_ACT     ActivationLayer<activation a> with its constructor
 *    argument param (ActivationLayer.hpp:42-136, activations.hpp: Tanh's
 *    constant, LeakyReLU's negative slope, LeakyHardTanh's outside slope,
 *    Cos's a; ignored by the others);
 *  - YQ_LAYER_MEAN    Mean({a, b}, c = axis) (layers/Mean.hpp:21-25, forward
 *    :92-103, backprop :106-125): the input as an a x b column-major matrix.
 * n is the input size (the spins); the last layer must leave one value
 * (FeedForward.hpp:165).  Parameters: the layers' getParams() concatenated
 * (Conv1D: kernel (ki + ic*kernelSize) + oc*inChannels*kernelSize, then the
 * bias; FullyConnected: weight i + o*inputDim, then the bias).  dtype 0 =
 * float, 1 = double.  The C4 stack (Conv1D(1, C, K) -> LnCosh ->
 * FullyConnected(C*n, H) -> Tanh -> Mean({H, 1}, 0), no biases) runs on the
 * kernels of yq_ff_create; every other stack on the general layer kernel.
 * Every yq_ff_* call below serves both. */
typedef enum { YQ_LAYER_CONV1D = 1, YQ_LAYER_FC = 2, YQ_LAYER_ACT = 3, YQ_LAYER_MEAN = 4 } yq_layer_kind;
typedef enum {
  YQ_ACT_IDENTITY = 0, YQ_ACT_LNCOSH = 1, YQ_ACT_TANH = 2, YQ_ACT_WEAKTANH = 3,
  YQ_ACT_SIGMOID = 4, YQ_ACT_RELU = 5, YQ_ACT_LEAKYRELU = 6, YQ_ACT_HARDTANH = 7,
  YQ_ACT_SOFTSHRINK = 8, YQ_ACT_LEAKYHARDTANH = 9, YQ_ACT_SOFTSIGN = 10, YQ_ACT_COS = 11
} yq_activation;
typedef struct {
  int kind;
  int a, b, c, d;
  int use_bias;
  double param;
} yq_ff_layer;
YQ_API int yq_ff_create_stack(yq_ctx* ctx, int n, int dtype, int n_layers,
                              const yq_ff_layer* layers);
/* Values per configuration that yq_ff_forward_dev stores in act (and
 * yq_ff_vjp_dev reads): 2*channels*n + hidden for the C4 network, else the
 * input and every layer's output (FeedForward::makeData, :170-184). */
YQ_API int yq_ff_act_len(const yq_ctx* ctx, int* len);
/* getDim: number of parameters */
YQ_API int yq_ff_npar(const yq_ctx* ctx, int* npar);
/* setParams / getParams / updateParams */
YQ_API int yq_ff_set_params(yq_ctx* ctx, const double* params);
YQ_API int yq_ff_get_params(yq_ctx* ctx, double* params);
YQ_API int yq_ff_update_params(yq_ctx* ctx, const double* delta);
/* initializeRandom(re, mode): FeedForward.hpp:43-73.  mode 0 LeCun, 1
 * Xavier (Gaussian sigma = sqrt(2/(fanIn+fanOut)), the reference's), 2 He,
 * 3 Xavier-uniform U(+-sqrt(6/(fanIn+fanOut))); Philox draws keyed by seed. */
YQ_API int yq_ff_init_random(yq_ctx* ctx, uint64_t seed, int mode);
/* forward(sigma) for ns configurations (FeedForward.hpp:152-168) */
YQ_API int yq_ff_forward(yq_ctx* ctx, int ns, const int8_t* sigma,
                         double* logpsi);
/* backward(makeData(sigma)) for ns configurations: rows of d log psi / d
 * params (ns x npar, FeedForward.hpp:186-205); logpsi may be null. */
YQ_API int yq_ff_jacobian(yq_ctx* ctx, int ns, const int8_t* sigma,
                          double* jac, double* logpsi);
/* Device-pointer forms for framework interop (e.g. a torch.autograd
 * Function).  Arrays are in the machine dtype on the context's device; the
 * caller finishes its own work on them before the call, which returns after
 * completion.  act (ns x yq_ff_act_len values) stores the activations in
 * yq_ff_forward_dev when non-null; yq_ff_vjp_dev reads them when non-null and
 * recomputes them otherwise.  grad = sum_s r_s d log psi_s / d params. */
YQ_API int yq_ff_forward_dev(yq_ctx* ctx, int ns, const int8_t* d_sigma,
                             void* d_logpsi, void* d_act);
YQ_API int yq_ff_jacobian_dev(yq_ctx* ctx, int ns, const int8_t* d_sigma,
                              void* d_jac, void* d_logpsi);
YQ_API int yq_ff_vjp_dev(yq_ctx* ctx, int ns, const int8_t* d_sigma,
                         const void* d_act, const void* d_r, void* d_grad);
/* yq_ff_vjp_dev without any device allocation.  No reference counterpart:
 * it is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  The per-block gradient partials live in
 * a fixed 16 MiB array that the library holds on every device it runs on
 * until the process ends, so npar x element size may be at most 16 MiB
 * (YQ_ERR_ARG beyond); with fewer blocks than yq_ff_vjp_dev uses when the
 * array is the limit, the result agrees with it to rounding.  Every call on a
 * device shares that array: calls that overlap on one GPU (several contexts
 * or host threads) race. */
YQ_API int yq_ff_vjp_dev_noalloc(yq_ctx* ctx, int ns, const int8_t* d_sigma,
                                 const void* d_act, const void* d_r, void* d_grad);

/* ---- AmplitudePhase machine: Machines/AmplitudePhase.hpp ---------------
 * psi(sigma) = sqrt(RBM(sigma)) exp(i pi FF(sigma)) (:127-138): the context's
 * real RBM (yq_rbm_create, YQ_DTYPE_F64) is the amplitude and its
 * FeedForward machine (yq_ff_create, dtype 1, same n) the phase; the
 * reference builds the same pair in its constructor (:37-40).  While enabled
 * the context's calls act on psi:
 *  - yq_sample draws from |psi|^2 = RBM(sigma), i.e. Metropolis on the RBM at
 *    beta = 1/2 (logRatioRe = logRatio / 2, States/AmplitudePhaseState.hpp:
 *    52-55); parallel tempering is not available;
 *  - yq_exact_batch / yq_exact_sample weight by RBM(sigma) / sum;
 *  - yq_eloc returns complex local energies (ns (re, im) pairs), the ratios
 *    being exp(logRatio / 2 + i pi (FF(sigma') - FF(sigma))) (:57-120);
 *  - yq_sr_build (YQ_SOLVE_EXPLICIT or YQ_SOLVE_CG, YQ_FP64) builds the real
 *    SR system of the concatenated parameters [RBM | FF] (:64-77, 153-158):
 *    rows O = [O_RBM / 2 | pi dFF/dp] (:42-58; the phase block multiplies i),
 *    S = Re(<O^* O^T> - <O^*><O^T>), which is block-diagonal (the blocks are
 *    GroundState/NGDExact.hpp:177-189's corrMatAmp and corrMatPhase), and
 *    F = Re(<O^* E_loc> - <O^*><E_loc>) = [cov(O_amp, Re E_loc) |
 *    cov(O_phase, Im E_loc)], half of NGDExact::energyGrad (NGDExact.hpp:139-145);
 *    yq_sr_get / yq_sr_apply / yq_sr_solve act on that system (vectors of
 *    dim_amp + dim_phase doubles; the explicit solve factors the two blocks
 *    separately, as RunAmplitudePhase.hpp:96-111 does);
 *  - yq_sr_step updates both halves.
 * yq_rbm_create or yq_ff_create disables it. */
YQ_API int yq_ap_enable(yq_ctx* ctx, int on);
/* getDimAmp / getDimPhase (:60-68) */
YQ_API int yq_ap_dims(const yq_ctx* ctx, int* dim_amp, int* dim_phase);
/* parameters [RBM | FF] (dim_amp + dim_phase doubles); updateParams :153-158 */
YQ_API int yq_ap_get_params(yq_ctx* ctx, double* params);
YQ_API int yq_ap_set_params(yq_ctx* ctx, const double* params);
YQ_API int yq_ap_update_params(yq_ctx* ctx, const double* delta);
/* log coeff(sigma) = (log RBM(sigma) / 2, pi FF(sigma)) per configuration:
 * out holds ns (re, im) pairs (:127-138) */
YQ_API int yq_ap_log_psi(yq_ctx* ctx, int ns, const int8_t* sigma, double* out);

/* ---- instrumentation (no reference counterpart) ------------------------- */
/* Time every launch of the named kernels with HIP events on the launch's
 * stream.  labels: comma-separated kernel labels (e.g. "sample,bwd_mfma"),
 * NULL or "" to disable.  yq_prof_read returns the accumulated device time
 * and launch count of one label and resets it. */
YQ_API int yq_prof_enable(yq_ctx* ctx, const char* labels);
YQ_API int yq_prof_read(yq_ctx* ctx, const char* label, double* total_ms,
                        long long* launches);

#ifdef __cplusplus
}
#endif
#endif /* YANNQ_AMD_H */
