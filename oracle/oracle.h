// CPU oracle for the VMC / SR step.  TEST INFRASTRUCTURE ONLY: it is loaded
// by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
// always as the checker or the timed CPU baseline, never by the library.
//
// Two layers:
//  * or_*  restate the reference algorithm in double precision with libm and
//          the reference's own summation order (citations per function).
//  * rp_*  replay the exact arithmetic the gfx950 kernels perform (canonical
//          lane-strided + tree reductions, the replayable exp/logcosh of
//          yannq_amd/csrc/yq_math.h, the same Philox counters), so spin
//          configurations, accept decisions and local energies can be
//          compared bit for bit.
//  * bl_*  the multi-threaded CPU baseline (std::thread over chains and
//          samples, mirroring the reference's TBB parallel_for sites).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_TFI = 0, OR_XXZ = 1, OR_J1J2 = 2 };

typedef struct {
  int kind;        // OR_TFI / OR_XXZ / OR_J1J2
  int n;           // sites (periodic chain)
  double J;        // TFI: J; XXZ: J; J1J2: J1
  double h;        // TFI: h (coefficient of the ratio term)
  double delta;    // XXZ: Delta
  double J2;       // J1J2: J2
  int sign_rule;   // XXZ / J1J2 sign rule (1 -> sign = -1)
} or_ham;

// ---- restatement (reference order, libm) ---------------------------------
double or_logcosh(double x);
void or_theta(int n, int m, const double* W, const double* b,
              const int8_t* sigma, double* theta);
double or_log_ratio1(int n, int m, const double* W, const double* a,
                     const int8_t* sigma, const double* theta, int k);
double or_log_ratio2(int n, int m, const double* W, const double* a,
                     const int8_t* sigma, const double* theta, int k, int l);
double or_log_coeff(int n, int m, const double* a, const int8_t* sigma,
                    const double* theta);
double or_log_ratio_vec(int n, int m, const double* W, const double* a, const int8_t* sigma,
                        const double* theta, int k, const int* v);
void or_logderiv(int n, int m, int use_bias, const int8_t* sigma,
                 const double* theta, double* out);
double or_eloc(const or_ham* H, int n, int m, const double* W, const double* a,
               const int8_t* sigma, const double* theta);
int or_ham_column(const or_ham* H, uint32_t col, uint32_t* rows, double* vals,
                  int cap);
// O (ns x dim row-major), centred in place; omean (dim); F (dim);
// S (dim x dim) if non-null; weights (ns) optional (null -> 1/ns, the
// unweighted SRMat::constructFromSamples path).
void or_sr_build(int ns, int n, int m, int use_bias, const int8_t* sigma,
                 const double* theta, const double* eloc, const double* weights,
                 double* O, double* omean, double* F, double* S);
int or_llt_solve(int dim, const double* S, double shift, const double* F,
                 double* x);
int or_cg_solve(int ns, int dim, const double* Oc, const double* weights,
                double shift, const double* b, double tol, int max_iter,
                double* x, int* iters, double* err);
void or_sgd_update(int dim, double alpha, double p, int t, const double* v,
                   double* out);

// ---- complex RBM (std::complex<double>, interleaved re/im arrays) ----------
// Parameters in the RBM layout [W col-major | a | b] as complex numbers.
void or_c_logcosh(double re, double im, double* out2);
void or_c_theta(int n, int m, const double* W, const double* b, const int8_t* sigma,
                double* theta);
void or_c_log_ratio1(int n, int m, const double* W, const double* a, const int8_t* sigma,
                     const double* theta, int k, double* out2);
void or_c_log_ratio2(int n, int m, const double* W, const double* a, const int8_t* sigma,
                     const double* theta, int k, int l, double* out2);
void or_c_log_coeff(int n, int m, const double* a, const int8_t* sigma, const double* theta,
                    double* out2);
void or_c_log_ratio_vec(int n, int m, const double* W, const double* a, const int8_t* sigma,
                        const double* theta, int k, const int* v, double* out2);
void or_c_logderiv(int n, int m, int use_bias, const int8_t* sigma, const double* theta,
                   double* out);
void or_c_eloc(const or_ham* H, int n, int m, const double* W, const double* a,
               const int8_t* sigma, const double* theta, double* out2);
// O (ns x dim complex), centred in place; omean; F = O_c^H W E; S = O_c^H W
// O_c (if S); W = diag(weights) (null -> 1/ns, SRMat::constructFromSamples).
void or_c_sr_build(int ns, int n, int m, int use_bias, const int8_t* sigma, const double* theta,
                   const double* eloc, const double* weights, double* O, double* omean,
                   double* F, double* S);

// ---- FeedForward: Conv1D(1->C, K, periodic) -> LnCosh -> FC(C*n -> H) ->
// Tanh -> Mean, no biases (the layer defaults).  Parameters [conv kernel
// (K*C, index ki + oc*K) | FC weight (C*n x H column-major)].
// A general layer stack (nl > 0: channels, kernel and hidden are unused;
// or_ff_npar / or_ff_forward / or_ff_jacobian and the AmplitudePhase
// functions below then run the stack of oracle_ffs.cpp).
typedef struct {
  int kind;  // 1 Conv1D(a in, b out, c kernel, d stride), 2 FullyConnected(a in, b out),
             // 3 activation a (param: Tanh / Cos constant, LeakyReLU / LeakyHardTanh slope),
             // 4 Mean({a, b}, axis c)
  int a, b, c, d;
  int use_bias;
  double param;
} or_ff_layer;
typedef struct {
  int n, channels, kernel, hidden;
  int nl;
  const or_ff_layer* layers;
} or_ff_desc;
// general stack on n inputs: 0 when the shapes chain to one output
int or_ffs_check(int n, int nl, const or_ff_layer* layers);
int or_ffs_npar(int nl, const or_ff_layer* layers);
double or_ffs_forward(int n, int nl, const or_ff_layer* layers, const double* params,
                      const int8_t* sigma);
double or_ffs_jacobian(int n, int nl, const or_ff_layer* layers, const double* params,
                       const int8_t* sigma, double* der);
int or_ff_npar(const or_ff_desc* d);
// log psi(sigma) = network output (FeedForward.hpp:152-168)
double or_ff_forward(const or_ff_desc* d, const double* params,
                     const int8_t* sigma);
// d log psi / d params (FeedForward.hpp:186-205); returns log psi
double or_ff_jacobian(const or_ff_desc* d, const double* params,
                      const int8_t* sigma, double* der);

// ---- AmplitudePhase: psi = sqrt(RBM) exp(i pi FF) with a real RBM [W | a | b]
// (a, b zero without bias) and the FeedForward network above.
// out2 = (Re, Im) of log coeff (Machines/AmplitudePhase.hpp:127-138)
void or_ap_log_psi(int n, int m, const double* W, const double* a, const double* b,
                   const or_ff_desc* d, const double* ffp, const int8_t* sigma, double* out2);
// log psi(sigma^{k,l}) - log psi(sigma) (l < 0: k only), AmplitudePhaseState.hpp:57-120
void or_ap_log_ratio(int n, int m, const double* W, const double* a, const or_ff_desc* d,
                     const double* ffp, const int8_t* sigma, const double* theta, int k, int l,
                     double* out2);
// complex local energy (re, im)
void or_ap_eloc(const or_ham* H, int n, int m, const double* W, const double* a,
                const or_ff_desc* d, const double* ffp, const int8_t* sigma,
                const double* theta, double* out2);
// real rows [RBM logDeriv / 2 | pi FF backward] (AmplitudePhase.hpp:42-58)
void or_ap_logderiv(int n, int m, int use_bias, const or_ff_desc* d, const double* ffp,
                    const int8_t* sigma, const double* theta, double* out);
// NGDExact::constructExact + corrMatAmp / corrMatPhase (NGDExact.hpp:137-184)
int or_ap_ngd_exact(const or_ham* H, int n, int m, int use_bias, const double* rbm_params,
                    const or_ff_desc* d, const double* ffp, const uint32_t* basis, int B,
                    double* E, double* olocAmp, double* olocPhase, double* grad, double* Samp,
                    double* Sphase);

// ---- replay of the kernels' arithmetic -------------------------------------
typedef struct {
  int n, m, use_bias;
  int kind;          // 0 local single-spin flips, 1 exchange (swap) moves
  int precision;     // 0 fp64, 1 mixed (fp32 theta refreshed from fp64)
  int lanes;         // reduction width of one chain (power of two)
  int sweep_factor;  // proposals per sweep = n * sweep_factor
  double beta;
  uint64_t seed;
} rp_cfg;

void rp_theta(int n, int m, const double* W, const double* b,
              const int8_t* sigma, double* theta);
double rp_logcosh_d(double x);
float rp_logcosh_f(float x);
double rp_exp_d(double x);
float rp_exp_f(float x);
void rp_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
               uint32_t k1, uint32_t* out4);
// chains: C; stream ids; W,a,b fp64 master parameters; sigma C x n in/out;
// ctr, acc per chain in/out; samples (n_sweeps*C rows) out.
int rp_sample(const rp_cfg* cfg, int C, const uint32_t* stream,
              const double* W, const double* a, const double* b,
              int8_t* sigma, uint64_t* ctr, uint64_t* acc, int n_sweeps,
              int n_discard, int8_t* out_sigma, double* out_theta);
// Parallel tempering replay: T temperatures per chain group, groups groups,
// C = T * groups chains (sigma, ctr, acc, stream of all C); n_sweeps * groups
// samples (the beta = 1 slot of every group) out.
int rp_sample_pt(const rp_cfg* cfg, int T, int groups, const uint32_t* stream,
                 const double* W, const double* a, const double* b, int8_t* sigma,
                 uint64_t* ctr, uint64_t* acc, int n_sweeps, int n_discard,
                 int8_t* out_sigma, double* out_theta);
// Local energies with the kernel's arithmetic. precision 0: fp64, 1: fp32.
void rp_eloc(const or_ham* H, int precision, int ns, int n, int m,
             const double* W, const double* a, const int8_t* sigma,
             const double* theta, double* eloc);
void rp_randomize(uint64_t seed, uint64_t epoch, int C, const uint32_t* stream,
                  int n, int nup, int8_t* sigma);
// Replay of the complex sampler: decisions on Re(log ratio) with the
// replayable Re logcosh; W, a, b, out_theta complex interleaved (re, im).
int rp_c_sample(const rp_cfg* cfg, int C, const uint32_t* stream, const double* W,
                const double* a, const double* b, int8_t* sigma, uint64_t* ctr, uint64_t* acc,
                int n_sweeps, int n_discard, int8_t* out_sigma, double* out_theta);
// Parallel tempering replay of a complex machine (the arrangement of
// rp_sample_pt, lp = Re a . sigma + sum_j Re logcosh(theta_j)).
int rp_c_sample_pt(const rp_cfg* cfg, int T, int groups, const uint32_t* stream,
                   const double* W, const double* a, const double* b, int8_t* sigma,
                   uint64_t* ctr, uint64_t* acc, int n_sweeps, int n_discard,
                   int8_t* out_sigma, double* out_theta);
// Complex local energies with the kernels' arithmetic (precision 0: fp64, 1:
// fp32 theta and log-ratios); eloc complex interleaved.  The phase uses the
// host atan2, so the device agrees to a few ulp, not bit for bit.
void rp_c_eloc(const or_ham* H, int precision, int ns, int n, int m, const double* W,
               const double* a, const int8_t* sigma, const double* theta, double* eloc);
// The kernels' complex functions: Re log cosh (replayed bit for bit), the
// full log cosh (out2 = re, im) and tanh of a complex argument.
double rp_logcosh_re_d(double x, double y);
float rp_logcosh_re_f(float x, float y);
void rp_logcosh_c_d(double x, double y, double* out2);
void rp_logcosh_c_f(float x, float y, float* out2);
void rp_tanh_c_d(double x, double y, double* out2);

// ---- multi-threaded CPU baseline --------------------------------------------
int bl_threads(void);
int bl_sample(const rp_cfg* cfg, int C, const uint32_t* stream,
              const double* W, const double* a, const double* b,
              int8_t* sigma, uint64_t* ctr, uint64_t* acc, int n_sweeps,
              int n_discard, int8_t* out_sigma, double* out_theta,
              int nthreads);
void bl_eloc(const or_ham* H, int ns, int n, int m, const double* W,
             const double* a, const int8_t* sigma, const double* theta,
             double* eloc, int nthreads);
void bl_sr_build(int ns, int n, int m, int use_bias, const int8_t* sigma,
                 const double* theta, const double* eloc, double* O,
                 double* omean, double* F, int nthreads);
int bl_cg_solve(int ns, int dim, const double* Oc, double shift,
                const double* b, double tol, int max_iter, double* x,
                int* iters, double* err, int nthreads);

#ifdef __cplusplus
}
#endif
