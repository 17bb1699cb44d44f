// Internal declarations shared by the kernels and the C-ABI implementation.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/yannq_amd.h"

namespace yq {

// Per-kernel HIP-event timing, enabled by label through yq_prof_enable.  A
// ProfGuard placed around a launch records a start/stop event pair on the
// launch's stream when its label is enabled, and costs one branch otherwise.
bool prof_enabled(const char* label);
void prof_begin(const char* label, hipStream_t st);
void prof_end(const char* label, hipStream_t st);
struct ProfGuard {
  const char* label;
  hipStream_t st;
  bool on;
  ProfGuard(const char* l, hipStream_t s) : label(l), st(s), on(prof_enabled(l)) {
    if (on) prof_begin(label, st);
  }
  ~ProfGuard() {
    if (on) prof_end(label, st);
  }
};
// registers an event pair that brackets one launch
void prof_pair(const char* label, hipEvent_t start, hipEvent_t stop);
// A launch timed, when its label is enabled, by events bound to the kernel's
// own dispatch (hipExtLaunchKernel: start and end of the kernel, without the
// queue time between separately recorded events and the kernel).
template <typename F, typename... Args>
void launch_prof(const char* label, F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t st,
                 Args... args) {
  if (prof_enabled(label)) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
      hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)lds, st, e0, e1, 0u, args...);
      prof_pair(label, e0, e1);
      return;
    }
    if (e0) (void)hipEventDestroy(e0);
  }
  hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

constexpr int kWave = 64;
// One wave per chain keeps up to 8 theta registers per lane (m <= 512); a
// wider machine runs one chain per block of G = 64 NW threads (chain_group),
// still at most 8 registers per thread, up to kMaxHidden hidden units.
constexpr int kWaveHidden = 512;
constexpr int kMaxHidden = 8192;
constexpr int kMaxSites = 4096;

// Threads that share one chain's hidden units for m hidden units: 64 (one
// wave) up to m = 512, else 64 NW with NW the smallest power of two that
// leaves at most 8 units per thread.  Hidden unit j lives in thread j % G of
// the group; the chain's log-ratio is the canonical tree over G lanes.
inline int chain_group(int m) {
  int g = kWave;
  while (g * 8 < m) g *= 2;
  return g;
}

// Sum of v over the G threads of a chain group in the canonical order: the
// tree with strides G/2, ..., 1 over the per-thread partials (what an xor
// butterfly produces within one wave; addition is commutative, so every
// thread ends with the same bits).  G = 64: the butterfly.  G > 64 (the
// group is the whole block): the waves' partials meet in LDS (red: 2 G
// values, alternating halves by call parity, so one barrier per call
// suffices: a half is written again only after every wave has passed the
// next call's barrier, i.e. finished reading it), every wave forms the tree
// levels above 64 for its own lane, then the butterfly.  Every thread of the
// block must call it.
template <typename R, int G>
__device__ __forceinline__ R group_sum(R v, R* red, int& par) {
  if constexpr (G > kWave) {
    R* buf = red + par * G;
    buf[threadIdx.x] = v;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    constexpr int NW = G / kWave;
    R w[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) w[q] = buf[lane + kWave * q];
#pragma unroll
    for (int s = NW / 2; s >= 1; s /= 2)
#pragma unroll
      for (int q = 0; q < s; ++q) w[q] = w[q] + w[q + s];
    v = w[0];
    par ^= 1;
  }
#pragma unroll
  for (int s = kWave / 2; s >= 1; s >>= 1) v = v + __shfl_xor(v, s);
  return v;
}

// ------------------------------------------------------------------ sampler
struct SampleArgs {
  const double* W64;  // m x n column-major, fp64 master
  const double* b64;  // m
  const void* Wr;     // working-precision W (double or float), m x n col-major
                      // (complex machine: real part)
  const void* Wi;     // complex machine: imaginary part of W, else unused
  const void* ar;     // working-precision a (n) (complex machine: Re a)
  int8_t* sigma;      // C x n chain state
  uint64_t* ctr;      // C proposal counters
  uint64_t* acc;      // C accepted-proposal counters
  int8_t* out_sigma;  // n_sweeps*C x n
  void* out_theta;    // n_sweeps*C x m (double for fp64, float for mixed)
  int C, n, m, nprop, n_sweeps, n_discard, npad, kind, mixed;
  double beta;
  uint64_t seed;
  int ntemps;  // parallel tempering: C = ntemps x chain groups (1: plain chains)
  // Real machine, ntemps == 1: theta of every chain's spins comes from the
  // MFMA theta pass (theta_buf: C x m scratch in the working precision) and
  // k_sample reads it instead of summing W itself; null: k_sample sums.
  void* theta_buf;
  const void* theta_in;  // set by launch_sample per launch
  int cplx;              // complex machine (Wr / Wi planes, interleaved theta)
};
void launch_sample(const SampleArgs& a, hipStream_t st);
// LDS of the parallel-tempering launch (spin strips, exchange areas)
size_t pt_lds_bytes(const SampleArgs& a);
// Many machines in one persistent launch: inst_dev holds one SampleArgs per
// machine (device memory), items the (machine, chain) pairs in decreasing
// cost order; all machines share the sweeper kind and precision.
void launch_sample_multi(const SampleArgs* inst_dev, const int2* items, int n_items, int npad,
                         int kind, int mixed, int blocks, hipStream_t st);

// theta = W sigma + b for a batch on FP64 MFMA (the sum over the spins in
// index order, exactly as a sequential FP64 loop adds it, then + b, optionally
// rounded to float); out is double or float.
// logRatio(vector) (RBMState.hpp:108-126) of ns configurations: out[s] (a
// complex pair when cplx) for flipping the up to k sites listed in row s of
// sites (entries < 0 skipped); W, a, theta in the machine's layout (complex
// interleaved when cplx), theta = W sigma + b of the unflipped rows.
void launch_log_ratio_vec(const double* W, const double* a, const int8_t* sigma,
                          const double* theta, const int* sites, int ns, int n, int m, int k,
                          int cplx, double* out, hipStream_t st);
void launch_theta_batch(const double* W64, const double* b64,
                        const int8_t* sigma, int ns, int n, int m, void* out,
                        int out_float, hipStream_t st);

// ------------------------------------------------------------------ E_loc
// Sites of off-diagonal term b in the order the E_loc functors visit them
// (TFIsing.hpp:33-37 single flips; XXZ.hpp:37-49 and the first n terms of
// J1J2.hpp:35-55 nearest-neighbour pairs; then J1J2's next-nearest pairs);
// l = -1 for a single flip.
__device__ inline void bond_sites(int kind, int n, int b, int& k, int& l) {
  if (kind == YQ_TFI) {
    k = b;
    l = -1;
  } else if (kind == YQ_XXZ || b < n) {
    k = b;
    l = (b + 1) % n;
  } else {
    k = b - n;
    l = (b - n + 2) % n;
  }
}
struct ElocArgs {
  const int8_t* sigma;  // Ns x n
  const void* theta;    // Ns x m (double or float)
  const void* WT;       // m x n row-major (WT[j*n+i] = W_ji), double or float
                        // (complex machine: real part)
  const void* WTi;      // complex machine: imaginary part of WT
  const void* ar;       // n, double or float (complex machine: Re a)
  const void* ai;       // complex machine: Im a
  double* ratio;        // Ns x B (complex machine: interleaved complex)
  double* eloc;         // Ns (complex machine: interleaved complex)
  int ns, n, m, B, use_float;
  yq_ham_desc ham;
  // AmplitudePhase machine (real RBM amplitude, FeedForward phase): the phase
  // network's output for every sample (ns) and for the neighbour of every
  // (sample, term) pair (ns x B); then ratio holds interleaved complex
  // exp(logRatio / 2 + i pi (ap_ffb - ap_ffs)).  Null otherwise.
  const double* ap_ffs;
  const double* ap_ffb;
};
int num_bonds(const yq_ham_desc& h);
void launch_eloc(const ElocArgs& a, hipStream_t st);
// the ratio pass alone (ratios of every (sample, term) pair)
void launch_eloc_ratio(const ElocArgs& a, hipStream_t st);
// LDS of the single-machine ratio launch: theta and logcosh(theta) of the
// block's samples are staged when they fit kElocStageBytes (eloc_staged),
// else only the spins
constexpr size_t kElocStageBytes = 48 * 1024;
bool eloc_staged(const ElocArgs& a);
size_t eloc_lds_bytes(const ElocArgs& a);
// LDS of the staged form whatever its size (the many-machine launch)
size_t eloc_stage_lds_bytes(const ElocArgs& a);
// Many machines in one launch per pass: inst_dev holds one ElocArgs per
// machine; stats[2 i], stats[2 i + 1] = mean and variance of machine i.
void launch_eloc_multi(const ElocArgs* inst_dev, int count, int max_blocks, int max_ns,
                       size_t lds, int use_float, double* stats, hipStream_t st);
// mean and variance (Energy.hpp:65-70) into out[0], out[1] on the device
void launch_eloc_stats(const double* eloc, int ns, double* out,
                       hipStream_t st);

// ------------------------------------------------------------------ SR
void launch_tanh(const void* theta, void* T, size_t count, int use_float,
                 hipStream_t st);
void launch_build_O(const int8_t* sigma, const double* T, int ns, int n, int m,
                    int use_bias, double* O, hipStream_t st);
// partial column sums of X (ns x dim row-major) over sample chunks, of
// w_s X[s][c] when w is given, of X^2 when square
int colsum_chunks(int ns);
void launch_colsum(const double* X, int ns, int dim, double* partial,
                   hipStream_t st, const double* w = nullptr, int square = 0);
// out[c] = scale * sum_k partial[k][c]
void launch_reduce_partials(const double* partial, int chunks, int dim,
                            double scale, double* out, hipStream_t st);
// O_c = rs_s (O - omean) in place (rs_s = 1 without rs) and
// partial[k][c] = sum_{s in chunk k} O_c[s][c] rs_s e[s]
void launch_center_force(double* O, const double* omean, const double* e,
                         int ns, int dim, double* partial, hipStream_t st,
                         const double* rs = nullptr);
void launch_add_diag(double* M, int dim, double shift, hipStream_t st);
void launch_copy_lower_shift(const double* S, double* M, int dim,
                             double shift, hipStream_t st);

// factored O (sigma, T), fp64: y = O v (uncentred), length ns
// scratch: factored_partial_doubles(ns, n, m) doubles (shared with the
// backward product's partials)
size_t factored_partial_doubles(int ns, int n, int m);
// tb: the augmented column of T' (1; 0 for the imaginary plane of a complex
// machine's T, whose bias column is real)
void launch_factored_fwd64(const int8_t* sigma, const double* T, int ns, int n,
                           int m, int use_bias, const double* v, double* y,
                           double* scratch, hipStream_t st, double tb = 1.0);
// g = O^T r (uncentred), length dim; partial has colsum_chunks(ns)*dim doubles
void launch_factored_bwd64(const int8_t* sigma, const double* T, int ns, int n,
                           int m, int use_bias, const double* r, double* g,
                           double* partial, hipStream_t st, double tb = 1.0);
// column mean of the factored O: omean = O^T 1 / ns
void launch_factored_mean64(const int8_t* sigma, const double* T, int ns,
                            int n, int m, int use_bias, double* omean,
                            double* partial, hipStream_t st, double tb = 1.0);

// diag(S) = Var(O_c) per parameter from T = tanh(theta) and the column mean;
// partial needs (colsum_chunks(ns) + 1) * m doubles
void launch_diag_S(const double* T, int ns, int n, int m, int use_bias, const double* omean,
                   double* diag, double* partial, hipStream_t st);
void launch_jacobi_inv64(const double* diag, double shift, double* out, size_t n,
                         hipStream_t st);  // out = 1 / (diag + shift)
void launch_mul(const double* a, const double* b, double* out, size_t n,
                hipStream_t st);  // out = a * b

// ------------------------------------------------------------------ complex RBM
// Complex parameters are interleaved (re, im) doubles in the [W | a | b]
// layout; the kernels read these planes of them.
struct CplxPlanes {
  double *Wr64, *Wi64, *WTr64, *WTi64, *ar64, *ai64;  // W: i*m+j, WT: j*n+i
  float *Wr32, *Wi32, *WTr32, *WTi32, *ar32, *ai32;
};
void launch_cplx_planes(const double* params, int n, int m, const CplxPlanes& P,
                        hipStream_t st);
// sampler with W64 / b64 interleaved, Wr / Wi / ar planes; theta out interleaved
void launch_sample_c(const SampleArgs& a, hipStream_t st);
void launch_theta_batch_c(const double* W64, const double* b64, const int8_t* sigma, int ns,
                          int n, int m, void* out, int out_float, hipStream_t st);
size_t eloc_c_lds_bytes(const ElocArgs& a);
void launch_eloc_c(const ElocArgs& a, hipStream_t st);
// the functor accumulation over interleaved complex ratios (a.ratio) alone
void launch_eloc_combine_c(const ElocArgs& a, hipStream_t st);
void launch_eloc_stats_c(const double* eloc, int ns, double* out, hipStream_t st);
// T = tanh(theta), complex, from an interleaved double or float theta
void launch_tanh_c(const void* theta, int in_float, double* T, size_t count, hipStream_t st);
// factored complex SR operator pieces (complex.hip): interleaved complex <->
// real and imaginary planes, the combinations of the four real forward and
// backward products, a one-block complex dot (conj(a) . b, or sum a when b is
// null) and the conjugate of a split vector, interleaved
void launch_split_c(const double* z, size_t n, double* re, double* im, hipStream_t st);
void launch_cfwd_combine(const double* P1, const double* P2, const double* P3, const double* P4,
                         const double* ov, size_t n, double scale, double* yr, double* yi,
                         double* yz, hipStream_t st);
void launch_cbwd_combine(const double* B1, const double* B2, const double* B3, const double* B4,
                         const double* a, const double* sy, const double* v, double shift,
                         size_t n, double* out, hipStream_t st);
void launch_cdot1(const double* a, const double* b, size_t n, double* out, hipStream_t st);
void launch_center_c(const double* e, const double* esum, size_t n, double* yr, double* yi,
                     hipStream_t st);
void launch_conj_merge(const double* re, const double* im, size_t n, double* z, hipStream_t st);
// Oh = conj(O), ns x dim complex row-major
void launch_build_Oh(const int8_t* sigma, const double* T, int ns, int n, int m, int use_bias,
                     double* Oh, hipStream_t st);
// Oh = rs_s (Oh - omean_h); partial[k][c] = sum_{s in chunk k} Oh_sc rs_s e_s
// (complex; rs_s = 1 without rs)
void launch_center_force_c(double* Oh, const double* omean_h, const double* e, int ns, int dim,
                           double* partial, hipStream_t st, const double* rs = nullptr);
// S = sum of np column-major dim x dim partial products, lower triangle
// (nc = 2: complex entries as pairs)
void launch_sum_lower(const double* P, int np, size_t dim, int nc, double* S, hipStream_t st);
// lower triangle: M = float(S) + shift I
void launch_lower_shift_f(const double* S, float* M, size_t dim, double shift, hipStream_t st);
// lower triangle: avg = beta avg + (1 - beta) S, S = avg / corr
void launch_ema_lower(double* avg, double* S, size_t dim, int nc, double beta, double corr,
                      hipStream_t st);
// S = alpha O_c^T O_c, lower triangle of the dim x dim column-major S, on
// FP64 MFMA (syrk.hip).  Factored: O_c[s][c] = (sigma'_{s,i(c)} T'_{s,j(c)} -
// omean_c) rs_s (rs may be null) from the int8 spins and T = tanh(theta)
// (ns x m); dense: O_c row-major ns x dim.  chunks > 1 splits the samples
// into that many chunks whose partial products (in part, chunks x dim x dim
// doubles) are summed in chunk order; syrk64_chunks picks the count for a
// shape (1 unless the tiles are too few to fill the chip) within
// part_cap_doubles.
int syrk64_chunks(int ns, size_t dim, size_t part_cap_doubles);
int launch_syrk64_factored(const int8_t* sigma, const double* T, const double* omean,
                           const double* rs, int ns, int n, int m, int use_bias, double alpha,
                           double* S, int chunks, double* part, hipStream_t st);
// The same S in split-bf16 (mixed precision): O_c rounded to float, three
// bf16 pieces, six pair products per sample on bf16 MFMA with fp32
// accumulation, stored in double.  rs must be null (YQ_ERR_UNSUPPORTED).
int launch_syrk_bf16x3(const int8_t* sigma, const double* T, const double* omean,
                       const double* rs, int ns, int n, int m, int use_bias, double alpha,
                       double* S, int chunks, double* part, hipStream_t st);
int launch_syrk64_dense(const double* O, int ns, size_t dim, double alpha, double* S, int chunks,
                        double* part, hipStream_t st);
// the split-bf16 S of a dense row-major O (ns x dim)
int launch_syrk_bf16x3_dense(const double* O, int ns, size_t dim, double alpha, double* S,
                             int chunks, double* part, hipStream_t st);
// complex S = A A^H (lower, interleaved, dim x dim) from XX = X^T X (lower,
// 2 dim x 2 dim) of the same rows read as 2 dim reals
void launch_herk_combine(const double* XX, size_t dim, double* S, hipStream_t st);
// ---------------------------------------------------------------- exact sums
// full-basis / sector configurations as spins, |psi|^2 weights (and their
// row scales sqrt(B w)), weighted E_loc statistics, the CDF and the draws of
// the exact sampler (exact.hip)
int exact_partials(int B);
int exact_scan_tiles(int B);
void launch_basis_spins(const uint32_t* basis, int B, int n, int8_t* sigma, hipStream_t st);
// (weights |psi|^2 ~ exp(pw Re log psi), pw = 2; 1 for an AmplitudePhase amplitude)
void launch_exact_weights(const int8_t* sigma, const double* theta, const double* a, int B, int n,
                          int m, int cplx, double* logpsi, double* w, double* rs, double* part,
                          hipStream_t st, double pw = 2.0);
// log psi of an RBM batch (a: visible bias or null; complex: (re, im) pairs)
void launch_log_psi(const int8_t* sigma, const double* theta, const double* a, int B, int n, int m,
                    int cplx, double* logpsi, hipStream_t st);
void launch_eloc_stats_w(const double* e, const double* w, int ns, int cplx, double* out,
                         hipStream_t st);
void launch_exact_cdf(const double* w, int B, double* cdf, double* tile_sums, hipStream_t st);
void launch_exact_draw(const double* cdf, int B, const uint32_t* basis, int n, int count,
                       uint64_t seed, int8_t* sigma, int* idx, hipStream_t st);
// partial[k][c] = sum_{s in chunk k} |Oh_sc|^2 (real)
void launch_absq_colsum(const double* Oh, int ns, int dim, double* partial, hipStream_t st);
void launch_dup2(const double* x, double* y, size_t n, hipStream_t st);  // y[2i] = y[2i+1] = x[i]
// M = S + shift I, complex dim x dim
void launch_copy_shift_c(const double* S, double* M, int dim, double shift, hipStream_t st);

// ------------------------------------------------------------------ FeedForward
struct FFArgs {
  int n, C, K, H, ns;
  const void* params;   // device, machine precision
  const void* wt;       // the FC weight transposed, WT[i * H + o] = W(i, o) (forward reads)
  const int8_t* sigma;  // ns x n
  void* out;            // log psi (ns)
  void* act_out;        // optional stored activations (ns x (2*C*n + H))
  const void* act_in;   // optional activations for the VJP (null: recompute)
  void* jac;            // ns x npar
  const void* r;        // VJP weights (ns)
  void* partial;        // VJP partials (ff_vjp_blocks x npar)
  void* grad;           // VJP result (npar)
  int vjp_blocks;       // VJP grid size (0: ff_vjp_blocks(ns))
  // flip_B > 0: configuration q is the term-(q % flip_B) neighbour (see
  // bond_sites, Hamiltonian kind flip_kind) of sigma row q / flip_B, so ns
  // counts (sample, term) pairs
  int flip_B, flip_kind;
  // General layer stack (layers non-null): the layer table on the device,
  // the per-sample activation strip (the input and every layer's output, in
  // order; act_out / act_in hold ns strips of act_len values), the largest
  // layer width, the parameter count and the samples per block.
  const struct FFLayerDev* layers;
  int nl, act_len, dmax, npar, sb;
};
// One layer of a general FeedForward stack, as the kernels read it: the
// yq_ff_layer fields resolved against the running width (in_dim, out_dim),
// with the offsets of its input and output in the activation strip and of
// its parameters in the parameter vector (npar = weights + bias).
struct FFLayerDev {
  int kind, act;
  int in_dim, out_dim, in_off, out_off;
  int p_off, npar;
  int ic, oc, K, stride, bias;  // Conv1D; FullyConnected uses ic = in, oc = out
  int rows, cols, axis;         // Mean
  double prm;                   // activation parameter
};
// LDS of the general kernel for sb samples a block
size_t ffs_lds_bytes(int act_len, int dmax, int sb, int use_double);
// the general kernel (FFArgs with layers set); launch_ff dispatches to it
int launch_ffs(const FFArgs& a, int mode, int use_double, hipStream_t st);
enum { kFFForward = 0, kFFJacobian = 1, kFFVjp = 2 };
int ff_vjp_blocks(int ns, int use_double);
// The fixed device array behind yq_ff_vjp_dev_noalloc (kFFScratchBytes per
// device, held until the process ends): ff_scratch_init looks up its address
// on the current device once (call it with that device current); ff_scratch
// returns the address recorded for a device, or null.
constexpr size_t kFFScratchBytes = (size_t)16 << 20;
int ff_scratch_init(int device);
void* ff_scratch(int device);
size_t ff_smem_bytes(const FFArgs& a, int use_double);
int launch_ff(const FFArgs& a, int mode, int use_double, hipStream_t st);
// a.wt from the FC weight of a.params (after every parameter upload)
void launch_ff_transpose(const FFArgs& a, int use_double, hipStream_t st);

// ------------------------------------------------------------------ AmplitudePhase
// out[s] = (lp[s] / 2, pi ff[s]): log coeff of psi = sqrt(RBM) exp(i pi FF)
void launch_ap_log_psi(const double* lp, const double* ff, int ns, double* out, hipStream_t st);
// re[s] = z[2 s], im[s] = z[2 s + 1]
void launch_split_complex(const double* z, int ns, double* re, double* im, hipStream_t st);

// ------------------------------------------------------------------ vectors
// deterministic dot product: partial sums per block; host finishes the sum
int dot_blocks(size_t n);
void launch_dot(const double* x, const double* y, size_t n, double* partial,
                hipStream_t st);
void launch_axpy(double alpha, const double* x, double* y, size_t n,
                 hipStream_t st);  // y += alpha x
void launch_xpby(const double* x, double beta, double* y, size_t n,
                 hipStream_t st);  // y = x + beta y
void launch_scale_add(double* y, double a, const double* x, double b, size_t n,
                      hipStream_t st);  // y = a*y + b*x
void launch_add_const(double* y, double c, size_t n, hipStream_t st);
// the same with the scalar on the device: *out = x . y summed in the host's
// order (no round trip), y += sign * *c, y += (sign * *a) x
void launch_dot_dev(const double* x, const double* y, size_t n, double* partial, double* out,
                    hipStream_t st);
void launch_add_const_p(double* y, const double* c, double sign, size_t n, hipStream_t st);
void launch_axpy_p(const double* a, double sign, const double* x, double* y, size_t n,
                   hipStream_t st);
void launch_to_float(const double* x, float* y, size_t n, hipStream_t st);
void launch_float_to_double(const float* x, double* y, size_t n, hipStream_t st);
void launch_transpose(const double* W, int m, int n, double* WT,
                      hipStream_t st);  // WT[j*n+i] = W[i*m+j]
void launch_transpose_f(const double* W, int m, int n, float* WT,
                        hipStream_t st);

}  // namespace yq
