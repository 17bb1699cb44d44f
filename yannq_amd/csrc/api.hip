// C-ABI implementation of include/yannq_amd.h: device state, argument
// checks, the SR solvers and the epoch body.  Reference citations are paths
// under the yannq source tree.
#include <math.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mixed.h"
#include "yq_internal.h"
#include "yq_math.h"

using namespace yq;

// ---------------------------------------------------------------- profiling
namespace {
struct ProfEntry {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> open;
  double total_ms = 0.0;
  long long launches = 0;
};
struct ProfState {
  std::vector<std::string> labels;
  std::vector<ProfEntry> entries;
  int find(const char* l) const {
    for (size_t i = 0; i < labels.size(); ++i)
      if (labels[i] == l) return (int)i;
    return -1;
  }
};
ProfState g_prof;
bool g_prof_any = false;
std::mutex g_prof_mu;  // worker threads of yq_multi_sr_step share the registry
}  // namespace

namespace yq {
bool prof_enabled(const char* label) { return g_prof_any && g_prof.find(label) >= 0; }
void prof_begin(const char* label, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  ProfEntry& e = g_prof.entries[g_prof.find(label)];
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return;
  (void)hipEventRecord(ev, st);
  e.open.push_back(ev);
}
void prof_pair(const char* label, hipEvent_t start, hipEvent_t stop) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.entries[g_prof.find(label)].pending.push_back({start, stop});
}
void prof_end(const char* label, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  ProfEntry& e = g_prof.entries[g_prof.find(label)];
  if (e.open.empty()) return;
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return;
  (void)hipEventRecord(ev, st);
  e.pending.push_back({e.open.back(), ev});
  e.open.pop_back();
}
}  // namespace yq

namespace {

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace

struct yq_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  std::string err;
  rocblas_handle blas = nullptr;
  std::vector<Buf*> bufs;

  // machine (cplx: complex double parameters, interleaved re / im)
  bool have_rbm = false;
  int n = 0, m = 0, use_bias = 1, cplx = 0;
  size_t dim = 0, nparam_alloc = 0;
  Buf params, WT64, W32, WT32, a32;
  Buf cpd, cpf;  // complex parameter planes (double, float)
  CplxPlanes planes{};

  // hamiltonian
  bool have_ham = false;
  yq_ham_desc ham{};

  // sampler
  bool have_sampler = false;
  int skind = 0, chains = 0, sweep_factor = 1, sprec = 0;
  int ntemps = 1;  // parallel tempering: chains = ntemps x chain groups
  uint64_t seed = 0, rand_epoch = 0;
  Buf sig_state, ctr, acc, tbuf, lrbuf;

  // current batch; weighted: an exact basis with |psi|^2 weights bw (row
  // scales brs = sqrt(ns bw)), see exact.hip
  bool have_batch = false, have_eloc = false, weighted = false;
  int ns = 0, bprec = 0;
  Buf bsig, btheta, ratio, eloc, estat;
  Buf bw, brs, blogpsi, bbasis, bcdf, bpart, xsig;
  double e_mean = 0.0, e_var = 0.0;

  // SR
  bool have_sr = false;
  int sr_solver = 0, sr_prec = 0;
  Buf T64, T32, O, omean, F, S, M, partial, ecent, info;
  Buf hxx;  // complex explicit build: X^T X of the rows read as reals
  Buf cft, cfs;  // complex factored build: T planes; operator scratch
  Buf vx, vr, vp, vAp, vy, vtmp, vb, dotp;
  // device scalars: [0, 1] <omean, v> and sum y of the FP64 factored operator,
  // [2, 3] <b, b> and <r, r> of the mixed refinement
  Buf dscal;
  yq::MixedWork mixed;
  // solver options: Jacobi preconditioning, warm start from the last solution
  int precond = 0, warm_start = 0;
  Buf diagS, dinv64, vz, xprev, ones;
  // fisherAvg (yq_sr_fisher_average): the running average of S and its count
  Buf savg;
  int fisher_t = 0;
  size_t fisher_len = 0;
  bool diag_ready = false;  // diagS holds diag(S) of the current build
  bool have_xprev = false;
  int ones_len = 0;

  // many machines (yq_multi_*): one sub-context per machine
  std::vector<yq_ctx*> subs;
  int multi_kind = 0, multi_prec = 0, m_nitems = 0, m_npad = 16;
  // the machines' inner CG in one set of launches per iteration (mixed SR)
  yq::MixedGroup group;
  bool grouped = true;
  Buf m_sargs, m_items, m_eargs, m_stats;

  // FeedForward machine
  bool have_ff = false;
  yq_ff_desc ff{};
  int ff_npar = 0;
  std::vector<double> ff_host;  // fp64 master parameters
  Buf ffp, ff_sig, ff_out, ff_jac, ff_part, ff_grad;
  // general layer stack (yq_ff_create_stack off the C4 pattern): the layer
  // table (host copy and device copy), the activation strip length, the
  // widest layer and the samples per block of the general kernel
  bool ff_stack = false;
  std::vector<FFLayerDev> ff_ldev;
  Buf ff_ltab;
  int ff_act_len = 0, ff_dmax = 0, ff_sb = 1;

  // AmplitudePhase machine (yq_ap_enable): the real RBM above is the
  // amplitude, the FeedForward machine the phase; ap_ffs / ap_ffb hold the
  // phase network's outputs on the batch and on its E_loc neighbours
  bool ap = false;
  Buf ap_ffs, ap_ffb;

  yq_ctx() {
    for (Buf* b : {&params, &WT64, &W32, &WT32, &a32, &sig_state, &ctr, &acc, &tbuf, &lrbuf, &bsig,
                   &btheta, &ratio, &eloc, &estat, &T64, &T32, &O, &omean, &F, &S, &M,
                   &partial, &ecent, &info, &vx, &vr, &vp, &vAp, &vy, &vtmp, &vb, &dotp, &dscal,
                   &ffp, &ff_sig, &ff_out, &ff_jac, &ff_part, &ff_grad, &ff_ltab, &diagS, &dinv64, &vz, &savg,
                   &xprev, &ones, &cpd, &cpf, &m_sargs, &m_items, &m_eargs, &m_stats, &ap_ffs,
                   &ap_ffb})
      bufs.push_back(b);
  }
};

namespace {

int fail(yq_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, x)                                                        \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess)                                                   \
      return fail(c, YQ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(c) HIPCHK(c, hipGetLastError())

#define BLASCHK(c, x)                                                        \
  do {                                                                       \
    rocblas_status s_ = (x);                                                 \
    if (s_ != rocblas_status_success)                                        \
      return fail(c, YQ_ERR_HIP, std::string(#x) + ": " + rocblas_status_to_string(s_)); \
  } while (0)

#define TRY(x)                 \
  do {                         \
    int rc_ = (x);             \
    if (rc_ != YQ_OK) return rc_; \
  } while (0)

int ensure(yq_ctx* c, Buf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.cap >= bytes) return YQ_OK;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (e != hipSuccess) return fail(c, YQ_ERR_HIP, "hipFree failed");
  }
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return fail(c, YQ_ERR_OOM,
                "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  b.cap = bytes;
  return YQ_OK;
}

int blas(yq_ctx* c) {
  if (c->blas) return YQ_OK;
  BLASCHK(c, rocblas_create_handle(&c->blas));
  BLASCHK(c, rocblas_set_stream(c->blas, c->st));
  return YQ_OK;
}

int sync(yq_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->st));
  return YQ_OK;
}

// doubles per parameter (2 for a complex machine) and vector lengths in doubles
size_t ncomp(const yq_ctx* c) { return c->cplx ? 2 : 1; }
size_t vlen(const yq_ctx* c) { return c->dim * ncomp(c); }
// doubles per local energy (complex for a complex or an AmplitudePhase machine)
size_t ecomp(const yq_ctx* c) { return (c->cplx || c->ap) ? 2 : 1; }
// length of the SR system's vectors: the AmplitudePhase machine's parameters
// are [RBM | FeedForward] (AmplitudePhase.hpp:64-77)
size_t srlen(const yq_ctx* c) { return c->ap ? c->dim + (size_t)c->ff_npar : vlen(c); }
double* W64(yq_ctx* c) { return c->params.as<double>(); }
double* a64(yq_ctx* c) { return c->params.as<double>() + ncomp(c) * c->n * c->m; }
double* b64(yq_ctx* c) { return c->params.as<double>() + ncomp(c) * ((size_t)c->n * c->m + c->n); }

// working copies derived from the fp64 master parameters
int refresh_params(yq_ctx* c) {
  const int n = c->n, m = c->m;
  if (c->cplx) {
    launch_cplx_planes(W64(c), n, m, c->planes, c->st);
    LAUNCHCHK(c);
    c->have_eloc = false;
    c->have_sr = false;
    return YQ_OK;
  }
  launch_transpose(W64(c), m, n, c->WT64.as<double>(), c->st);
  launch_to_float(W64(c), c->W32.as<float>(), (size_t)n * m, c->st);
  launch_transpose_f(W64(c), m, n, c->WT32.as<float>(), c->st);
  launch_to_float(a64(c), c->a32.as<float>(), (size_t)n, c->st);
  LAUNCHCHK(c);
  c->have_eloc = false;
  c->have_sr = false;
  return YQ_OK;
}

int need_rbm(yq_ctx* c) {
  if (!c) return YQ_ERR_ARG;
  if (!c->have_rbm) return fail(c, YQ_ERR_STATE, "no machine: call yq_rbm_create first");
  return YQ_OK;
}

// deterministic dot: device partials, host sum in block order
int dot(yq_ctx* c, const double* x, const double* y, size_t n, double* out) {
  const int nb = dot_blocks(n);
  TRY(ensure(c, c->dotp, sizeof(double) * 512));
  launch_dot(x, y, n, c->dotp.as<double>(), c->st);
  LAUNCHCHK(c);
  double h[512];
  HIPCHK(c, hipMemcpyAsync(h, c->dotp.p, sizeof(double) * nb, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += h[i];
  *out = s;
  return YQ_OK;
}

// FeedForward kernel arguments for the context's network (defined below)
FFArgs ff_args(yq_ctx* c, int ns, const int8_t* d_sigma);

// ---------------------------------------------------------------- E_loc
// Arguments of the E_loc kernels for the current batch; allocates outputs.
int eloc_args(yq_ctx* c, ElocArgs* out) {
  if (!c->have_batch) return fail(c, YQ_ERR_STATE, "no samples: call yq_sample or yq_samples_set");
  if (!c->have_ham) return fail(c, YQ_ERR_STATE, "no Hamiltonian: call yq_ham_set");
  if (c->ham.n != c->n) return fail(c, YQ_ERR_ARG, "Hamiltonian size differs from the machine's n");
  const int B = num_bonds(c->ham);
  TRY(ensure(c, c->ratio, sizeof(double) * ecomp(c) * (size_t)c->ns * B));
  TRY(ensure(c, c->eloc, sizeof(double) * ecomp(c) * (size_t)c->ns));
  TRY(ensure(c, c->estat, sizeof(double) * 2));
  ElocArgs& a = *out;
  a = ElocArgs{};
  a.sigma = c->bsig.as<int8_t>();
  a.theta = c->btheta.p;
  a.use_float = c->bprec == YQ_MIXED;
  a.WT = a.use_float ? (const void*)c->WT32.p : (const void*)c->WT64.p;
  a.ar = a.use_float ? (const void*)c->a32.p : (const void*)a64(c);
  a.ratio = c->ratio.as<double>();
  a.eloc = c->eloc.as<double>();
  a.ns = c->ns;
  a.n = c->n;
  a.m = c->m;
  a.B = B;
  a.ham = c->ham;
  if (c->cplx) {
    const CplxPlanes& P = c->planes;
    a.WT = a.use_float ? (const void*)P.WTr32 : (const void*)P.WTr64;
    a.WTi = a.use_float ? (const void*)P.WTi32 : (const void*)P.WTi64;
    a.ar = a.use_float ? (const void*)P.ar32 : (const void*)P.ar64;
    a.ai = a.use_float ? (const void*)P.ai32 : (const void*)P.ai64;
  }
  return YQ_OK;
}

int compute_eloc(yq_ctx* c) {
  ElocArgs a;
  TRY(eloc_args(c, &a));
  if (c->ap) {
    // AmplitudePhaseState.hpp:57-120: the phase network on the batch and on
    // every (sample, term) neighbour, then the RBM ratio pass folds
    // exp(logRatio / 2 + i pi dFF) into complex ratios for the functors
    if (eloc_lds_bytes(a) > 64 * 1024) return fail(c, YQ_ERR_UNSUPPORTED, "E_loc tile exceeds LDS");
    const size_t pairs = (size_t)c->ns * a.B;
    if (pairs > (size_t)INT32_MAX) return fail(c, YQ_ERR_ARG, "too many (sample, term) pairs");
    TRY(ensure(c, c->ap_ffs, sizeof(double) * c->ns));
    TRY(ensure(c, c->ap_ffb, sizeof(double) * pairs));
    FFArgs f = ff_args(c, c->ns, c->bsig.as<int8_t>());
    f.out = c->ap_ffs.p;
    TRY(launch_ff(f, kFFForward, 1, c->st));
    f.ns = (int)pairs;
    f.out = c->ap_ffb.p;
    f.flip_B = a.B;
    f.flip_kind = c->ham.kind;
    TRY(launch_ff(f, kFFForward, 1, c->st));
    a.ap_ffs = c->ap_ffs.as<double>();
    a.ap_ffb = c->ap_ffb.as<double>();
    launch_eloc_ratio(a, c->st);
    launch_eloc_combine_c(a, c->st);
    LAUNCHCHK(c);
    if (c->weighted)
      launch_eloc_stats_w(c->eloc.as<double>(), c->bw.as<double>(), c->ns, 1,
                          c->estat.as<double>(), c->st);
    else
      launch_eloc_stats_c(c->eloc.as<double>(), c->ns, c->estat.as<double>(), c->st);
  } else if (c->cplx) {
    if (eloc_c_lds_bytes(a) > 64 * 1024)
      return fail(c, YQ_ERR_UNSUPPORTED, "E_loc tile exceeds LDS");
    launch_eloc_c(a, c->st);
    LAUNCHCHK(c);
    if (c->weighted)
      launch_eloc_stats_w(c->eloc.as<double>(), c->bw.as<double>(), c->ns, 1,
                          c->estat.as<double>(), c->st);
    else
      launch_eloc_stats_c(c->eloc.as<double>(), c->ns, c->estat.as<double>(), c->st);
  } else {
    if (eloc_lds_bytes(a) > 64 * 1024) return fail(c, YQ_ERR_UNSUPPORTED, "E_loc tile exceeds LDS");
    launch_eloc(a, c->st);
    LAUNCHCHK(c);
    if (c->weighted)
      launch_eloc_stats_w(c->eloc.as<double>(), c->bw.as<double>(), c->ns, 0,
                          c->estat.as<double>(), c->st);
    else
      launch_eloc_stats(c->eloc.as<double>(), c->ns, c->estat.as<double>(), c->st);
  }
  LAUNCHCHK(c);
  double h[2];
  HIPCHK(c, hipMemcpyAsync(h, c->estat.p, sizeof(h), hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  c->e_mean = h[0];
  c->e_var = h[1];
  c->have_eloc = true;
  return YQ_OK;
}

// ---------------------------------------------------------------- SR build
size_t aug_partial_doubles(int ns, int n, int m) { return factored_partial_doubles(ns, n, m); }

// Complex S = alpha A A^H (lower triangle, interleaved) with A = conj(O_c)
// (the stored rows, ns x dim complex row-major) on the hand-written SYRKs:
// the rows read as 2 dim reals X give XX = alpha X^T X (FP64 MFMA, or split-
// bf16 with fp32 accumulation for the mixed build), from whose 2 x 2 blocks
// k_herk_combine forms Re S and Im S.  A batch with few tiles and many samples
// is split into sample chunks summed in chunk order (partials in c->M).
int herk_mfma(yq_ctx* c, const double* A, size_t dim, int ns, double alpha, double* S,
              bool bf16x3) {
  const size_t d2 = 2 * dim;
  TRY(ensure(c, c->hxx, sizeof(double) * d2 * d2));
  const size_t cap = ((size_t)2 << 30) / sizeof(double);
  const int P = syrk64_chunks(ns, d2, cap);
  double* part = nullptr;
  if (P > 1) {
    TRY(ensure(c, c->M, sizeof(double) * (size_t)P * d2 * d2));
    part = c->M.as<double>();
  }
  const int rc = bf16x3 ? launch_syrk_bf16x3_dense(A, ns, d2, alpha, c->hxx.as<double>(), P,
                                                   part, c->st)
                        : launch_syrk64_dense(A, ns, d2, alpha, c->hxx.as<double>(), P, part,
                                              c->st);
  if (rc) return fail(c, rc, "SYRK arguments");
  launch_herk_combine(c->hxx.as<double>(), dim, S, c->st);
  LAUNCHCHK(c);
  return YQ_OK;
}

// S = alpha O_c^T O_c (lower triangle) of a real machine on the hand-written
// FP64 MFMA SYRK (syrk.hip): from the factors (spins, T64, column mean and
// the exact batch's row scales) when Odense is null, else from the stored
// row-major centred O.  A batch with few output tiles and many samples (an
// exact basis) is split into sample chunks whose partial products go to
// c->M and are summed in chunk order.
int syrk_real(yq_ctx* c, const double* Odense, size_t dim, int ns, double alpha, double* S) {
  const size_t cap = ((size_t)2 << 30) / sizeof(double);  // partial products, doubles
  const int P = syrk64_chunks(ns, dim, cap);
  double* part = nullptr;
  if (P > 1) {
    TRY(ensure(c, c->M, sizeof(double) * (size_t)P * dim * dim));
    part = c->M.as<double>();
  }
  int rc;
  if (Odense)
    rc = launch_syrk64_dense(Odense, ns, dim, alpha, S, P, part, c->st);
  else
    rc = launch_syrk64_factored(c->bsig.as<int8_t>(), c->T64.as<double>(),
                                c->omean.as<double>(), c->weighted ? c->brs.as<double>() : nullptr,
                                ns, c->n, c->m, c->use_bias, alpha, S, P, part, c->st);
  if (rc) return fail(c, rc, "SYRK arguments");
  LAUNCHCHK(c);
  return YQ_OK;
}

// Complex machine: dense conj(O) (see complex.hip); S only for the explicit
// solver, in FP64, or with YQ_MIXED in split-bf16 (solved by FP64-refined
// Cholesky); the dense CG runs in FP64 whatever the precision.
// Scratch of the complex factored operator (doubles from c->cfs): the planes
// of v (2 dim), the four forward products (4 ns), y as planes and interleaved
// (4 ns), the four backward products (4 dim) and three complex scalars.
struct CfScratch {
  double *vr, *vi, *P[4], *yr, *yi, *yz, *B[4], *ov, *sy, *zero;
};
CfScratch cf_scratch(yq_ctx* c) {
  const size_t ns = (size_t)c->ns, dim = c->dim;
  double* q = c->cfs.as<double>();
  CfScratch w;
  w.vr = q;
  w.vi = q + dim;
  q += 2 * dim;
  for (int k = 0; k < 4; ++k) w.P[k] = q + k * ns;
  q += 4 * ns;
  w.yr = q;
  w.yi = q + ns;
  w.yz = q + 2 * ns;
  q += 4 * ns;
  for (int k = 0; k < 4; ++k) w.B[k] = q + k * dim;
  q += 4 * dim;
  w.ov = q;
  w.sy = q + 2;
  w.zero = q + 4;
  return w;
}
size_t cf_scratch_doubles(int ns, size_t dim) { return 6 * dim + 8 * (size_t)ns + 6; }

// The four real backward products of a complex y (planes yr, yi) on the
// factored T planes: B0 = O_r^T yr, B1 = O_i^T yi, B2 = O_r^T yi, B3 = O_i^T yr
// with O_r / O_i the factored rows of Re / Im T
int cf_backward(yq_ctx* c, const double* yr, const double* yi, double* const (&B)[4]) {
  const int ns = c->ns, n = c->n, m = c->m;
  const double* Tr = c->cft.as<double>();
  const double* Ti = Tr + (size_t)ns * m;
  const int8_t* sg = c->bsig.as<int8_t>();
  double* part = c->partial.as<double>();
  launch_factored_bwd64(sg, Tr, ns, n, m, c->use_bias, yr, B[0], part, c->st, 1.0);
  launch_factored_bwd64(sg, Ti, ns, n, m, c->use_bias, yi, B[1], part, c->st, 0.0);
  launch_factored_bwd64(sg, Tr, ns, n, m, c->use_bias, yi, B[2], part, c->st, 1.0);
  launch_factored_bwd64(sg, Ti, ns, n, m, c->use_bias, yr, B[3], part, c->st, 0.0);
  LAUNCHCHK(c);
  return YQ_OK;
}

// Factored complex build (FP64): O_s = sigma'_s (x) T'_s with complex T' =
// [tanh theta | 1] is never stored; every product is four real factored
// products on the planes of T (sr.hip's FP64 MFMA kernels).  The device keeps
// conj(omean) and F as the dense build does.
int sr_build_cf(yq_ctx* c) {
  const int ns = c->ns, n = c->n, m = c->m;
  const size_t dim = c->dim;
  TRY(ensure(c, c->cft, sizeof(double) * 2 * (size_t)ns * m));
  TRY(ensure(c, c->cfs, sizeof(double) * cf_scratch_doubles(ns, dim)));
  TRY(ensure(c, c->partial, sizeof(double) * aug_partial_doubles(ns, n, m)));
  double* Tr = c->cft.as<double>();
  double* Ti = Tr + (size_t)ns * m;
  launch_split_c(c->T64.as<double>(), (size_t)ns * m, Tr, Ti, c->st);
  const CfScratch w = cf_scratch(c);
  // omean = (mean of the Re T rows, mean of the Im T rows); kept conjugated
  launch_factored_mean64(c->bsig.as<int8_t>(), Tr, ns, n, m, c->use_bias, w.B[0],
                         c->partial.as<double>(), c->st);
  launch_factored_mean64(c->bsig.as<int8_t>(), Ti, ns, n, m, c->use_bias, w.B[1],
                         c->partial.as<double>(), c->st, 0.0);
  launch_conj_merge(w.B[0], w.B[1], dim, c->omean.as<double>(), c->st);
  LAUNCHCHK(c);
  // F = conj(O)^T (E - mean E) / ns = O_c^H E / ns (SRMat.hpp:89-103), the
  // energies centred on the device (their sum held in w.sy, which every
  // apply overwrites)
  launch_cdot1(c->eloc.as<double>(), nullptr, (size_t)ns, w.sy, c->st);
  launch_center_c(c->eloc.as<double>(), w.sy, (size_t)ns, w.yr, w.yi, c->st);
  LAUNCHCHK(c);
  TRY(cf_backward(c, w.yr, w.yi, w.B));
  HIPCHK(c, hipMemsetAsync(w.zero, 0, sizeof(double) * 2, c->st));
  // (B0 + B1, B2 - B3): conj(O)^T of the centred energies
  launch_cbwd_combine(w.B[0], w.B[1], w.B[2], w.B[3], c->omean.as<double>(), w.zero,
                      c->omean.as<double>(), 0.0, dim, c->F.as<double>(), c->st);
  LAUNCHCHK(c);
  TRY(sync(c));
  c->have_sr = true;
  c->diag_ready = false;  // the Jacobi option is not offered for this build
  c->sr_solver = YQ_SOLVE_CG_FACTORED;
  c->sr_prec = YQ_FP64;
  return YQ_OK;
}

int sr_build_c(yq_ctx* c, int solver, int precision) {
  const int ns = c->ns, n = c->n, m = c->m;
  const size_t dim = c->dim;
  if (solver != YQ_SOLVE_EXPLICIT && solver != YQ_SOLVE_CG && solver != YQ_SOLVE_CG_FACTORED)
    return fail(c, YQ_ERR_ARG, "unknown solver");
  if (solver == YQ_SOLVE_CG_FACTORED && c->weighted)
    return fail(c, YQ_ERR_UNSUPPORTED,
                "an exact weighted batch uses the explicit or the dense CG solver");
  TRY(ensure(c, c->omean, sizeof(double) * 2 * dim));
  TRY(ensure(c, c->F, sizeof(double) * 2 * dim));
  TRY(ensure(c, c->T64, sizeof(double) * 2 * (size_t)ns * m));
  launch_tanh_c(c->btheta.p, c->bprec == YQ_MIXED, c->T64.as<double>(), (size_t)ns * m, c->st);
  if (solver == YQ_SOLVE_CG_FACTORED) return sr_build_cf(c);  // O is not stored
  TRY(ensure(c, c->O, sizeof(double) * 2 * (size_t)ns * dim));
  const int chunks = colsum_chunks(ns);
  TRY(ensure(c, c->partial, sizeof(double) * 2 * (size_t)chunks * dim));
  double* Oh = c->O.as<double>();
  launch_build_Oh(c->bsig.as<int8_t>(), c->T64.as<double>(), ns, n, m, c->use_bias, Oh, c->st);
  // column means of conj(O) (complex columns as pairs of doubles); an exact
  // weighted batch takes the |psi|^2 weights and scales the centred rows by
  // sqrt(ns w_s), so that every product below (/ ns) is the weighted sum
  const double* w = c->weighted ? c->bw.as<double>() : nullptr;
  const double* rs = c->weighted ? c->brs.as<double>() : nullptr;
  launch_colsum(Oh, ns, (int)(2 * dim), c->partial.as<double>(), c->st, w);
  launch_reduce_partials(c->partial.as<double>(), chunks, (int)(2 * dim), w ? 1.0 : 1.0 / ns,
                         c->omean.as<double>(), c->st);
  // centre and F = conj(O_c)^T E / ns = O_c^H E / ns
  launch_center_force_c(Oh, c->omean.as<double>(), c->eloc.as<double>(), ns, (int)dim,
                        c->partial.as<double>(), c->st, rs);
  launch_reduce_partials(c->partial.as<double>(), chunks, (int)(2 * dim), 1.0 / ns,
                         c->F.as<double>(), c->st);
  LAUNCHCHK(c);
  if (solver == YQ_SOLVE_EXPLICIT) {
    TRY(ensure(c, c->S, sizeof(double) * 2 * dim * dim));
    TRY(herk_mfma(c, Oh, dim, ns, 1.0 / ns, c->S.as<double>(), precision == YQ_MIXED));
  }
  // diag(S) = mean |O_c|^2, duplicated per (re, im) for the Jacobi option
  TRY(ensure(c, c->diagS, sizeof(double) * 2 * dim));
  TRY(ensure(c, c->vtmp, sizeof(double) * dim));
  launch_absq_colsum(Oh, ns, (int)dim, c->partial.as<double>(), c->st);
  launch_reduce_partials(c->partial.as<double>(), chunks, (int)dim, 1.0 / ns,
                         c->vtmp.as<double>(), c->st);
  launch_dup2(c->vtmp.as<double>(), c->diagS.as<double>(), dim, c->st);
  LAUNCHCHK(c);
  TRY(sync(c));
  c->have_sr = true;
  c->diag_ready = true;  // formed above
  c->sr_solver = solver;
  c->sr_prec = solver == YQ_SOLVE_EXPLICIT ? precision : YQ_FP64;
  return YQ_OK;
}

// The SR system of a real machine as independent diagonal blocks: one block
// for an RBM; for an AmplitudePhase machine the amplitude and the phase
// parameters, whose metric Re S has no cross block (NGDExact.hpp:177-189:
// with O = [O_amp | i O_phase] the cross terms of <O^* O^T> are imaginary).
// Block b: parameters [off, off + d), centred rows O (ns x d row-major, i.e.
// d x ns column-major) and the lower triangle of its S (d x d) at S.
struct SrBlock {
  size_t off, d;
  double* O;
  double* S;
};
std::vector<SrBlock> sr_blocks(yq_ctx* c) {
  if (!c->ap) return {SrBlock{0, c->dim, c->O.as<double>(), c->S.as<double>()}};
  const size_t dA = c->dim, dP = (size_t)c->ff_npar;
  return {SrBlock{0, dA, c->O.as<double>(), c->S.as<double>()},
          SrBlock{dA, dP, c->O.as<double>() + (size_t)c->ns * dA,
                  c->S.p ? c->S.as<double>() + dA * dA : nullptr}};
}

// AmplitudePhase machine, FP64: rows [O_RBM / 2 | pi dFF] (AmplitudePhase.hpp:
// 42-58; the phase block multiplies i), F = Re(<O^* E_loc> - <O^*><E_loc>) =
// [cov(O_amp, Re E_loc) | cov(O_phase, Im E_loc)], S = Re(S) block-diagonal
// (the blocks of NGDExact::corrMatAmp / corrMatPhase, NGDExact.hpp:177-189).
int sr_build_ap(yq_ctx* c, int solver, int precision) {
  if (solver != YQ_SOLVE_EXPLICIT && solver != YQ_SOLVE_CG)
    return fail(c, YQ_ERR_UNSUPPORTED,
                "AmplitudePhase machines use the explicit or the dense CG solver");
  if (precision != YQ_FP64)
    return fail(c, YQ_ERR_UNSUPPORTED, "AmplitudePhase SR runs in FP64");
  const int ns = c->ns, n = c->n, m = c->m;
  const size_t dA = c->dim, dP = (size_t)c->ff_npar, D = dA + dP;
  TRY(ensure(c, c->omean, sizeof(double) * D));
  TRY(ensure(c, c->F, sizeof(double) * D));
  TRY(ensure(c, c->diagS, sizeof(double) * D));
  TRY(ensure(c, c->T64, sizeof(double) * (size_t)ns * m));
  TRY(ensure(c, c->O, sizeof(double) * (size_t)ns * D));
  TRY(ensure(c, c->ecent, sizeof(double) * 2 * (size_t)ns));
  TRY(ensure(c, c->ff_out, sizeof(double) * (size_t)ns));
  const int chunks = colsum_chunks(ns);
  TRY(ensure(c, c->partial, sizeof(double) * (size_t)chunks * std::max(dA, dP)));
  if (solver == YQ_SOLVE_EXPLICIT) TRY(ensure(c, c->S, sizeof(double) * (dA * dA + dP * dP)));
  if (c->bprec == YQ_MIXED) {
    TRY(ensure(c, c->vtmp, sizeof(double) * (size_t)ns * m));
    launch_float_to_double(c->btheta.as<float>(), c->vtmp.as<double>(), (size_t)ns * m, c->st);
    launch_tanh(c->vtmp.p, c->T64.p, (size_t)ns * m, 0, c->st);
  } else {
    launch_tanh(c->btheta.p, c->T64.p, (size_t)ns * m, 0, c->st);
  }
  const std::vector<SrBlock> blk = sr_blocks(c);
  // logDerivAmp = RBM logDeriv / 2, logDerivPhase = pi FF backward
  launch_build_O(c->bsig.as<int8_t>(), c->T64.as<double>(), ns, n, m, c->use_bias, blk[0].O,
                 c->st);
  launch_scale_add(blk[0].O, 0.5, blk[0].O, 0.0, (size_t)ns * dA, c->st);
  FFArgs f = ff_args(c, ns, c->bsig.as<int8_t>());
  f.jac = blk[1].O;
  f.out = c->ff_out.p;
  TRY(launch_ff(f, kFFJacobian, 1, c->st));
  launch_scale_add(blk[1].O, M_PI, blk[1].O, 0.0, (size_t)ns * dP, c->st);
  double* er = c->ecent.as<double>();
  launch_split_complex(c->eloc.as<double>(), ns, er, er + ns, c->st);
  LAUNCHCHK(c);
  const double* w = c->weighted ? c->bw.as<double>() : nullptr;
  const double* rs = c->weighted ? c->brs.as<double>() : nullptr;
  for (size_t b = 0; b < blk.size(); ++b) {
    const int d = (int)blk[b].d;
    double* part = c->partial.as<double>();
    launch_colsum(blk[b].O, ns, d, part, c->st, w);
    launch_reduce_partials(part, chunks, d, w ? 1.0 : 1.0 / ns, c->omean.as<double>() + blk[b].off,
                           c->st);
    launch_center_force(blk[b].O, c->omean.as<double>() + blk[b].off, er + b * ns, ns, d, part,
                        c->st, rs);
    launch_reduce_partials(part, chunks, d, 1.0 / ns, c->F.as<double>() + blk[b].off, c->st);
    // diag(S) = sum of the (row-scaled) centred squares / ns
    launch_colsum(blk[b].O, ns, d, part, c->st, nullptr, 1);
    launch_reduce_partials(part, chunks, d, 1.0 / ns, c->diagS.as<double>() + blk[b].off, c->st);
    LAUNCHCHK(c);
    if (solver == YQ_SOLVE_EXPLICIT)
      TRY(syrk_real(c, blk[b].O, blk[b].d, ns, 1.0 / ns, blk[b].S));
  }
  TRY(sync(c));
  c->have_sr = true;
  c->diag_ready = true;  // formed above
  c->sr_solver = solver;
  c->sr_prec = YQ_FP64;
  return YQ_OK;
}

int sr_build(yq_ctx* c, int solver, int precision) {
  if (!c->have_eloc) TRY(compute_eloc(c));
  if (precision != YQ_FP64 && precision != YQ_MIXED)
    return fail(c, YQ_ERR_ARG, "unknown precision");
  if (c->ap) return sr_build_ap(c, solver, precision);
  if (c->cplx) return sr_build_c(c, solver, precision);
  const int ns = c->ns, n = c->n, m = c->m;
  const size_t dim = c->dim;
  if (solver != YQ_SOLVE_EXPLICIT && solver != YQ_SOLVE_CG &&
      solver != YQ_SOLVE_CG_FACTORED)
    return fail(c, YQ_ERR_ARG, "unknown solver");
  if (precision != YQ_FP64 && precision != YQ_MIXED)
    return fail(c, YQ_ERR_ARG, "unknown precision");
  if (precision == YQ_MIXED && solver == YQ_SOLVE_CG)
    return fail(c, YQ_ERR_UNSUPPORTED,
                "mixed precision is implemented for the factored CG and explicit solvers");
  if (precision == YQ_MIXED && solver == YQ_SOLVE_EXPLICIT && c->weighted)
    return fail(c, YQ_ERR_UNSUPPORTED, "an exact weighted batch builds S in FP64");
  if (c->weighted && solver == YQ_SOLVE_CG_FACTORED)
    return fail(c, YQ_ERR_UNSUPPORTED,
                "an exact weighted batch uses the explicit or the dense CG solver");
  TRY(ensure(c, c->omean, sizeof(double) * dim));
  TRY(ensure(c, c->F, sizeof(double) * dim));
  TRY(ensure(c, c->T64, sizeof(double) * (size_t)ns * m));
  // T64 = tanh(theta) in double from the batch's theta (float batches promote)
  if (c->bprec == YQ_MIXED) {
    TRY(ensure(c, c->vtmp, sizeof(double) * (size_t)ns * m));
    launch_float_to_double(c->btheta.as<float>(), c->vtmp.as<double>(), (size_t)ns * m, c->st);
    launch_tanh(c->vtmp.p, c->T64.p, (size_t)ns * m, 0, c->st);
  } else {
    launch_tanh(c->btheta.p, c->T64.p, (size_t)ns * m, 0, c->st);
  }
  LAUNCHCHK(c);
  if (solver == YQ_SOLVE_CG_FACTORED) {
    TRY(ensure(c, c->partial, sizeof(double) * aug_partial_doubles(ns, n, m)));
    TRY(ensure(c, c->ecent, sizeof(double) * ns));
    launch_factored_mean64(c->bsig.as<int8_t>(), c->T64.as<double>(), ns, n, m, c->use_bias,
                           c->omean.as<double>(), c->partial.as<double>(), c->st);
    // F = O_c^T E / Ns = O^T (E - <E>) / Ns
    HIPCHK(c, hipMemcpyAsync(c->ecent.p, c->eloc.p, sizeof(double) * ns,
                             hipMemcpyDeviceToDevice, c->st));
    launch_add_const(c->ecent.as<double>(), -c->e_mean, ns, c->st);
    launch_scale_add(c->ecent.as<double>(), 1.0 / ns, c->ecent.as<double>(), 0.0, ns, c->st);
    launch_factored_bwd64(c->bsig.as<int8_t>(), c->T64.as<double>(), ns, n, m, c->use_bias,
                          c->ecent.as<double>(), c->F.as<double>(), c->partial.as<double>(),
                          c->st);
    LAUNCHCHK(c);
    if (precision == YQ_MIXED) {
      TRY(ensure(c, c->T32, sizeof(float) * (size_t)ns * m));
      launch_to_float(c->T64.as<double>(), c->T32.as<float>(), (size_t)ns * m, c->st);
      LAUNCHCHK(c);
      int rc = mixed_prepare(c->mixed, c->bsig.as<int8_t>(), c->T32.as<float>(), ns, n, m,
                             c->use_bias, c->st, c->err);
      if (rc) return rc;
    }
  } else {
    const size_t obytes = sizeof(double) * (size_t)ns * dim;
    TRY(ensure(c, c->O, obytes));
    const int chunks = colsum_chunks(ns);
    TRY(ensure(c, c->partial, sizeof(double) * std::max((size_t)chunks * dim,
                                                       aug_partial_doubles(ns, n, m))));
    launch_build_O(c->bsig.as<int8_t>(), c->T64.as<double>(), ns, n, m, c->use_bias,
                   c->O.as<double>(), c->st);
    // an exact weighted batch: weighted mean, rows of O_c scaled by sqrt(ns w_s)
    const double* w = c->weighted ? c->bw.as<double>() : nullptr;
    const double* rs = c->weighted ? c->brs.as<double>() : nullptr;
    launch_colsum(c->O.as<double>(), ns, (int)dim, c->partial.as<double>(), c->st, w);
    launch_reduce_partials(c->partial.as<double>(), chunks, (int)dim, w ? 1.0 : 1.0 / ns,
                           c->omean.as<double>(), c->st);
    launch_center_force(c->O.as<double>(), c->omean.as<double>(), c->eloc.as<double>(), ns,
                        (int)dim, c->partial.as<double>(), c->st, rs);
    launch_reduce_partials(c->partial.as<double>(), chunks, (int)dim, 1.0 / ns,
                           c->F.as<double>(), c->st);
    LAUNCHCHK(c);
    if (solver == YQ_SOLVE_EXPLICIT) {
      // S = O_c^T O_c / ns formed from the factors (O_c is not read): FP64
      // MFMA, or split-bf16 MFMA for the mixed explicit solve
      TRY(ensure(c, c->S, sizeof(double) * dim * dim));
      if (precision == YQ_MIXED) {
        const size_t cap = ((size_t)2 << 30) / sizeof(double);
        const int P = syrk64_chunks(ns, dim, cap);
        double* part = nullptr;
        if (P > 1) {
          TRY(ensure(c, c->M, sizeof(double) * (size_t)P * dim * dim));
          part = c->M.as<double>();
        }
        const int rc = launch_syrk_bf16x3(c->bsig.as<int8_t>(), c->T64.as<double>(),
                                          c->omean.as<double>(), nullptr, ns, n, m, c->use_bias,
                                          1.0 / ns, c->S.as<double>(), P, part, c->st);
        if (rc) return fail(c, rc, "split-bf16 SYRK arguments");
        LAUNCHCHK(c);
      } else {
        TRY(syrk_real(c, nullptr, dim, ns, 1.0 / ns, c->S.as<double>()));
      }
    }
  }
  c->diag_ready = false;  // diag(S) is formed when a preconditioned solve needs it
  TRY(sync(c));
  c->have_sr = true;
  c->sr_solver = solver;
  c->sr_prec = precision;
  return YQ_OK;
}

// diag(S) of a real machine's build for the Jacobi preconditioner, formed on
// first use (one pass over T; for a weighted batch the column sums of the
// scaled O_c^2); the complex and AmplitudePhase builds form it themselves
int ensure_diag(yq_ctx* c) {
  if (c->diag_ready) return YQ_OK;
  if (c->cplx)  // the dense complex builds form it; the factored one does not
    return fail(c, YQ_ERR_UNSUPPORTED,
                "the Jacobi option is not offered for the factored build of a complex machine");
  const int ns = c->ns, n = c->n, m = c->m;
  const size_t dim = c->dim;
  TRY(ensure(c, c->diagS, sizeof(double) * dim));
  TRY(ensure(c, c->vtmp, sizeof(double) * std::max(dim, (size_t)(colsum_chunks(ns) + 1) * m)));
  if (c->weighted) {
    launch_colsum(c->O.as<double>(), ns, (int)dim, c->partial.as<double>(), c->st, nullptr, 1);
    launch_reduce_partials(c->partial.as<double>(), colsum_chunks(ns), (int)dim, 1.0 / ns,
                           c->diagS.as<double>(), c->st);
  } else {
    launch_diag_S(c->T64.as<double>(), ns, n, m, c->use_bias, c->omean.as<double>(),
                  c->diagS.as<double>(), c->vtmp.as<double>(), c->st);
  }
  LAUNCHCHK(c);
  c->diag_ready = true;
  return YQ_OK;
}

// out = (S + shift) v on device pointers, FP64
int apply64(yq_ctx* c, double shift, const double* v, double* out) {
  const int ns = c->ns;
  const size_t dim = c->dim;
  TRY(ensure(c, c->vy, sizeof(double) * ns));
  double* y = c->vy.as<double>();
  if (c->ap) {
    // block by block: y = A_b^T v_b / ns, out_b = A_b y + shift v_b
    TRY(blas(c));
    const double one = 1.0, zero = 0.0, inv = 1.0 / ns;
    for (const SrBlock& b : sr_blocks(c)) {
      const rocblas_int d = (rocblas_int)b.d;
      BLASCHK(c, rocblas_dgemv(c->blas, rocblas_operation_transpose, d, ns, &inv, b.O, d,
                               v + b.off, 1, &zero, y, 1));
      HIPCHK(c, hipMemcpyAsync(out + b.off, v + b.off, sizeof(double) * b.d,
                               hipMemcpyDeviceToDevice, c->st));
      BLASCHK(c, rocblas_dgemv(c->blas, rocblas_operation_none, d, ns, &one, b.O, d, y, 1,
                               &shift, out + b.off, 1));
    }
  } else if (c->cplx && c->sr_solver == YQ_SOLVE_CG_FACTORED) {
    // y = O_c v / ns = (O v - (omean . v)) / ns, out = conj(O)^T y -
    // conj(omean) sum(y) + shift v, each complex product four real factored
    // ones on the planes of T; the device keeps conj(omean)
    const int n = c->n, m = c->m;
    const double* Tr = c->cft.as<double>();
    const double* Ti = Tr + (size_t)ns * m;
    const int8_t* sg = c->bsig.as<int8_t>();
    double* part = c->partial.as<double>();
    const CfScratch w = cf_scratch(c);
    launch_split_c(v, dim, w.vr, w.vi, c->st);
    launch_factored_fwd64(sg, Tr, ns, n, m, c->use_bias, w.vr, w.P[0], part, c->st, 1.0);
    launch_factored_fwd64(sg, Ti, ns, n, m, c->use_bias, w.vi, w.P[1], part, c->st, 0.0);
    launch_factored_fwd64(sg, Tr, ns, n, m, c->use_bias, w.vi, w.P[2], part, c->st, 1.0);
    launch_factored_fwd64(sg, Ti, ns, n, m, c->use_bias, w.vr, w.P[3], part, c->st, 0.0);
    launch_cdot1(c->omean.as<double>(), v, dim, w.ov, c->st);  // omean . v
    launch_cfwd_combine(w.P[0], w.P[1], w.P[2], w.P[3], w.ov, ns, 1.0 / ns, w.yr, w.yi, w.yz,
                        c->st);
    launch_cdot1(w.yz, nullptr, ns, w.sy, c->st);  // sum y
    LAUNCHCHK(c);
    TRY(cf_backward(c, w.yr, w.yi, w.B));
    launch_cbwd_combine(w.B[0], w.B[1], w.B[2], w.B[3], c->omean.as<double>(), w.sy, v, shift,
                        dim, out, c->st);
    LAUNCHCHK(c);
  } else if (c->cplx) {
    // y = A^H v / ns = O_c v / ns, out = A y + shift v = O_c^H y + shift v
    TRY(ensure(c, c->vy, sizeof(double) * 2 * ns));
    TRY(blas(c));
    const rocblas_double_complex inv(1.0 / ns, 0.0), one(1.0, 0.0), zero(0.0, 0.0),
        sh(shift, 0.0);
    const auto* A = reinterpret_cast<const rocblas_double_complex*>(c->O.p);
    auto* yc = reinterpret_cast<rocblas_double_complex*>(c->vy.p);
    BLASCHK(c, rocblas_zgemv(c->blas, rocblas_operation_conjugate_transpose, (rocblas_int)dim,
                             ns, &inv, A, (rocblas_int)dim,
                             reinterpret_cast<const rocblas_double_complex*>(v), 1, &zero, yc,
                             1));
    HIPCHK(c, hipMemcpyAsync(out, v, sizeof(double) * 2 * dim, hipMemcpyDeviceToDevice, c->st));
    BLASCHK(c, rocblas_zgemv(c->blas, rocblas_operation_none, (rocblas_int)dim, ns, &one, A,
                             (rocblas_int)dim, yc, 1, &sh,
                             reinterpret_cast<rocblas_double_complex*>(out), 1));
  } else if (c->sr_solver == YQ_SOLVE_CG_FACTORED) {
    // the scalars <omean, v> and sum y stay on the device (summed in the order
    // of a host loop over the dot partials), so the operator needs no round trip
    TRY(ensure(c, c->dotp, sizeof(double) * 512));
    TRY(ensure(c, c->dscal, sizeof(double) * 4));
    double* ov = c->dscal.as<double>();
    double* sy = ov + 1;
    launch_factored_fwd64(c->bsig.as<int8_t>(), c->T64.as<double>(), ns, c->n, c->m,
                          c->use_bias, v, y, c->partial.as<double>(), c->st);
    launch_dot_dev(c->omean.as<double>(), v, dim, c->dotp.as<double>(), ov, c->st);
    LAUNCHCHK(c);
    launch_add_const_p(y, ov, -1.0, ns, c->st);           // O_c v
    launch_scale_add(y, 1.0 / ns, y, 0.0, ns, c->st);     // / Ns
    launch_factored_bwd64(c->bsig.as<int8_t>(), c->T64.as<double>(), ns, c->n, c->m,
                          c->use_bias, y, out, c->partial.as<double>(), c->st);
    LAUNCHCHK(c);
    // sum of y as a dot with a ones vector (its own buffer: vb may hold the
    // caller's right-hand side)
    if (c->ones_len < ns) {
      TRY(ensure(c, c->ones, sizeof(double) * ns));
      std::vector<double> h(ns, 1.0);
      HIPCHK(c, hipMemcpyAsync(c->ones.p, h.data(), sizeof(double) * ns,
                               hipMemcpyHostToDevice, c->st));
      TRY(sync(c));
      c->ones_len = ns;
    }
    launch_dot_dev(y, c->ones.as<double>(), ns, c->dotp.as<double>(), sy, c->st);
    launch_axpy_p(sy, -1.0, c->omean.as<double>(), out, dim, c->st);
    launch_axpy(shift, v, out, dim, c->st);
    LAUNCHCHK(c);
  } else if (c->sr_solver == YQ_SOLVE_CG || c->sr_solver == YQ_SOLVE_EXPLICIT) {
    TRY(blas(c));
    const double one = 1.0, zero = 0.0, inv = 1.0 / ns;
    const double* A = c->O.as<double>();  // dim x ns column-major
    BLASCHK(c, rocblas_dgemv(c->blas, rocblas_operation_transpose, (rocblas_int)dim, ns,
                             &inv, A, (rocblas_int)dim, v, 1, &zero, y, 1));
    HIPCHK(c, hipMemcpyAsync(out, v, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
    BLASCHK(c, rocblas_dgemv(c->blas, rocblas_operation_none, (rocblas_int)dim, ns, &one, A,
                             (rocblas_int)dim, y, 1, &shift, out, 1));
  }
  return YQ_OK;
}

// Conjugate gradients (SRMat.hpp:197-223 through Eigen's ConjugateGradient):
// stop when |r| < tol |b|.  x starts at 0, or at x on entry when x0 is set
// (warm start); with c->precond the Jacobi preconditioner 1/(diag S + shift)
// is applied (Eigen's IdentityPreconditioner when 0).
int cg64(yq_ctx* c, double shift, const double* b, double* x, double tol, int max_iter,
         int* iters, double* err, bool x0) {
  // complex vectors run as 2*dim doubles: every dot here is Re(a^H b), which
  // is what CG on a Hermitian operator needs (p^H A p and r^H z are real)
  const size_t dim = srlen(c);
  TRY(ensure(c, c->vr, sizeof(double) * dim));
  TRY(ensure(c, c->vp, sizeof(double) * dim));
  TRY(ensure(c, c->vAp, sizeof(double) * dim));
  TRY(ensure(c, c->vz, sizeof(double) * dim));
  TRY(ensure(c, c->dinv64, sizeof(double) * dim));
  double *r = c->vr.as<double>(), *p = c->vp.as<double>(), *Ap = c->vAp.as<double>();
  double* z = c->vz.as<double>();
  const double* dinv = nullptr;
  if (c->precond) {
    TRY(ensure_diag(c));
    launch_jacobi_inv64(c->diagS.as<double>(), shift, c->dinv64.as<double>(), dim, c->st);
    dinv = c->dinv64.as<double>();
  }
  double bn2;
  TRY(dot(c, b, b, dim, &bn2));
  *iters = 0;
  *err = 0.0;
  if (bn2 == 0.0) {
    HIPCHK(c, hipMemsetAsync(x, 0, sizeof(double) * dim, c->st));
    return YQ_OK;
  }
  if (x0) {  // r = b - (S + shift) x
    TRY(apply64(c, shift, x, Ap));
    HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
    launch_axpy(-1.0, Ap, r, dim, c->st);
  } else {
    HIPCHK(c, hipMemsetAsync(x, 0, sizeof(double) * dim, c->st));
    HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
  }
  if (dinv) launch_mul(dinv, r, z, dim, c->st);
  else HIPCHK(c, hipMemcpyAsync(z, r, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(p, z, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
  LAUNCHCHK(c);
  const double thr = tol * tol * bn2;
  double rn2, rho;
  TRY(dot(c, r, r, dim, &rn2));
  TRY(dot(c, r, z, dim, &rho));
  int it = 0;
  while (rn2 >= thr && it < max_iter) {
    TRY(apply64(c, shift, p, Ap));
    double pAp;
    TRY(dot(c, p, Ap, dim, &pAp));
    const double alpha = rho / pAp;
    launch_axpy(alpha, p, x, dim, c->st);
    launch_axpy(-alpha, Ap, r, dim, c->st);
    LAUNCHCHK(c);
    TRY(dot(c, r, r, dim, &rn2));
    ++it;
    if (rn2 < thr) break;
    double rho_new = rn2;
    if (dinv) {
      launch_mul(dinv, r, z, dim, c->st);
      TRY(dot(c, r, z, dim, &rho_new));
      launch_xpby(z, rho_new / rho, p, dim, c->st);
    } else {
      launch_xpby(r, rho_new / rho, p, dim, c->st);
    }
    rho = rho_new;
    LAUNCHCHK(c);
  }
  *iters = it;
  *err = sqrt(rn2 / bn2);
  if (!std::isfinite(*err))
    return fail(c, YQ_ERR_NAN, "non-finite right-hand side or residual in the CG");
  return rn2 < thr ? YQ_OK : YQ_ERR_NOT_CONVERGED;
}

// out = (S + shift) x with the explicit real S, block by block (device pointers)
int symv_blocks(yq_ctx* c, double shift, const double* x, double* out) {
  TRY(blas(c));
  const double one = 1.0;
  HIPCHK(c, hipMemcpyAsync(out, x, sizeof(double) * srlen(c), hipMemcpyDeviceToDevice, c->st));
  for (const SrBlock& b : sr_blocks(c))
    BLASCHK(c, rocblas_dsymv(c->blas, rocblas_fill_lower, (rocblas_int)b.d, &one, b.S,
                             (rocblas_int)b.d, x + b.off, 1, &shift, out + b.off, 1));
  return YQ_OK;
}

// |(S + shift) x - b| / |b| in FP64
int residual64(yq_ctx* c, double shift, const double* x, const double* b, double* out) {
  const size_t dim = srlen(c);
  TRY(ensure(c, c->vtmp, sizeof(double) * std::max(dim, (size_t)c->ns * c->m)));
  double* t = c->vtmp.as<double>();
  if (c->sr_solver == YQ_SOLVE_EXPLICIT && c->cplx && c->sr_prec != YQ_MIXED) {
    TRY(blas(c));
    HIPCHK(c, hipMemcpyAsync(t, x, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
    const rocblas_double_complex one(1.0, 0.0), sh(shift, 0.0);
    BLASCHK(c, rocblas_zhemv(c->blas, rocblas_fill_lower, (rocblas_int)c->dim, &one,
                             reinterpret_cast<const rocblas_double_complex*>(c->S.p),
                             (rocblas_int)c->dim,
                             reinterpret_cast<const rocblas_double_complex*>(x), 1, &sh,
                             reinterpret_cast<rocblas_double_complex*>(t), 1));
  } else if (c->sr_solver == YQ_SOLVE_EXPLICIT && c->sr_prec != YQ_MIXED) {
    TRY(symv_blocks(c, shift, x, t));
  } else {  // the FP64 operator on O (a mixed explicit S is not exact)
    TRY(apply64(c, shift, x, t));
  }
  launch_axpy(-1.0, b, t, dim, c->st);
  LAUNCHCHK(c);
  double rr, bb;
  TRY(dot(c, t, t, dim, &rr));
  TRY(dot(c, b, b, dim, &bb));
  *out = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
  return YQ_OK;
}

int solve_explicit_c(yq_ctx* c, double shift, const double* b, double* x) {
  const size_t dim = c->dim;
  TRY(blas(c));
  TRY(ensure(c, c->M, sizeof(double) * 2 * dim * dim));
  TRY(ensure(c, c->info, sizeof(rocblas_int)));
  launch_copy_shift_c(c->S.as<double>(), c->M.as<double>(), (int)dim, shift, c->st);
  LAUNCHCHK(c);
  auto* M = reinterpret_cast<rocblas_double_complex*>(c->M.p);
  {
    ProfGuard g("zpotrf", c->st);
    BLASCHK(c, rocsolver_zpotrf(c->blas, rocblas_fill_lower, (rocblas_int)dim, M,
                                (rocblas_int)dim, c->info.as<rocblas_int>()));
  }
  rocblas_int info = 0;
  HIPCHK(c, hipMemcpyAsync(&info, c->info.p, sizeof(info), hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  if (info != 0)
    return fail(c, YQ_ERR_NOT_SPD,
                "Cholesky of S + shift I failed at pivot " + std::to_string(info));
  HIPCHK(c, hipMemcpyAsync(x, b, sizeof(double) * 2 * dim, hipMemcpyDeviceToDevice, c->st));
  BLASCHK(c, rocsolver_zpotrs(c->blas, rocblas_fill_lower, (rocblas_int)dim, 1, M,
                              (rocblas_int)dim, reinterpret_cast<rocblas_double_complex*>(x),
                              (rocblas_int)dim));
  return YQ_OK;
}

// Cholesky of S + shift I and its solve, block by block (one block unless
// the machine is AmplitudePhase; its blocks are the separate amplitude and
// phase solves of RunAmplitudePhase.hpp:96-111)
int solve_explicit(yq_ctx* c, double shift, const double* b, double* x) {
  if (c->cplx) return solve_explicit_c(c, shift, b, x);
  const std::vector<SrBlock> blk = sr_blocks(c);
  size_t msize = 0;
  for (const SrBlock& k : blk) msize += k.d * k.d;
  TRY(blas(c));
  TRY(ensure(c, c->M, sizeof(double) * msize));
  TRY(ensure(c, c->info, sizeof(rocblas_int) * blk.size()));
  HIPCHK(c, hipMemcpyAsync(x, b, sizeof(double) * srlen(c), hipMemcpyDeviceToDevice, c->st));
  double* M = c->M.as<double>();
  for (size_t q = 0; q < blk.size(); ++q) {
    const rocblas_int d = (rocblas_int)blk[q].d;
    launch_copy_lower_shift(blk[q].S, M, d, shift, c->st);
    LAUNCHCHK(c);
    {
      ProfGuard g("dpotrf", c->st);
      BLASCHK(c, rocsolver_dpotrf(c->blas, rocblas_fill_lower, d, M, d,
                                  c->info.as<rocblas_int>() + q));
    }
    BLASCHK(c, rocsolver_dpotrs(c->blas, rocblas_fill_lower, d, 1, M, d, x + blk[q].off, d));
    M += blk[q].d * blk[q].d;
  }
  std::vector<rocblas_int> info(blk.size(), 0);
  HIPCHK(c, hipMemcpyAsync(info.data(), c->info.p, sizeof(rocblas_int) * blk.size(),
                           hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  for (size_t q = 0; q < blk.size(); ++q)
    if (info[q] != 0)
      return fail(c, YQ_ERR_NOT_SPD,
                  "Cholesky of S + shift I failed at pivot " +
                      std::to_string(info[q] + (int)blk[q].off));
  return YQ_OK;
}

// Mixed explicit solve: Cholesky of the split-bf16 S + shift I in FP32
// (rocSOLVER spotrf / spotrs) inside FP64 iterative refinement,
// x += (S32 + shift)^-1 r, r = b - (S + shift) x with the FP64 operator on
// the stored O, until |r| / |b| < tol (30 rounds at most).
int solve_explicit_mixed(yq_ctx* c, double shift, const double* b, double* x, double tol,
                         int* rounds, double* err) {
  const size_t dim = c->dim;
  TRY(blas(c));
  TRY(ensure(c, c->M, sizeof(float) * dim * dim));
  TRY(ensure(c, c->info, sizeof(rocblas_int)));
  TRY(ensure(c, c->vr, sizeof(double) * dim));
  TRY(ensure(c, c->vp, sizeof(double) * dim));
  TRY(ensure(c, c->vAp, sizeof(double) * dim));
  TRY(ensure(c, c->vz, sizeof(double) * dim));
  float* M = c->M.as<float>();
  launch_lower_shift_f(c->S.as<double>(), M, dim, shift, c->st);
  LAUNCHCHK(c);
  {
    ProfGuard g("spotrf", c->st);
    BLASCHK(c, rocsolver_spotrf(c->blas, rocblas_fill_lower, (rocblas_int)dim, M,
                                (rocblas_int)dim, c->info.as<rocblas_int>()));
  }
  rocblas_int info = 0;
  HIPCHK(c, hipMemcpyAsync(&info, c->info.p, sizeof(info), hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  if (info != 0)
    return fail(c, YQ_ERR_NOT_SPD,
                "FP32 Cholesky of S + shift I failed at pivot " + std::to_string(info));
  double *r = c->vr.as<double>(), *d = c->vp.as<double>(), *Ax = c->vAp.as<double>();
  float* rf = c->vz.as<float>();
  double bn2;
  TRY(dot(c, b, b, dim, &bn2));
  HIPCHK(c, hipMemsetAsync(x, 0, sizeof(double) * dim, c->st));
  HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
  *rounds = 0;
  *err = bn2 > 0.0 ? 1.0 : 0.0;
  if (!std::isfinite(bn2)) return fail(c, YQ_ERR_NAN, "non-finite right-hand side");
  for (int k = 0; k < 30 && bn2 > 0.0; ++k) {
    launch_to_float(r, rf, dim, c->st);
    LAUNCHCHK(c);
    BLASCHK(c, rocsolver_spotrs(c->blas, rocblas_fill_lower, (rocblas_int)dim, 1, M,
                                (rocblas_int)dim, rf, (rocblas_int)dim));
    launch_float_to_double(rf, d, dim, c->st);
    launch_axpy(1.0, d, x, dim, c->st);
    LAUNCHCHK(c);
    TRY(apply64(c, shift, x, Ax));
    HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
    launch_axpy(-1.0, Ax, r, dim, c->st);
    LAUNCHCHK(c);
    double rn2;
    TRY(dot(c, r, r, dim, &rn2));
    *rounds = k + 1;
    *err = sqrt(rn2 / bn2);
    if (!std::isfinite(*err)) return fail(c, YQ_ERR_NAN, "non-finite residual");
    if (*err < tol) break;
  }
  return *err < tol ? YQ_OK : YQ_ERR_NOT_CONVERGED;
}

// Mixed explicit solve of a complex machine: FP64 Cholesky (rocSOLVER zpotrf)
// of the split-bf16 S + shift I, then FP64 iterative refinement x += (S16 +
// shift)^-1 r, r = b - (S + shift) x with the exact FP64 operator on the
// stored conj(O) (apply64), until |r| / |b| < tol (30 rounds at most).
int solve_explicit_mixed_c(yq_ctx* c, double shift, const double* b, double* x, double tol,
                           int* rounds, double* err) {
  const size_t dim = c->dim, len = 2 * dim;
  TRY(blas(c));
  TRY(ensure(c, c->M, sizeof(double) * 2 * dim * dim));
  TRY(ensure(c, c->info, sizeof(rocblas_int)));
  TRY(ensure(c, c->vr, sizeof(double) * len));
  TRY(ensure(c, c->vp, sizeof(double) * len));
  TRY(ensure(c, c->vAp, sizeof(double) * len));
  launch_copy_shift_c(c->S.as<double>(), c->M.as<double>(), (int)dim, shift, c->st);
  LAUNCHCHK(c);
  auto* M = reinterpret_cast<rocblas_double_complex*>(c->M.p);
  {
    ProfGuard g("zpotrf", c->st);
    BLASCHK(c, rocsolver_zpotrf(c->blas, rocblas_fill_lower, (rocblas_int)dim, M,
                                (rocblas_int)dim, c->info.as<rocblas_int>()));
  }
  rocblas_int info = 0;
  HIPCHK(c, hipMemcpyAsync(&info, c->info.p, sizeof(info), hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  if (info != 0)
    return fail(c, YQ_ERR_NOT_SPD,
                "Cholesky of the split-bf16 S + shift I failed at pivot " + std::to_string(info));
  double *r = c->vr.as<double>(), *d = c->vp.as<double>(), *Ax = c->vAp.as<double>();
  double bn2;
  TRY(dot(c, b, b, len, &bn2));
  HIPCHK(c, hipMemsetAsync(x, 0, sizeof(double) * len, c->st));
  HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * len, hipMemcpyDeviceToDevice, c->st));
  *rounds = 0;
  *err = bn2 > 0.0 ? 1.0 : 0.0;
  if (!std::isfinite(bn2)) return fail(c, YQ_ERR_NAN, "non-finite right-hand side");
  for (int k = 0; k < 30 && bn2 > 0.0; ++k) {
    HIPCHK(c, hipMemcpyAsync(d, r, sizeof(double) * len, hipMemcpyDeviceToDevice, c->st));
    BLASCHK(c, rocsolver_zpotrs(c->blas, rocblas_fill_lower, (rocblas_int)dim, 1, M,
                                (rocblas_int)dim, reinterpret_cast<rocblas_double_complex*>(d),
                                (rocblas_int)dim));
    launch_axpy(1.0, d, x, len, c->st);
    LAUNCHCHK(c);
    TRY(apply64(c, shift, x, Ax));
    HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * len, hipMemcpyDeviceToDevice, c->st));
    launch_axpy(-1.0, Ax, r, len, c->st);
    LAUNCHCHK(c);
    double rn2;
    TRY(dot(c, r, r, len, &rn2));
    *rounds = k + 1;
    *err = sqrt(rn2 / bn2);
    if (!std::isfinite(*err)) return fail(c, YQ_ERR_NAN, "non-finite residual");
    if (*err < tol) break;
  }
  return *err < tol ? YQ_OK : YQ_ERR_NOT_CONVERGED;
}

// FP64 iterative refinement around the mixed-precision inner CG:
// x += CG_mixed((S+shift) d = r), r = b - (S+shift) x in FP64.
int solve_mixed(yq_ctx* c, double shift, const double* b, double* x, double tol,
                int max_iter, int* iters, double* err, bool x0) {
  const size_t dim = c->dim;
  TRY(ensure(c, c->vr, sizeof(double) * dim));
  TRY(ensure(c, c->vAp, sizeof(double) * dim));
  TRY(ensure(c, c->vp, sizeof(double) * dim));
  TRY(ensure(c, c->dotp, sizeof(double) * 512));
  TRY(ensure(c, c->dscal, sizeof(double) * 4));
  double *r = c->vr.as<double>(), *Ax = c->vAp.as<double>(), *d = c->vp.as<double>();
  // <b, b> and <r, r> are formed on the device (in the order of a host sum
  // over the dot partials) and read together once per refinement round
  double* bb = c->dscal.as<double>() + 2;
  double* rr = bb + 1;
  launch_dot_dev(b, b, dim, c->dotp.as<double>(), bb, c->st);
  LAUNCHCHK(c);
  double bn2 = 0.0;
  auto read_rel = [&](double* rel) -> int {
    launch_dot_dev(r, r, dim, c->dotp.as<double>(), rr, c->st);
    LAUNCHCHK(c);
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, bb, sizeof(h), hipMemcpyDeviceToHost, c->st));
    TRY(sync(c));
    bn2 = h[0];
    *rel = bn2 == 0.0 ? 0.0 : sqrt(h[1] / bn2);
    if (!std::isfinite(*rel))
      return fail(c, YQ_ERR_NAN, "non-finite right-hand side or residual in the mixed solve");
    return YQ_OK;
  };
  *iters = 0;
  *err = 0.0;
  // b = 0: x = 0 (the inner CG on r = b = 0 stops at once and leaves d = 0,
  // so x stays 0 either way)
  double rel = 1.0;
  if (x0) {  // warm start: r = b - (S + shift) x
    TRY(apply64(c, shift, x, Ax));
    HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
    launch_axpy(-1.0, Ax, r, dim, c->st);
    LAUNCHCHK(c);
    TRY(read_rel(&rel));
    if (bn2 == 0.0) {
      HIPCHK(c, hipMemsetAsync(x, 0, sizeof(double) * dim, c->st));
      return YQ_OK;
    }
  } else {
    HIPCHK(c, hipMemsetAsync(x, 0, sizeof(double) * dim, c->st));
    HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
  }
  if (c->precond) TRY(ensure_diag(c));
  int total = 0;
  // (a NaN or infinite b also stops the inner CG at once with YQ_ERR_NAN)
  for (int outer = 0; outer < 30 && total < max_iter && !(rel < tol); ++outer) {
    int it = 0;
    const double inner_tol = std::max(tol / rel, 1e-4) < 1.0 ? std::max(tol / rel, 1e-4) : 0.5;
    int rc = mixed_cg(c->mixed, shift, c->omean.as<double>(),
                      c->precond ? c->diagS.as<double>() : nullptr, r, d, inner_tol,
                      max_iter - total, &it, c->st, c->err);
    total += it;
    if (rc != YQ_OK && rc != YQ_ERR_NOT_CONVERGED) return rc;
    launch_axpy(1.0, d, x, dim, c->st);
    LAUNCHCHK(c);
    TRY(apply64(c, shift, x, Ax));
    HIPCHK(c, hipMemcpyAsync(r, b, sizeof(double) * dim, hipMemcpyDeviceToDevice, c->st));
    launch_axpy(-1.0, Ax, r, dim, c->st);
    LAUNCHCHK(c);
    TRY(read_rel(&rel));
    if (bn2 == 0.0) break;
    if (rel < tol) break;
  }
  *iters = total;
  *err = rel;
  return rel < tol ? YQ_OK : YQ_ERR_NOT_CONVERGED;
}

int solve(yq_ctx* c, double shift, double tol, int max_iter, const double* b_dev,
          double* x_dev, double* resid, int* iters) {
  if (!c->have_sr) return fail(c, YQ_ERR_STATE, "call yq_sr_build first");
  if (!(shift >= 0.0)) return fail(c, YQ_ERR_ARG, "shift must be >= 0");
  int it = 0;
  double err = 0.0;
  int rc = YQ_OK;
  if (c->sr_solver == YQ_SOLVE_EXPLICIT && c->sr_prec == YQ_MIXED && !c->ap) {
    rc = c->cplx ? solve_explicit_mixed_c(c, shift, b_dev, x_dev, tol > 0 ? tol : 1e-12, &it,
                                          &err)
                 : solve_explicit_mixed(c, shift, b_dev, x_dev, tol > 0 ? tol : 1e-12, &it,
                                        &err);
    if (rc != YQ_OK && rc != YQ_ERR_NOT_CONVERGED) return rc;
  } else if (c->sr_solver == YQ_SOLVE_EXPLICIT) {
    rc = solve_explicit(c, shift, b_dev, x_dev);
    if (rc) return rc;
    TRY(residual64(c, shift, x_dev, b_dev, &err));
  } else {
    if (tol <= 0 || max_iter <= 0) return fail(c, YQ_ERR_ARG, "tol and max_iter must be > 0");
    // warm start from the previous iterative solution (same dimension)
    const bool x0 = c->warm_start && c->have_xprev;
    if (x0)
      HIPCHK(c, hipMemcpyAsync(x_dev, c->xprev.p, sizeof(double) * srlen(c),
                               hipMemcpyDeviceToDevice, c->st));
    if (c->sr_prec == YQ_MIXED) {
      rc = solve_mixed(c, shift, b_dev, x_dev, tol, max_iter, &it, &err, x0);
      if (rc != YQ_OK && rc != YQ_ERR_NOT_CONVERGED) return rc;
    } else {
      rc = cg64(c, shift, b_dev, x_dev, tol, max_iter, &it, &err, x0);
      if (rc != YQ_OK && rc != YQ_ERR_NOT_CONVERGED) return rc;
      double e2;
      TRY(residual64(c, shift, x_dev, b_dev, &e2));
      err = e2;
    }
    TRY(ensure(c, c->xprev, sizeof(double) * srlen(c)));
    HIPCHK(c, hipMemcpyAsync(c->xprev.p, x_dev, sizeof(double) * srlen(c),
                             hipMemcpyDeviceToDevice, c->st));
    c->have_xprev = true;
  }
  TRY(sync(c));
  if (resid) *resid = err;
  if (iters) *iters = it;
  if (rc == YQ_ERR_NOT_CONVERGED)
    return fail(c, rc, "iterative solve reached max_iter with relative residual " +
                           std::to_string(err));
  return YQ_OK;
}

int has_nan(yq_ctx* c, const double* dev, size_t n, int* out) {
  std::vector<double> h(n);
  HIPCHK(c, hipMemcpyAsync(h.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  *out = 0;
  for (double v : h)
    if (!std::isfinite(v)) {
      *out = 1;
      break;
    }
  return YQ_OK;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

YQ_API int yq_ctx_create(int device, yq_ctx** out) {
  if (!out) return YQ_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return YQ_ERR_HIP;
  }
  if (device < 0 || device >= count) return YQ_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return YQ_ERR_HIP;
  // the fixed scratch array of the allocation-free entry point: its address
  // on this device, looked up once per device with the device current (a
  // failure is reported by that entry point, which is its only user)
  (void)ff_scratch_init(device);
  yq_ctx* c = new yq_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return YQ_ERR_HIP;
  }
  *out = c;
  return YQ_OK;
}

YQ_API int yq_ctx_destroy(yq_ctx* c) {
  if (!c) return YQ_OK;
  for (yq_ctx* sub : c->subs) yq_ctx_destroy(sub);
  c->subs.clear();
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->st);
  mixed_group_release(c->group);
  mixed_release(c->mixed);
  for (Buf* b : c->bufs)
    if (b->p) (void)hipFree(b->p);
  if (c->blas) rocblas_destroy_handle(c->blas);
  (void)hipStreamDestroy(c->st);
  delete c;
  return YQ_OK;
}

YQ_API const char* yq_last_error(const yq_ctx* c) { return c ? c->err.c_str() : "null context"; }

// ---------------------------------------------------------------- machine
YQ_API int yq_rbm_create(yq_ctx* c, int dtype, int n, int m, int use_bias) {
  if (!c) return YQ_ERR_ARG;
  if (dtype != YQ_DTYPE_F64 && dtype != YQ_DTYPE_C128)
    return fail(c, YQ_ERR_UNSUPPORTED, "dtype must be 1 (double) or 2 (complex double)");
  if (n < 2 || n > kMaxSites || m < 1 || m > kMaxHidden)
    return fail(c, YQ_ERR_ARG, "need 2 <= n <= 4096 and 1 <= m <= 8192");
  c->n = n;
  c->m = m;
  c->use_bias = use_bias ? 1 : 0;
  c->cplx = dtype == YQ_DTYPE_C128;
  c->ap = false;
  c->dim = (size_t)n * m + (use_bias ? (size_t)n + m : 0);
  const size_t np = ((size_t)n * m + n + m) * ncomp(c);
  TRY(ensure(c, c->params, sizeof(double) * np));
  if (c->cplx) {
    const size_t nm = (size_t)n * m;
    TRY(ensure(c, c->cpd, sizeof(double) * (4 * nm + 2 * n)));
    TRY(ensure(c, c->cpf, sizeof(float) * (4 * nm + 2 * n)));
    double* d = c->cpd.as<double>();
    float* f = c->cpf.as<float>();
    c->planes = CplxPlanes{d, d + nm, d + 2 * nm, d + 3 * nm, d + 4 * nm, d + 4 * nm + n,
                           f, f + nm, f + 2 * nm, f + 3 * nm, f + 4 * nm, f + 4 * nm + n};
  } else {
    TRY(ensure(c, c->WT64, sizeof(double) * (size_t)n * m));
    TRY(ensure(c, c->W32, sizeof(float) * (size_t)n * m));
    TRY(ensure(c, c->WT32, sizeof(float) * (size_t)n * m));
    TRY(ensure(c, c->a32, sizeof(float) * (size_t)n));
  }
  HIPCHK(c, hipMemsetAsync(c->params.p, 0, sizeof(double) * np, c->st));
  c->have_rbm = true;
  c->have_batch = false;
  c->have_sampler = false;
  c->have_xprev = false;
  TRY(refresh_params(c));
  return sync(c);
}

YQ_API int yq_chain_lanes(int m) { return chain_group(m < 1 ? 1 : m); }

YQ_API int yq_rbm_dims(const yq_ctx* c, int* n, int* m, int* dim) {
  if (!c || !c->have_rbm) return YQ_ERR_STATE;
  if (n) *n = c->n;
  if (m) *m = c->m;
  if (dim) *dim = (int)c->dim;
  return YQ_OK;
}

YQ_API int yq_rbm_set_params(yq_ctx* c, const double* p) {
  TRY(need_rbm(c));
  if (!p) return fail(c, YQ_ERR_ARG, "null params");
  const size_t nm = (size_t)c->n * c->m;
  std::vector<double> h((nm + c->n + c->m) * ncomp(c), 0.0);
  memcpy(h.data(), p, sizeof(double) * vlen(c));
  for (double v : h)
    if (!std::isfinite(v)) return fail(c, YQ_ERR_NAN, "non-finite parameter");
  HIPCHK(c, hipMemcpyAsync(c->params.p, h.data(), sizeof(double) * h.size(),
                           hipMemcpyHostToDevice, c->st));
  TRY(refresh_params(c));
  return sync(c);
}

YQ_API int yq_rbm_get_params(yq_ctx* c, double* p) {
  TRY(need_rbm(c));
  if (!p) return fail(c, YQ_ERR_ARG, "null output");
  HIPCHK(c, hipMemcpyAsync(p, c->params.p, sizeof(double) * vlen(c), hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

YQ_API int yq_rbm_update_params(yq_ctx* c, const double* delta) {
  TRY(need_rbm(c));
  if (!delta) return fail(c, YQ_ERR_ARG, "null delta");
  TRY(ensure(c, c->vtmp, sizeof(double) * vlen(c)));
  HIPCHK(c, hipMemcpyAsync(c->vtmp.p, delta, sizeof(double) * vlen(c), hipMemcpyHostToDevice, c->st));
  launch_axpy(1.0, c->vtmp.as<double>(), c->params.as<double>(), vlen(c), c->st);
  LAUNCHCHK(c);
  TRY(refresh_params(c));
  return sync(c);
}

YQ_API int yq_rbm_init_random(yq_ctx* c, uint64_t seed, double sigma) {
  TRY(need_rbm(c));
  const int n = c->n, m = c->m;
  const size_t nm = (size_t)n * m;
  const size_t nc = ncomp(c);
  std::vector<double> h((nm + n + m) * nc, 0.0);
  // draw order of RBM::initializeRandom: a, b, then W column by column; a
  // complex entry takes (re, im) from two consecutive draws (RBM.hpp:327-349)
  std::vector<double*> order;
  auto take = [&](size_t idx) {
    for (size_t k = 0; k < nc; ++k) order.push_back(&h[idx * nc + k]);
  };
  if (c->use_bias) {
    for (int i = 0; i < n; ++i) take(nm + i);
    for (int j = 0; j < m; ++j) take(nm + n + j);
  }
  for (size_t q = 0; q < nm; ++q) take(q);
  for (size_t q = 0; q < order.size(); ++q) {
    const u32x4 x = draw(seed, q / 2, 0u, kDomParams);
    const double u1 = 1.0 - u01_double(x.x, x.y);
    const double u2 = u01_double(x.z, x.w);
    const double rad = sqrt(-2.0 * log(u1));
    const double z = (q & 1) ? rad * sin(2.0 * M_PI * u2) : rad * cos(2.0 * M_PI * u2);
    *order[q] = sigma * z;
  }
  HIPCHK(c, hipMemcpyAsync(c->params.p, h.data(), sizeof(double) * h.size(),
                           hipMemcpyHostToDevice, c->st));
  TRY(refresh_params(c));
  return sync(c);
}

YQ_API int yq_rbm_has_nan(yq_ctx* c, int* out) {
  TRY(need_rbm(c));
  if (!out) return fail(c, YQ_ERR_ARG, "null output");
  return has_nan(c, c->params.as<double>(), vlen(c), out);
}

// ---------------------------------------------------------------- checkpoints
// RunRBM.hpp:155-165 (w%04d.dat through a unique_ptr) and
// AbstractRunner.hpp:83-94 (initializeFrom); encoding in checkpoint.hip
YQ_API int yq_rbm_save(yq_ctx* c, const char* path, int elem, int pointer_wrapper) {
  TRY(need_rbm(c));
  if (!path) return fail(c, YQ_ERR_ARG, "null path");
  const bool cplx_elem = elem == YQ_CKPT_C64 || elem == YQ_CKPT_C128;
  if (elem < YQ_CKPT_F32 || elem > YQ_CKPT_C128 || cplx_elem != (c->cplx != 0))
    return fail(c, YQ_ERR_ARG, "checkpoint element type does not match the machine");
  std::vector<double> h(vlen(c));
  TRY(yq_rbm_get_params(c, h.data()));
  const size_t bytes = yq_ckpt_rbm_bytes(elem, c->n, c->m, c->use_bias, pointer_wrapper);
  std::vector<unsigned char> buf(bytes);
  if (yq_ckpt_rbm_encode(elem, c->n, c->m, c->use_bias, h.data(), pointer_wrapper, buf.data(),
                         bytes) != YQ_OK)
    return fail(c, YQ_ERR_ARG, "checkpoint encoding failed");
  FILE* f = fopen(path, "wb");
  if (!f) return fail(c, YQ_ERR_ARG, std::string("cannot open ") + path + " for writing");
  const size_t wr = fwrite(buf.data(), 1, bytes, f);
  const int cl = fclose(f);
  if (wr != bytes || cl != 0) return fail(c, YQ_ERR_ARG, std::string("short write to ") + path);
  return YQ_OK;
}

YQ_API int yq_rbm_load(yq_ctx* c, const char* path, int elem, int pointer_wrapper) {
  if (!c) return YQ_ERR_ARG;
  if (!path) return fail(c, YQ_ERR_ARG, "null path");
  if (elem < YQ_CKPT_F32 || elem > YQ_CKPT_C128)
    return fail(c, YQ_ERR_ARG, "unknown checkpoint element type");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(c, YQ_ERR_ARG, std::string("cannot open ") + path);
  std::vector<unsigned char> buf;
  unsigned char chunk[1 << 16];
  size_t got;
  while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  fclose(f);
  int n = 0, m = 0, ub = 0;
  if (yq_ckpt_rbm_decode(elem, buf.data(), buf.size(), pointer_wrapper, &n, &m, &ub, nullptr,
                         0) != YQ_OK)
    return fail(c, YQ_ERR_ARG, std::string(path) +
                                   ": not an RBM checkpoint of this element type and form");
  const bool cplx_elem = elem == YQ_CKPT_C64 || elem == YQ_CKPT_C128;
  const size_t dim = (size_t)n * m + (ub ? (size_t)n + m : 0);
  std::vector<double> h(dim * (cplx_elem ? 2 : 1));
  if (yq_ckpt_rbm_decode(elem, buf.data(), buf.size(), pointer_wrapper, &n, &m, &ub, h.data(),
                         h.size()) != YQ_OK)
    return fail(c, YQ_ERR_ARG, std::string(path) + ": malformed RBM checkpoint");
  TRY(yq_rbm_create(c, cplx_elem ? YQ_DTYPE_C128 : YQ_DTYPE_F64, n, m, ub));
  return yq_rbm_set_params(c, h.data());
}

// ---------------------------------------------------------------- Hamiltonian
YQ_API int yq_ham_set(yq_ctx* c, const yq_ham_desc* d) {
  if (!c) return YQ_ERR_ARG;
  if (!d) return fail(c, YQ_ERR_ARG, "null descriptor");
  if (d->kind < YQ_TFI || d->kind > YQ_J1J2) return fail(c, YQ_ERR_ARG, "unknown Hamiltonian kind");
  if (d->n < 3) return fail(c, YQ_ERR_ARG, "Hamiltonian needs n >= 3");
  for (double v : {d->J, d->h, d->delta, d->J2})
    if (!std::isfinite(v)) return fail(c, YQ_ERR_ARG, "non-finite coupling");
  c->ham = *d;
  c->have_ham = true;
  c->have_eloc = false;
  c->have_sr = false;
  return YQ_OK;
}

// ---------------------------------------------------------------- sampler
YQ_API int yq_sampler_config(yq_ctx* c, int sweeper, int n_temps, int n_chains_per,
                             int sweep_factor, int precision, uint64_t seed) {
  TRY(need_rbm(c));
  if (sweeper != YQ_LOCAL && sweeper != YQ_EXCHANGE) return fail(c, YQ_ERR_ARG, "unknown sweeper");
  // SamplerMT.hpp:48: nTmps even or 1; one block of nTmps wavefronts per chain group
  if (n_temps < 1 || n_temps > 16 || (n_temps > 1 && n_temps % 2))
    return fail(c, YQ_ERR_ARG, "n_temps must be 1 or even and <= 16");
  if (n_temps > 1 && c->ap)
    return fail(c, YQ_ERR_UNSUPPORTED, "parallel tempering of AmplitudePhase machines");
  if (n_chains_per < 1 || sweep_factor < 1) return fail(c, YQ_ERR_ARG, "chains and sweep factor >= 1");
  if ((long long)n_temps * n_chains_per > (1 << 24)) return fail(c, YQ_ERR_ARG, "too many chains");
  if (precision != YQ_FP64 && precision != YQ_MIXED) return fail(c, YQ_ERR_ARG, "unknown precision");
  c->skind = sweeper;
  c->ntemps = n_temps;
  c->chains = n_temps * n_chains_per;
  c->sweep_factor = sweep_factor;
  c->sprec = precision;
  c->seed = seed;
  c->rand_epoch = 0;
  TRY(ensure(c, c->sig_state, (size_t)c->chains * c->n));
  TRY(ensure(c, c->ctr, sizeof(uint64_t) * c->chains));
  TRY(ensure(c, c->acc, sizeof(uint64_t) * c->chains));
  HIPCHK(c, hipMemsetAsync(c->ctr.p, 0, sizeof(uint64_t) * c->chains, c->st));
  HIPCHK(c, hipMemsetAsync(c->acc.p, 0, sizeof(uint64_t) * c->chains, c->st));
  c->have_sampler = true;
  TRY(sync(c));
  return yq_sampler_randomize(c, sweeper == YQ_EXCHANGE ? c->n / 2 : -1);
}

YQ_API int yq_sampler_randomize(yq_ctx* c, int nup) {
  TRY(need_rbm(c));
  if (!c->have_sampler) return fail(c, YQ_ERR_STATE, "call yq_sampler_config first");
  const int n = c->n;
  if (nup > n) return fail(c, YQ_ERR_ARG, "nup > n");
  std::vector<int8_t> h((size_t)c->chains * n);
  const uint64_t base = c->rand_epoch << 20;
  for (int ch = 0; ch < c->chains; ++ch) {
    int8_t* s = h.data() + (size_t)ch * n;
    if (nup < 0) {
      for (int i = 0; i < n; i += 128) {
        const u32x4 x = draw(c->seed, base + (uint64_t)(i / 128), (uint32_t)ch, kDomRandomize);
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
        for (int q = 0; q < 128 && i + q < n; ++q) s[i + q] = ((w[q >> 5] >> (q & 31)) & 1u) ? -1 : 1;
      }
    } else {
      for (int i = 0; i < n; ++i) s[i] = i < nup ? -1 : 1;
      for (int i = n - 1; i > 0; --i) {
        const u32x4 x = draw(c->seed, base + (uint64_t)(n - 1 - i), (uint32_t)ch, kDomRandomize);
        const int j = (int)index_from(x.x, (uint32_t)(i + 1));
        std::swap(s[i], s[j]);
      }
    }
  }
  ++c->rand_epoch;
  HIPCHK(c, hipMemcpyAsync(c->sig_state.p, h.data(), h.size(), hipMemcpyHostToDevice, c->st));
  return sync(c);
}

YQ_API int yq_sampler_set_sigma(yq_ctx* c, const int8_t* sigma) {
  TRY(need_rbm(c));
  if (!c->have_sampler) return fail(c, YQ_ERR_STATE, "call yq_sampler_config first");
  if (!sigma) return fail(c, YQ_ERR_ARG, "null sigma");
  const size_t cnt = (size_t)c->chains * c->n;
  for (size_t q = 0; q < cnt; ++q)
    if (sigma[q] != 1 && sigma[q] != -1) return fail(c, YQ_ERR_ARG, "spins must be +1 or -1");
  HIPCHK(c, hipMemcpyAsync(c->sig_state.p, sigma, cnt, hipMemcpyHostToDevice, c->st));
  return sync(c);
}

YQ_API int yq_sampler_get_state(yq_ctx* c, int8_t* sigma, uint64_t* ctr, uint64_t* acc) {
  TRY(need_rbm(c));
  if (!c->have_sampler) return fail(c, YQ_ERR_STATE, "call yq_sampler_config first");
  if (sigma)
    HIPCHK(c, hipMemcpyAsync(sigma, c->sig_state.p, (size_t)c->chains * c->n,
                             hipMemcpyDeviceToHost, c->st));
  if (ctr)
    HIPCHK(c, hipMemcpyAsync(ctr, c->ctr.p, sizeof(uint64_t) * c->chains, hipMemcpyDeviceToHost, c->st));
  if (acc)
    HIPCHK(c, hipMemcpyAsync(acc, c->acc.p, sizeof(uint64_t) * c->chains, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

}  // extern "C"

namespace {
// Arguments of one sample() call; allocates the batch.  The caller launches.
int sample_args(yq_ctx* c, int n_sweeps, int n_discard, SampleArgs* out) {
  TRY(need_rbm(c));
  if (!c->have_sampler) return fail(c, YQ_ERR_STATE, "call yq_sampler_config first");
  if (n_sweeps < 1 || n_discard < 0) return fail(c, YQ_ERR_ARG, "n_sweeps >= 1, n_discard >= 0");
  // one sample per chain group and sweep (the beta = 1 slot, SamplerMT.hpp:166-173)
  const long long ns = (long long)(c->chains / c->ntemps) * n_sweeps;
  if (ns > (1ll << 30)) return fail(c, YQ_ERR_ARG, "too many samples");
  c->ns = (int)ns;
  c->bprec = c->sprec;
  const size_t rs = c->sprec == YQ_MIXED ? sizeof(float) : sizeof(double);
  TRY(ensure(c, c->bsig, (size_t)ns * c->n));
  TRY(ensure(c, c->btheta, rs * ncomp(c) * (size_t)ns * c->m));
  SampleArgs& a = *out;
  a = SampleArgs{};
  a.W64 = W64(c);
  a.b64 = b64(c);
  const bool f32 = c->sprec == YQ_MIXED;
  if (c->cplx) {
    a.Wr = f32 ? (const void*)c->planes.Wr32 : (const void*)c->planes.Wr64;
    a.Wi = f32 ? (const void*)c->planes.Wi32 : (const void*)c->planes.Wi64;
    a.ar = f32 ? (const void*)c->planes.ar32 : (const void*)c->planes.ar64;
  } else {
    a.Wr = f32 ? (const void*)c->W32.p : (const void*)W64(c);
    a.ar = f32 ? (const void*)c->a32.p : (const void*)a64(c);
  }
  a.sigma = c->sig_state.as<int8_t>();
  a.ctr = c->ctr.as<uint64_t>();
  a.acc = c->acc.as<uint64_t>();
  a.out_sigma = c->bsig.as<int8_t>();
  a.out_theta = c->btheta.p;
  a.C = c->chains;
  a.n = c->n;
  a.m = c->m;
  a.nprop = c->n * c->sweep_factor;
  a.n_sweeps = n_sweeps;
  a.n_discard = n_discard;
  a.npad = (c->n + 15) & ~15;
  a.kind = c->skind;
  a.mixed = c->sprec == YQ_MIXED;
  // an AmplitudePhase machine samples |sqrt(RBM)|^2 = RBM: the RBM chains at
  // beta = 1/2 (logRatioRe = logRatio / 2, AmplitudePhaseState.hpp:52-55)
  a.beta = c->ap ? 0.5 : 1.0;
  a.ntemps = c->ntemps;
  a.cplx = c->cplx ? 1 : 0;
  if (a.ntemps > 1) {
    if (c->m > kWaveHidden)
      return fail(c, YQ_ERR_UNSUPPORTED, "parallel tempering keeps one wave per slot: m <= 512");
    int dev = 0, lds_max = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess || lds_max <= 0)
      lds_max = 64 * 1024;
    if (pt_lds_bytes(a) + 256 > (size_t)lds_max)
      return fail(c, YQ_ERR_UNSUPPORTED, "parallel tempering: exchange area exceeds LDS");
  }
  a.seed = c->seed;
  return YQ_OK;
}
void batch_ready(yq_ctx* c) {
  c->have_batch = true;
  c->have_eloc = false;
  c->have_sr = false;
  c->weighted = false;
}
}  // namespace

extern "C" {

YQ_API int yq_sample(yq_ctx* c, int n_sweeps, int n_discard, int* n_samples) {
  SampleArgs a;
  TRY(sample_args(c, n_sweeps, n_discard, &a));
  if (!c->cplx && c->ntemps == 1) {  // theta of the chains from the MFMA pass
    const size_t rs = c->sprec == YQ_MIXED ? sizeof(float) : sizeof(double);
    TRY(ensure(c, c->tbuf, rs * (size_t)c->chains * c->m));
    a.theta_buf = c->tbuf.p;
  }
  if (c->cplx) launch_sample_c(a, c->st);
  else launch_sample(a, c->st);
  LAUNCHCHK(c);
  TRY(sync(c));
  batch_ready(c);
  if (n_samples) *n_samples = c->ns;
  return YQ_OK;
}

YQ_API int yq_samples_get(yq_ctx* c, int8_t* sigma, double* theta) {
  TRY(need_rbm(c));
  if (!c->have_batch) return fail(c, YQ_ERR_STATE, "no samples");
  const size_t cnt = (size_t)c->ns * c->m * ncomp(c);
  if (sigma)
    HIPCHK(c, hipMemcpyAsync(sigma, c->bsig.p, (size_t)c->ns * c->n, hipMemcpyDeviceToHost, c->st));
  if (theta) {
    if (c->bprec == YQ_MIXED) {
      std::vector<float> f(cnt);
      HIPCHK(c, hipMemcpyAsync(f.data(), c->btheta.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, c->st));
      TRY(sync(c));
      for (size_t q = 0; q < cnt; ++q) theta[q] = f[q];
    } else {
      HIPCHK(c, hipMemcpyAsync(theta, c->btheta.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->st));
    }
  }
  return sync(c);
}

YQ_API int yq_rbm_log_ratio(yq_ctx* c, int ns, const int8_t* sigma, int k, const int* sites,
                            double* out) {
  TRY(need_rbm(c));
  if (ns < 1 || k < 1 || !sigma || !sites || !out)
    return fail(c, YQ_ERR_ARG, "need ns >= 1, k >= 1 and arrays");
  const int n = c->n, m = c->m;
  for (size_t q = 0; q < (size_t)ns * n; ++q)
    if (sigma[q] != 1 && sigma[q] != -1) return fail(c, YQ_ERR_ARG, "spins must be +1 or -1");
  for (size_t q = 0; q < (size_t)ns * k; ++q)
    if (sites[q] >= n) return fail(c, YQ_ERR_ARG, "flip site out of range");
  const size_t nc = ncomp(c);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_sig = up((size_t)ns * n), b_sites = up(sizeof(int) * (size_t)ns * k);
  const size_t b_th = up(sizeof(double) * nc * (size_t)ns * m);
  const size_t b_out = up(sizeof(double) * nc * (size_t)ns);
  TRY(ensure(c, c->lrbuf, b_sig + b_sites + b_th + b_out));
  char* base = static_cast<char*>(c->lrbuf.p);
  int8_t* d_sig = reinterpret_cast<int8_t*>(base);
  int* d_sites = reinterpret_cast<int*>(base + b_sig);
  double* d_th = reinterpret_cast<double*>(base + b_sig + b_sites);
  double* d_out = reinterpret_cast<double*>(base + b_sig + b_sites + b_th);
  HIPCHK(c, hipMemcpyAsync(d_sig, sigma, (size_t)ns * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(d_sites, sites, sizeof(int) * (size_t)ns * k, hipMemcpyHostToDevice,
                           c->st));
  if (c->cplx) launch_theta_batch_c(W64(c), b64(c), d_sig, ns, n, m, d_th, 0, c->st);
  else launch_theta_batch(W64(c), b64(c), d_sig, ns, n, m, d_th, 0, c->st);
  launch_log_ratio_vec(W64(c), a64(c), d_sig, d_th, d_sites, ns, n, m, k, c->cplx ? 1 : 0, d_out,
                       c->st);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(out, d_out, sizeof(double) * nc * ns, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

YQ_API int yq_samples_set(yq_ctx* c, int ns, const int8_t* sigma, int precision) {
  TRY(need_rbm(c));
  if (ns < 1 || !sigma) return fail(c, YQ_ERR_ARG, "need ns >= 1 and sigma");
  if (precision != YQ_FP64 && precision != YQ_MIXED) return fail(c, YQ_ERR_ARG, "unknown precision");
  const size_t cnt = (size_t)ns * c->n;
  for (size_t q = 0; q < cnt; ++q)
    if (sigma[q] != 1 && sigma[q] != -1) return fail(c, YQ_ERR_ARG, "spins must be +1 or -1");
  const size_t rs = precision == YQ_MIXED ? sizeof(float) : sizeof(double);
  TRY(ensure(c, c->bsig, cnt));
  TRY(ensure(c, c->btheta, rs * ncomp(c) * (size_t)ns * c->m));
  HIPCHK(c, hipMemcpyAsync(c->bsig.p, sigma, cnt, hipMemcpyHostToDevice, c->st));
  if (c->cplx)
    launch_theta_batch_c(W64(c), b64(c), c->bsig.as<int8_t>(), ns, c->n, c->m, c->btheta.p,
                         precision == YQ_MIXED, c->st);
  else
    launch_theta_batch(W64(c), b64(c), c->bsig.as<int8_t>(), ns, c->n, c->m, c->btheta.p,
                       precision == YQ_MIXED, c->st);
  LAUNCHCHK(c);
  c->ns = ns;
  c->bprec = precision;
  batch_ready(c);
  return sync(c);
}

// ---------------------------------------------------------------- exact sums
namespace {
// Basis/BasisFull.hpp (nup < 0: 0 .. 2^n - 1) or Basis/BasisJz.hpp (the
// integers with nup set bits, increasing, by the same next-permutation step)
int exact_basis(yq_ctx* c, int nup, std::vector<uint32_t>& out) {
  const int n = c->n;
  if (n > 24) return fail(c, YQ_ERR_UNSUPPORTED, "exact sums need n <= 24");
  if (nup > n) return fail(c, YQ_ERR_ARG, "nup exceeds n");
  out.clear();
  if (nup < 0) {
    out.resize((size_t)1 << n);
    for (size_t i = 0; i < out.size(); ++i) out[i] = (uint32_t)i;
    return YQ_OK;
  }
  if (nup == 0) {
    out.push_back(0u);
    return YQ_OK;
  }
  uint32_t v = (1u << nup) - 1u;
  const uint32_t last = v << (n - nup);
  for (;;) {
    out.push_back(v);
    if (v == last) break;
    const uint32_t t = v | (v - 1u);
    v = (t + 1u) | (((~t & (0u - ~t)) - 1u) >> (__builtin_ctz(v) + 1));
  }
  return YQ_OK;
}

// the basis as the batch (FP64 theta) with its |psi|^2 weights
int exact_prepare(yq_ctx* c, int nup) {
  std::vector<uint32_t> basis;
  TRY(exact_basis(c, nup, basis));
  const int B = (int)basis.size();
  const size_t nc = ncomp(c);
  TRY(ensure(c, c->bbasis, sizeof(uint32_t) * B));
  TRY(ensure(c, c->bsig, (size_t)B * c->n));
  TRY(ensure(c, c->btheta, sizeof(double) * nc * (size_t)B * c->m));
  TRY(ensure(c, c->blogpsi, sizeof(double) * nc * B));
  TRY(ensure(c, c->bw, sizeof(double) * B));
  TRY(ensure(c, c->brs, sizeof(double) * B));
  TRY(ensure(c, c->bpart, sizeof(double) * 2 * exact_partials(B)));
  HIPCHK(c, hipMemcpyAsync(c->bbasis.p, basis.data(), sizeof(uint32_t) * B,
                           hipMemcpyHostToDevice, c->st));
  launch_basis_spins(c->bbasis.as<uint32_t>(), B, c->n, c->bsig.as<int8_t>(), c->st);
  if (c->cplx)
    launch_theta_batch_c(W64(c), b64(c), c->bsig.as<int8_t>(), B, c->n, c->m, c->btheta.p, 0,
                         c->st);
  else
    launch_theta_batch(W64(c), b64(c), c->bsig.as<int8_t>(), B, c->n, c->m, c->btheta.p, 0,
                       c->st);
  launch_exact_weights(c->bsig.as<int8_t>(), c->btheta.as<double>(),
                       c->use_bias ? a64(c) : nullptr, B, c->n, c->m, (int)c->cplx,
                       c->blogpsi.as<double>(), c->bw.as<double>(), c->brs.as<double>(),
                       c->bpart.as<double>(), c->st, c->ap ? 1.0 : 2.0);
  LAUNCHCHK(c);
  c->ns = B;
  c->bprec = YQ_FP64;
  batch_ready(c);
  c->weighted = true;
  return YQ_OK;
}
}  // namespace

YQ_API int yq_exact_batch(yq_ctx* c, int nup, int* basis_size) {
  TRY(need_rbm(c));
  TRY(exact_prepare(c, nup));
  if (basis_size) *basis_size = c->ns;
  return sync(c);
}

YQ_API int yq_exact_probs(yq_ctx* c, double* probs) {
  TRY(need_rbm(c));
  if (!c->have_batch || !c->weighted) return fail(c, YQ_ERR_STATE, "no exact batch");
  if (!probs) return fail(c, YQ_ERR_ARG, "null output");
  HIPCHK(c, hipMemcpyAsync(probs, c->bw.p, sizeof(double) * c->ns, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

YQ_API int yq_exact_sample(yq_ctx* c, int nup, int n_samples, uint64_t seed, int* indices) {
  TRY(need_rbm(c));
  if (n_samples < 1) return fail(c, YQ_ERR_ARG, "need n_samples >= 1");
  TRY(exact_prepare(c, nup));
  const int B = c->ns;
  TRY(ensure(c, c->bcdf, sizeof(double) * B));
  TRY(ensure(c, c->bpart, sizeof(double) * std::max(2 * exact_partials(B), exact_scan_tiles(B))));
  TRY(ensure(c, c->xsig, (size_t)n_samples * c->n));
  TRY(ensure(c, c->vtmp, sizeof(int) * (size_t)n_samples));
  launch_exact_cdf(c->bw.as<double>(), B, c->bcdf.as<double>(), c->bpart.as<double>(), c->st);
  launch_exact_draw(c->bcdf.as<double>(), B, c->bbasis.as<uint32_t>(), c->n, n_samples, seed,
                    c->xsig.as<int8_t>(), c->vtmp.as<int>(), c->st);
  LAUNCHCHK(c);
  // the draws become an ordinary (unweighted) batch
  const size_t cnt = (size_t)n_samples * c->n;
  TRY(ensure(c, c->bsig, cnt));
  TRY(ensure(c, c->btheta, sizeof(double) * ncomp(c) * (size_t)n_samples * c->m));
  HIPCHK(c, hipMemcpyAsync(c->bsig.p, c->xsig.p, cnt, hipMemcpyDeviceToDevice, c->st));
  if (c->cplx)
    launch_theta_batch_c(W64(c), b64(c), c->bsig.as<int8_t>(), n_samples, c->n, c->m,
                         c->btheta.p, 0, c->st);
  else
    launch_theta_batch(W64(c), b64(c), c->bsig.as<int8_t>(), n_samples, c->n, c->m,
                       c->btheta.p, 0, c->st);
  LAUNCHCHK(c);
  if (indices)
    HIPCHK(c, hipMemcpyAsync(indices, c->vtmp.p, sizeof(int) * n_samples, hipMemcpyDeviceToHost,
                             c->st));
  c->ns = n_samples;
  c->bprec = YQ_FP64;
  batch_ready(c);
  return sync(c);
}

// ---------------------------------------------------------------- energy
YQ_API int yq_eloc(yq_ctx* c, double* eloc, double* mean, double* var) {
  TRY(need_rbm(c));
  TRY(compute_eloc(c));
  if (eloc)
    HIPCHK(c, hipMemcpyAsync(eloc, c->eloc.p, sizeof(double) * ecomp(c) * c->ns,
                             hipMemcpyDeviceToHost, c->st));
  if (mean) *mean = c->e_mean;
  if (var) *var = c->e_var;
  return sync(c);
}

// ---------------------------------------------------------------- SR
YQ_API int yq_sr_build(yq_ctx* c, int solver, int precision) {
  TRY(need_rbm(c));
  if (!c->have_batch) return fail(c, YQ_ERR_STATE, "no samples");
  return sr_build(c, solver, precision);
}

YQ_API int yq_sr_fisher_average(yq_ctx* c, double beta2, int reset) {
  TRY(need_rbm(c));
  if (!c->have_sr || c->sr_solver != YQ_SOLVE_EXPLICIT)
    return fail(c, YQ_ERR_STATE, "fisher average needs an explicit yq_sr_build");
  if (!(beta2 >= 0.0 && beta2 < 1.0)) return fail(c, YQ_ERR_ARG, "need 0 <= beta2 < 1");
  const int nc = c->cplx ? 2 : 1;
  std::vector<SrBlock> blk;
  if (c->cplx) blk = {SrBlock{0, c->dim, nullptr, c->S.as<double>()}};
  else blk = sr_blocks(c);
  size_t len = 0;
  for (const SrBlock& b : blk) len += nc * b.d * b.d;
  if (reset || len != c->fisher_len) {
    TRY(ensure(c, c->savg, sizeof(double) * len));
    HIPCHK(c, hipMemsetAsync(c->savg.p, 0, sizeof(double) * len, c->st));
    c->fisher_t = 0;
    c->fisher_len = len;
  }
  ++c->fisher_t;
  const double corr = 1.0 - std::pow(beta2, c->fisher_t);
  size_t off = 0;
  for (const SrBlock& b : blk) {
    launch_ema_lower(c->savg.as<double>() + off, b.S, b.d, nc, beta2, corr, c->st);
    off += nc * b.d * b.d;
  }
  LAUNCHCHK(c);
  c->diag_ready = false;
  return sync(c);
}

YQ_API int yq_sr_get(yq_ctx* c, int item, double* out) {
  TRY(need_rbm(c));
  if (!out) return fail(c, YQ_ERR_ARG, "null output");
  if (item == YQ_SR_ELOC) {
    if (!c->have_eloc) return fail(c, YQ_ERR_STATE, "no local energies");
    HIPCHK(c, hipMemcpyAsync(out, c->eloc.p, sizeof(double) * ecomp(c) * c->ns,
                             hipMemcpyDeviceToHost, c->st));
    return sync(c);
  }
  if (!c->have_sr) return fail(c, YQ_ERR_STATE, "call yq_sr_build first");
  const size_t dim = c->dim;
  if (c->ap) {
    // the device keeps the amplitude and phase blocks apart; the host gets
    // rows [O_amp | O_phase] and the full (block-diagonal) S
    const size_t D = srlen(c), ns = (size_t)c->ns;
    const std::vector<SrBlock> blk = sr_blocks(c);
    switch (item) {
      case YQ_SR_F:
      case YQ_SR_OMEAN:
        HIPCHK(c, hipMemcpyAsync(out, (item == YQ_SR_F ? c->F : c->omean).p, sizeof(double) * D,
                                 hipMemcpyDeviceToHost, c->st));
        return sync(c);
      case YQ_SR_O: {
        std::vector<double> h(ns * D);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->O.p, sizeof(double) * h.size(),
                                 hipMemcpyDeviceToHost, c->st));
        TRY(sync(c));
        for (const SrBlock& b : blk) {
          const double* src = h.data() + ns * b.off;
          for (size_t r = 0; r < ns; ++r)
            std::copy(src + r * b.d, src + (r + 1) * b.d, out + r * D + b.off);
        }
        return YQ_OK;
      }
      case YQ_SR_S: {
        if (c->sr_solver != YQ_SOLVE_EXPLICIT)
          return fail(c, YQ_ERR_STATE, "S exists only for the explicit solver");
        std::fill(out, out + D * D, 0.0);
        for (const SrBlock& b : blk) {
          std::vector<double> h(b.d * b.d);
          HIPCHK(c, hipMemcpyAsync(h.data(), b.S, sizeof(double) * h.size(),
                                   hipMemcpyDeviceToHost, c->st));
          TRY(sync(c));
          for (size_t j = 0; j < b.d; ++j)
            for (size_t i = 0; i < b.d; ++i)  // lower triangle, mirrored
              out[(b.off + j) * D + b.off + i] = i >= j ? h[j * b.d + i] : h[i * b.d + j];
        }
        return YQ_OK;
      }
      default:
        return fail(c, YQ_ERR_ARG, "unknown item");
    }
  }
  if (c->cplx) {
    // the device holds conj(O_c), conj(omean) and the lower triangle of S
    // (column-major); the host gets O_c, omean and the full S (column-major)
    switch (item) {
      case YQ_SR_F:
        HIPCHK(c, hipMemcpyAsync(out, c->F.p, sizeof(double) * 2 * dim, hipMemcpyDeviceToHost,
                                 c->st));
        return sync(c);
      case YQ_SR_OMEAN:
      case YQ_SR_O: {
        if (item == YQ_SR_O && c->sr_solver == YQ_SOLVE_CG_FACTORED)
          return fail(c, YQ_ERR_STATE, "the factored solver does not materialise O");
        const size_t cnt = item == YQ_SR_O ? 2 * dim * c->ns : 2 * dim;
        const Buf& b = item == YQ_SR_O ? c->O : c->omean;
        HIPCHK(c, hipMemcpyAsync(out, b.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->st));
        TRY(sync(c));
        for (size_t q = 1; q < cnt; q += 2) out[q] = -out[q];
        return YQ_OK;
      }
      case YQ_SR_S: {
        if (c->sr_solver != YQ_SOLVE_EXPLICIT)
          return fail(c, YQ_ERR_STATE, "S exists only for the explicit solver");
        std::vector<double> h(2 * dim * dim);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->S.p, sizeof(double) * h.size(),
                                 hipMemcpyDeviceToHost, c->st));
        TRY(sync(c));
        for (size_t j = 0; j < dim; ++j)
          for (size_t i = 0; i < dim; ++i) {  // S_ij = conj(S_ji) above the diagonal
            const bool low = i >= j;
            const size_t src = low ? i + j * dim : j + i * dim;
            out[2 * (i + j * dim)] = h[2 * src];
            out[2 * (i + j * dim) + 1] = low ? h[2 * src + 1] : -h[2 * src + 1];
          }
        return YQ_OK;
      }
      default:
        return fail(c, YQ_ERR_ARG, "unknown item");
    }
  }
  switch (item) {
    case YQ_SR_F:
      HIPCHK(c, hipMemcpyAsync(out, c->F.p, sizeof(double) * dim, hipMemcpyDeviceToHost, c->st));
      break;
    case YQ_SR_OMEAN:
      HIPCHK(c, hipMemcpyAsync(out, c->omean.p, sizeof(double) * dim, hipMemcpyDeviceToHost, c->st));
      break;
    case YQ_SR_O:
      if (c->sr_solver == YQ_SOLVE_CG_FACTORED)
        return fail(c, YQ_ERR_STATE, "the factored solver does not materialise O");
      HIPCHK(c, hipMemcpyAsync(out, c->O.p, sizeof(double) * dim * c->ns, hipMemcpyDeviceToHost, c->st));
      break;
    case YQ_SR_S:
      if (c->sr_solver != YQ_SOLVE_EXPLICIT)
        return fail(c, YQ_ERR_STATE, "S exists only for the explicit solver");
      HIPCHK(c, hipMemcpyAsync(out, c->S.p, sizeof(double) * dim * dim, hipMemcpyDeviceToHost, c->st));
      TRY(sync(c));
      for (size_t j = 0; j < dim; ++j)  // mirror the lower triangle (column-major)
        for (size_t i = 0; i < j; ++i) out[j * dim + i] = out[i * dim + j];
      return YQ_OK;
    default:
      return fail(c, YQ_ERR_ARG, "unknown item");
  }
  return sync(c);
}

YQ_API int yq_sr_apply(yq_ctx* c, double shift, const double* v, double* out) {
  TRY(need_rbm(c));
  if (!c->have_sr) return fail(c, YQ_ERR_STATE, "call yq_sr_build first");
  if (!v || !out) return fail(c, YQ_ERR_ARG, "null vector");
  const size_t dim = srlen(c);
  TRY(ensure(c, c->vx, sizeof(double) * dim));
  TRY(ensure(c, c->vAp, sizeof(double) * dim));
  HIPCHK(c, hipMemcpyAsync(c->vx.p, v, sizeof(double) * dim, hipMemcpyHostToDevice, c->st));
  if (c->cplx && c->sr_solver == YQ_SOLVE_EXPLICIT) {
    TRY(blas(c));
    HIPCHK(c, hipMemcpyAsync(c->vAp.p, c->vx.p, sizeof(double) * dim, hipMemcpyDeviceToDevice,
                             c->st));
    const rocblas_double_complex one(1.0, 0.0), sh(shift, 0.0);
    BLASCHK(c, rocblas_zhemv(c->blas, rocblas_fill_lower, (rocblas_int)c->dim, &one,
                             reinterpret_cast<const rocblas_double_complex*>(c->S.p),
                             (rocblas_int)c->dim,
                             reinterpret_cast<const rocblas_double_complex*>(c->vx.p), 1, &sh,
                             reinterpret_cast<rocblas_double_complex*>(c->vAp.p), 1));
  } else if (c->sr_solver == YQ_SOLVE_EXPLICIT) {
    TRY(symv_blocks(c, shift, c->vx.as<double>(), c->vAp.as<double>()));
  } else if (c->sr_prec == YQ_MIXED) {
    const int rc = mixed_apply64(c->mixed, shift, c->omean.as<double>(), c->vx.as<double>(),
                                 c->vAp.as<double>(), c->st, c->err);
    if (rc) return rc;
  } else {
    TRY(apply64(c, shift, c->vx.as<double>(), c->vAp.as<double>()));
  }
  HIPCHK(c, hipMemcpyAsync(out, c->vAp.p, sizeof(double) * dim, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

YQ_API int yq_sr_solve(yq_ctx* c, double shift, double tol, int max_iter, const double* rhs,
                       double* x, double* resid, int* iters) {
  TRY(need_rbm(c));
  if (!c->have_sr) return fail(c, YQ_ERR_STATE, "call yq_sr_build first");
  const size_t dim = srlen(c);
  TRY(ensure(c, c->vx, sizeof(double) * dim));
  TRY(ensure(c, c->vb, sizeof(double) * std::max(dim, (size_t)c->ns)));
  const double* b = c->F.as<double>();
  if (rhs) {
    HIPCHK(c, hipMemcpyAsync(c->vb.p, rhs, sizeof(double) * dim, hipMemcpyHostToDevice, c->st));
    b = c->vb.as<double>();
  }
  int rc = solve(c, shift, tol, max_iter, b, c->vx.as<double>(), resid, iters);
  if (rc != YQ_OK && rc != YQ_ERR_NOT_CONVERGED) return rc;
  if (x) HIPCHK(c, hipMemcpyAsync(x, c->vx.p, sizeof(double) * dim, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  return rc;
}

// ---------------------------------------------------------------- FeedForward
}  // extern "C"

namespace {
int need_ff(yq_ctx* c) {
  if (!c) return YQ_ERR_ARG;
  if (!c->have_ff) return fail(c, YQ_ERR_STATE, "no FeedForward machine: call yq_ff_create");
  return YQ_OK;
}
size_t ff_rs(const yq_ctx* c) { return c->ff.dtype == 1 ? sizeof(double) : sizeof(float); }
// upload the fp64 master parameters in the machine dtype
int ff_upload(yq_ctx* c) {
  const size_t np = (size_t)c->ff_npar;
  if (c->ap) {  // the phase of an AmplitudePhase machine moved
    c->have_eloc = false;
    c->have_sr = false;
  }
  if (c->ff.dtype == 1) {
    HIPCHK(c, hipMemcpyAsync(c->ffp.p, c->ff_host.data(), sizeof(double) * np,
                             hipMemcpyHostToDevice, c->st));
  } else {
    std::vector<float> f(np);
    for (size_t q = 0; q < np; ++q) f[q] = (float)c->ff_host[q];
    HIPCHK(c, hipMemcpyAsync(c->ffp.p, f.data(), sizeof(float) * np, hipMemcpyHostToDevice, c->st));
    TRY(sync(c));
  }
  if (!c->ff_stack) launch_ff_transpose(ff_args(c, 1, nullptr), c->ff.dtype, c->st);
  LAUNCHCHK(c);
  return sync(c);
}
FFArgs ff_args(yq_ctx* c, int ns, const int8_t* d_sigma) {
  FFArgs a{};
  a.n = c->ff.n;
  a.C = c->ff.channels;
  a.K = c->ff.kernel;
  a.H = c->ff.hidden;
  a.ns = ns;
  a.params = c->ffp.p;
  a.sigma = d_sigma;
  if (c->ff_stack) {
    a.layers = c->ff_ltab.as<FFLayerDev>();
    a.nl = (int)c->ff_ldev.size();
    a.act_len = c->ff_act_len;
    a.dmax = c->ff_dmax;
    a.npar = c->ff_npar;
    a.sb = c->ff_sb;
  } else {
    // the transposed FC weight follows the parameters in the same buffer
    a.wt = static_cast<const char*>(c->ffp.p) + ff_rs(c) * (size_t)c->ff_npar;
  }
  return a;
}
// VJP partial rows (one per block) for ns configurations
int ff_vjp_rows(const yq_ctx* c, int ns) {
  if (!c->ff_stack) return ff_vjp_blocks(ns, c->ff.dtype);
  const int groups = (ns + c->ff_sb - 1) / c->ff_sb;
  return groups < 512 ? groups : 512;
}
// (paramDim, fanIn, fanOut) of the layers with parameters, in order
// (FeedForward.hpp:43-73; Conv1D.hpp:227-234, FullyConnected.hpp:181-188)
std::vector<std::array<long long, 3>> ff_param_layers(const yq_ctx* c) {
  std::vector<std::array<long long, 3>> v;
  if (c->ff_stack) {
    for (const FFLayerDev& L : c->ff_ldev) {
      if (L.kind == YQ_LAYER_CONV1D)
        v.push_back({L.npar, (long long)L.ic * L.K, (long long)L.oc * L.K});
      else if (L.kind == YQ_LAYER_FC)
        v.push_back({L.npar, L.ic, L.oc});
    }
    return v;
  }
  const yq_ff_desc& d = c->ff;
  const long long CN = (long long)d.channels * d.n;
  v.push_back({(long long)d.kernel * d.channels, d.kernel, (long long)d.channels * d.kernel});
  v.push_back({CN * d.hidden, CN, d.hidden});
  return v;
}
int ff_check_sigma(yq_ctx* c, int ns, const int8_t* sigma) {
  if (ns < 1 || !sigma) return fail(c, YQ_ERR_ARG, "need ns >= 1 and sigma");
  for (size_t q = 0; q < (size_t)ns * c->ff.n; ++q)
    if (sigma[q] != 1 && sigma[q] != -1) return fail(c, YQ_ERR_ARG, "spins must be +1 or -1");
  return YQ_OK;
}
// host double arrays <-> device machine-dtype arrays
int ff_to_host(yq_ctx* c, const void* dev, size_t count, double* out) {
  if (c->ff.dtype == 1) {
    HIPCHK(c, hipMemcpyAsync(out, dev, sizeof(double) * count, hipMemcpyDeviceToHost, c->st));
    return sync(c);
  }
  std::vector<float> f(count);
  HIPCHK(c, hipMemcpyAsync(f.data(), dev, sizeof(float) * count, hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  for (size_t q = 0; q < count; ++q) out[q] = f[q];
  return YQ_OK;
}
}  // namespace

extern "C" {

YQ_API int yq_ff_create(yq_ctx* c, const yq_ff_desc* d) {
  if (!c) return YQ_ERR_ARG;
  if (!d) return fail(c, YQ_ERR_ARG, "null descriptor");
  if (d->n < 2 || d->n > 1024 || d->channels < 1 || d->kernel < 1 || d->kernel > d->n ||
      d->hidden < 1 || (d->dtype != 0 && d->dtype != 1))
    return fail(c, YQ_ERR_ARG, "invalid FeedForward shape or dtype");
  const long long np = (long long)d->kernel * d->channels + (long long)d->channels * d->n * d->hidden;
  if (np > 96 * 256) return fail(c, YQ_ERR_UNSUPPORTED, "FeedForward limited to 24576 parameters");
  c->ap = false;
  c->ff = *d;
  c->ff_stack = false;
  c->ff_ldev.clear();
  c->ff_npar = (int)np;
  c->ff_host.assign((size_t)np, 0.0);
  FFArgs probe = ff_args(c, 1, nullptr);
  if (ff_smem_bytes(probe, d->dtype) > 64 * 1024)
    return fail(c, YQ_ERR_UNSUPPORTED, "FeedForward activations exceed the LDS budget");
  TRY(ensure(c, c->ffp, ff_rs(c) * ((size_t)np + (size_t)d->channels * d->n * d->hidden)));
  c->have_ff = true;
  return ff_upload(c);
}

YQ_API int yq_ff_create_stack(yq_ctx* c, int n, int dtype, int nl, const yq_ff_layer* ly) {
  if (!c) return YQ_ERR_ARG;
  if (!ly || nl < 1 || nl > 256) return fail(c, YQ_ERR_ARG, "need 1 to 256 layers");
  if (n < 1 || n > (1 << 20) || (dtype != 0 && dtype != 1))
    return fail(c, YQ_ERR_ARG, "invalid input size or dtype");
  // the running width through FeedForward::makeData (:170-184), each layer's
  // outputDim (Conv1D.hpp:103-105, FullyConnected.hpp, ActivationLayer.hpp:
  // 98-101, Mean.hpp:71-75)
  std::vector<FFLayerDev> t((size_t)nl);
  long long width = n, off = n, poff = 0, dmax = n;
  for (int i = 0; i < nl; ++i) {
    const yq_ff_layer& l = ly[i];
    FFLayerDev& L = t[(size_t)i];
    L = FFLayerDev{};
    L.kind = l.kind;
    L.in_dim = (int)width;
    L.in_off = (int)(off - width);
    L.p_off = (int)poff;
    long long out = 0, np = 0;
    if (l.kind == YQ_LAYER_CONV1D) {
      if (l.a < 1 || l.b < 1 || l.c < 1 || l.d < 1 || width % l.a)
        return fail(c, YQ_ERR_ARG, "Conv1D: invalid channels, kernel or stride for its input");
      const long long inSize = width / l.a;
      if (l.c / 2 > inSize || inSize / l.d < 1)
        return fail(c, YQ_ERR_ARG, "Conv1D: kernel or stride too large for its input");
      L.ic = l.a, L.oc = l.b, L.K = l.c, L.stride = l.d, L.bias = l.use_bias ? 1 : 0;
      out = (inSize / l.d) * l.b;
      np = (long long)l.a * l.b * l.c + (l.use_bias ? l.b : 0);
    } else if (l.kind == YQ_LAYER_FC) {
      if (l.a != width || l.b < 1)
        return fail(c, YQ_ERR_ARG, "FullyConnected: inputDim must equal the width before it");
      L.ic = l.a, L.oc = l.b, L.bias = l.use_bias ? 1 : 0;
      out = l.b;
      np = (long long)l.a * l.b + (l.use_bias ? l.b : 0);
    } else if (l.kind == YQ_LAYER_ACT) {
      if (l.a < YQ_ACT_IDENTITY || l.a > YQ_ACT_COS)
        return fail(c, YQ_ERR_ARG, "unknown activation");
      if ((l.a == YQ_ACT_TANH || l.a == YQ_ACT_COS) && !(l.param != 0.0 && std::isfinite(l.param)))
        return fail(c, YQ_ERR_ARG, "Tanh / Cos need a finite non-zero constant");
      if (!std::isfinite(l.param)) return fail(c, YQ_ERR_ARG, "non-finite activation parameter");
      L.act = l.a;
      L.prm = l.param;
      out = width;
    } else if (l.kind == YQ_LAYER_MEAN) {
      if (l.a < 1 || l.b < 1 || (long long)l.a * l.b != width || (l.c != 0 && l.c != 1))
        return fail(c, YQ_ERR_ARG, "Mean: input dims must multiply to the width, axis 0 or 1");
      L.rows = l.a, L.cols = l.b, L.axis = l.c;
      out = l.c == 0 ? l.b : l.a;
    } else {
      return fail(c, YQ_ERR_ARG, "unknown layer kind");
    }
    L.out_dim = (int)out;
    L.out_off = (int)off;
    L.npar = (int)np;
    width = out;
    off += out;
    poff += np;
    dmax = std::max(dmax, out);
    if (off > (1 << 24) || poff > (1LL << 30)) return fail(c, YQ_ERR_ARG, "network too large");
  }
  if (width != 1) return fail(c, YQ_ERR_ARG, "the last layer must leave one value");
  if (poff < 1) return fail(c, YQ_ERR_ARG, "the network has no parameters");
  // The C4 stack runs on its own kernels (ff.hip) when they can hold it.
  const bool c4 = nl == 5 && ly[0].kind == YQ_LAYER_CONV1D && ly[0].a == 1 && ly[0].d == 1 &&
                  !ly[0].use_bias && ly[1].kind == YQ_LAYER_ACT && ly[1].a == YQ_ACT_LNCOSH &&
                  ly[2].kind == YQ_LAYER_FC && !ly[2].use_bias && ly[3].kind == YQ_LAYER_ACT &&
                  ly[3].a == YQ_ACT_TANH && ly[3].param == 1.0 && ly[4].kind == YQ_LAYER_MEAN &&
                  ly[4].b == 1 && ly[4].c == 0 && ly[0].c <= n && n >= 2 && n <= 1024;
  if (c4) {
    const yq_ff_desc d{n, ly[0].b, ly[0].c, ly[2].b, dtype};
    FFArgs probe{};
    probe.n = d.n;
    probe.C = d.channels;
    probe.K = d.kernel;
    probe.H = d.hidden;
    if (poff <= 96 * 256 && ff_smem_bytes(probe, dtype) <= 64 * 1024) return yq_ff_create(c, &d);
  }
  const size_t rs = dtype == 1 ? sizeof(double) : sizeof(float);
  const size_t per = ffs_lds_bytes((int)off, (int)dmax, 1, dtype);
  int sb = (int)std::min<size_t>(8, (64 * 1024) / per);
  if (sb < 1) {
    int dev = 0, lds_max = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess || lds_max <= 0)
      lds_max = 64 * 1024;
    if (per > (size_t)lds_max)
      return fail(c, YQ_ERR_UNSUPPORTED, "FeedForward activations of one sample exceed the LDS");
    sb = 1;
  }
  c->ap = false;
  c->have_ff = false;
  c->ff = yq_ff_desc{n, 0, 0, 0, dtype};
  c->ff_stack = true;
  c->ff_ldev = t;
  c->ff_act_len = (int)off;
  c->ff_dmax = (int)dmax;
  c->ff_sb = sb;
  c->ff_npar = (int)poff;
  c->ff_host.assign((size_t)poff, 0.0);
  TRY(ensure(c, c->ff_ltab, sizeof(FFLayerDev) * t.size()));
  HIPCHK(c, hipMemcpyAsync(c->ff_ltab.p, c->ff_ldev.data(), sizeof(FFLayerDev) * t.size(),
                           hipMemcpyHostToDevice, c->st));
  TRY(ensure(c, c->ffp, rs * (size_t)poff));
  c->have_ff = true;
  return ff_upload(c);
}

YQ_API int yq_ff_act_len(const yq_ctx* c, int* len) {
  if (!c || !c->have_ff) return YQ_ERR_STATE;
  if (len)
    *len = c->ff_stack ? c->ff_act_len : 2 * c->ff.channels * c->ff.n + c->ff.hidden;
  return YQ_OK;
}

YQ_API int yq_ff_npar(const yq_ctx* c, int* npar) {
  if (!c || !c->have_ff) return YQ_ERR_STATE;
  if (npar) *npar = c->ff_npar;
  return YQ_OK;
}

YQ_API int yq_ff_set_params(yq_ctx* c, const double* p) {
  TRY(need_ff(c));
  if (!p) return fail(c, YQ_ERR_ARG, "null params");
  for (int q = 0; q < c->ff_npar; ++q)
    if (!std::isfinite(p[q])) return fail(c, YQ_ERR_NAN, "non-finite parameter");
  c->ff_host.assign(p, p + c->ff_npar);
  return ff_upload(c);
}

YQ_API int yq_ff_get_params(yq_ctx* c, double* p) {
  TRY(need_ff(c));
  if (!p) return fail(c, YQ_ERR_ARG, "null output");
  std::copy(c->ff_host.begin(), c->ff_host.end(), p);
  return YQ_OK;
}

YQ_API int yq_ff_update_params(yq_ctx* c, const double* delta) {
  TRY(need_ff(c));
  if (!delta) return fail(c, YQ_ERR_ARG, "null delta");
  for (int q = 0; q < c->ff_npar; ++q) c->ff_host[q] += delta[q];
  return ff_upload(c);
}

YQ_API int yq_ff_init_random(yq_ctx* c, uint64_t seed, int mode) {
  TRY(need_ff(c));
  if (mode < 0 || mode > 3) return fail(c, YQ_ERR_ARG, "mode must be 0..3");
  // layers with parameters, in order: (paramDim, fanIn, fanOut)
  const std::vector<std::array<long long, 3>> layers = ff_param_layers(c);
  size_t q = 0;
  for (const auto& L : layers) {
    const double fi = L[1], fo = L[2];
    const double s = mode == 0 ? sqrt(1.0 / fi) : mode == 1 ? sqrt(2.0 / (fi + fo))
                   : mode == 2 ? sqrt(2.0 / fi) : sqrt(6.0 / (fi + fo));
    for (long long k = 0; k < L[0]; ++k, ++q) {
      const u32x4 x = draw(seed, q / 2, 1u, kDomParams);
      if (mode == 3) {
        const double u = (q & 1) ? u01_double(x.z, x.w) : u01_double(x.x, x.y);
        c->ff_host[q] = s * (2.0 * u - 1.0);
      } else {
        const double u1 = 1.0 - u01_double(x.x, x.y), u2 = u01_double(x.z, x.w);
        const double rad = sqrt(-2.0 * log(u1));
        c->ff_host[q] = s * ((q & 1) ? rad * sin(2.0 * M_PI * u2) : rad * cos(2.0 * M_PI * u2));
      }
    }
  }
  return ff_upload(c);
}

YQ_API int yq_ff_forward_dev(yq_ctx* c, int ns, const int8_t* d_sigma, void* d_logpsi,
                             void* d_act) {
  TRY(need_ff(c));
  if (ns < 1 || !d_sigma || !d_logpsi) return fail(c, YQ_ERR_ARG, "need ns >= 1 and buffers");
  FFArgs a = ff_args(c, ns, d_sigma);
  a.out = d_logpsi;
  a.act_out = d_act;
  TRY(launch_ff(a, kFFForward, c->ff.dtype, c->st));
  LAUNCHCHK(c);
  return sync(c);
}

YQ_API int yq_ff_jacobian_dev(yq_ctx* c, int ns, const int8_t* d_sigma, void* d_jac,
                              void* d_logpsi) {
  TRY(need_ff(c));
  if (ns < 1 || !d_sigma || !d_jac) return fail(c, YQ_ERR_ARG, "need ns >= 1 and buffers");
  TRY(ensure(c, c->ff_out, ff_rs(c) * (size_t)ns));
  FFArgs a = ff_args(c, ns, d_sigma);
  a.jac = d_jac;
  a.out = d_logpsi ? d_logpsi : c->ff_out.p;
  TRY(launch_ff(a, kFFJacobian, c->ff.dtype, c->st));
  LAUNCHCHK(c);
  return sync(c);
}

YQ_API int yq_ff_vjp_dev(yq_ctx* c, int ns, const int8_t* d_sigma, const void* d_act,
                         const void* d_r, void* d_grad) {
  TRY(need_ff(c));
  if (ns < 1 || !d_sigma || !d_r || !d_grad) return fail(c, YQ_ERR_ARG, "need ns >= 1 and buffers");
  TRY(ensure(c, c->ff_part, ff_rs(c) * (size_t)ff_vjp_rows(c, ns) * c->ff_npar));
  FFArgs a = ff_args(c, ns, d_sigma);
  a.act_in = d_act;
  a.r = d_r;
  a.partial = c->ff_part.p;
  a.grad = d_grad;
  const int rc = launch_ff(a, kFFVjp, c->ff.dtype, c->st);
  if (rc) return fail(c, rc, "FeedForward VJP shape unsupported");
  LAUNCHCHK(c);
  return sync(c);
}

YQ_API int yq_ff_vjp_dev_noalloc(yq_ctx* c, int ns, const int8_t* d_sigma, const void* d_act,
                                 const void* d_r, void* d_grad) {
  TRY(need_ff(c));
  if (ns < 1 || !d_sigma || !d_r || !d_grad) return fail(c, YQ_ERR_ARG, "need ns >= 1 and buffers");
  void* scr = ff_scratch(c->device);
  if (!scr) return fail(c, YQ_ERR_HIP, "the fixed VJP scratch array has no address on this device");
  // one block's partial row is npar values; the array holds as many rows as
  // fit, and the grid-stride kernel covers any ns with that many blocks
  const size_t row = ff_rs(c) * (size_t)c->ff_npar;
  const size_t rows = kFFScratchBytes / row;
  if (rows < 1)
    return fail(c, YQ_ERR_ARG, "FeedForward parameters beyond the fixed VJP scratch (npar x "
                               "element size must be at most 16 MiB)");
  FFArgs a = ff_args(c, ns, d_sigma);
  a.act_in = d_act;
  a.r = d_r;
  a.partial = scr;
  a.grad = d_grad;
  const int nb = ff_vjp_rows(c, ns);
  a.vjp_blocks = (size_t)nb < rows ? nb : (int)rows;
  const int rc = launch_ff(a, kFFVjp, c->ff.dtype, c->st);
  if (rc) return fail(c, rc, "FeedForward VJP shape unsupported");
  LAUNCHCHK(c);
  return sync(c);
}

YQ_API int yq_ff_forward(yq_ctx* c, int ns, const int8_t* sigma, double* logpsi) {
  TRY(need_ff(c));
  TRY(ff_check_sigma(c, ns, sigma));
  if (!logpsi) return fail(c, YQ_ERR_ARG, "null output");
  TRY(ensure(c, c->ff_sig, (size_t)ns * c->ff.n));
  TRY(ensure(c, c->ff_out, ff_rs(c) * (size_t)ns));
  HIPCHK(c, hipMemcpyAsync(c->ff_sig.p, sigma, (size_t)ns * c->ff.n, hipMemcpyHostToDevice, c->st));
  TRY(yq_ff_forward_dev(c, ns, c->ff_sig.as<int8_t>(), c->ff_out.p, nullptr));
  return ff_to_host(c, c->ff_out.p, (size_t)ns, logpsi);
}

YQ_API int yq_ff_jacobian(yq_ctx* c, int ns, const int8_t* sigma, double* jac, double* logpsi) {
  TRY(need_ff(c));
  TRY(ff_check_sigma(c, ns, sigma));
  if (!jac) return fail(c, YQ_ERR_ARG, "null output");
  const size_t cnt = (size_t)ns * c->ff_npar;
  TRY(ensure(c, c->ff_sig, (size_t)ns * c->ff.n));
  TRY(ensure(c, c->ff_jac, ff_rs(c) * cnt));
  TRY(ensure(c, c->ff_grad, ff_rs(c) * (size_t)ns));
  HIPCHK(c, hipMemcpyAsync(c->ff_sig.p, sigma, (size_t)ns * c->ff.n, hipMemcpyHostToDevice, c->st));
  TRY(yq_ff_jacobian_dev(c, ns, c->ff_sig.as<int8_t>(), c->ff_jac.p, c->ff_grad.p));
  TRY(ff_to_host(c, c->ff_jac.p, cnt, jac));
  if (logpsi) TRY(ff_to_host(c, c->ff_grad.p, (size_t)ns, logpsi));
  return YQ_OK;
}

// ---------------------------------------------------------------- AmplitudePhase
}  // extern "C"

namespace {
int need_ap(yq_ctx* c) {
  if (!c) return YQ_ERR_ARG;
  if (!c->ap) return fail(c, YQ_ERR_STATE, "no AmplitudePhase machine: call yq_ap_enable");
  return YQ_OK;
}
void ap_reset(yq_ctx* c) {  // the wave function changed meaning: no batch carries over
  c->have_batch = false;
  c->have_eloc = false;
  c->have_sr = false;
  c->have_xprev = false;
}
}  // namespace

extern "C" {

YQ_API int yq_ap_enable(yq_ctx* c, int on) {
  if (!c) return YQ_ERR_ARG;
  if (!on) {
    if (c->ap) ap_reset(c);
    c->ap = false;
    return YQ_OK;
  }
  TRY(need_rbm(c));
  if (c->cplx)
    return fail(c, YQ_ERR_UNSUPPORTED, "the amplitude is a real RBM (YQ_DTYPE_F64)");
  TRY(need_ff(c));
  if (c->ff.dtype != 1)
    return fail(c, YQ_ERR_UNSUPPORTED, "the phase network must be double (dtype 1)");
  if (c->ff.n != c->n)
    return fail(c, YQ_ERR_ARG, "amplitude and phase act on different numbers of spins");
  if (c->have_sampler && c->ntemps > 1)
    return fail(c, YQ_ERR_UNSUPPORTED, "parallel tempering of AmplitudePhase machines");
  ap_reset(c);
  c->ap = true;
  return YQ_OK;
}

YQ_API int yq_ap_dims(const yq_ctx* c, int* dim_amp, int* dim_phase) {
  if (!c || !c->ap) return YQ_ERR_STATE;
  if (dim_amp) *dim_amp = (int)c->dim;
  if (dim_phase) *dim_phase = c->ff_npar;
  return YQ_OK;
}

YQ_API int yq_ap_get_params(yq_ctx* c, double* p) {
  TRY(need_ap(c));
  if (!p) return fail(c, YQ_ERR_ARG, "null output");
  TRY(yq_rbm_get_params(c, p));
  return yq_ff_get_params(c, p + c->dim);
}

YQ_API int yq_ap_set_params(yq_ctx* c, const double* p) {
  TRY(need_ap(c));
  if (!p) return fail(c, YQ_ERR_ARG, "null params");
  for (size_t q = 0; q < srlen(c); ++q)
    if (!std::isfinite(p[q])) return fail(c, YQ_ERR_NAN, "non-finite parameter");
  TRY(yq_rbm_set_params(c, p));
  return yq_ff_set_params(c, p + c->dim);
}

YQ_API int yq_ap_update_params(yq_ctx* c, const double* delta) {
  TRY(need_ap(c));
  if (!delta) return fail(c, YQ_ERR_ARG, "null delta");
  TRY(yq_rbm_update_params(c, delta));
  return yq_ff_update_params(c, delta + c->dim);
}

YQ_API int yq_ap_log_psi(yq_ctx* c, int ns, const int8_t* sigma, double* out) {
  TRY(need_ap(c));
  TRY(ff_check_sigma(c, ns, sigma));
  if (!out) return fail(c, YQ_ERR_ARG, "null output");
  const int n = c->n, m = c->m;
  TRY(ensure(c, c->ff_sig, (size_t)ns * n));
  TRY(ensure(c, c->vtmp, sizeof(double) * (size_t)ns * m));
  TRY(ensure(c, c->ff_out, sizeof(double) * (size_t)ns));
  TRY(ensure(c, c->ap_ffs, sizeof(double) * (size_t)ns));
  TRY(ensure(c, c->vy, sizeof(double) * 2 * (size_t)ns));
  HIPCHK(c, hipMemcpyAsync(c->ff_sig.p, sigma, (size_t)ns * n, hipMemcpyHostToDevice, c->st));
  const int8_t* ds = c->ff_sig.as<int8_t>();
  launch_theta_batch(W64(c), b64(c), ds, ns, n, m, c->vtmp.p, 0, c->st);
  launch_log_psi(ds, c->vtmp.as<double>(), c->use_bias ? a64(c) : nullptr, ns, n, m, 0,
                 c->ff_out.as<double>(), c->st);
  FFArgs f = ff_args(c, ns, ds);
  f.out = c->ap_ffs.p;
  TRY(launch_ff(f, kFFForward, 1, c->st));
  launch_ap_log_psi(c->ff_out.as<double>(), c->ap_ffs.as<double>(), ns, c->vy.as<double>(),
                    c->st);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(out, c->vy.p, sizeof(double) * 2 * (size_t)ns, hipMemcpyDeviceToHost,
                           c->st));
  return sync(c);
}

// ---------------------------------------------------------------- profiling
YQ_API int yq_prof_enable(yq_ctx* c, const char* labels) {
  if (!c) return YQ_ERR_ARG;
  TRY(sync(c));
  for (ProfEntry& e : g_prof.entries) {
    for (auto& p : e.pending) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
    for (hipEvent_t ev : e.open) (void)hipEventDestroy(ev);
  }
  g_prof = ProfState();
  g_prof_any = false;
  if (!labels || !*labels) return YQ_OK;
  std::string s(labels);
  size_t pos = 0;
  while (pos <= s.size()) {
    const size_t comma = s.find(',', pos);
    const std::string l = s.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
    if (!l.empty()) {
      g_prof.labels.push_back(l);
      g_prof.entries.emplace_back();
    }
    if (comma == std::string::npos) break;
    pos = comma + 1;
  }
  g_prof_any = !g_prof.labels.empty();
  return YQ_OK;
}

YQ_API int yq_prof_read(yq_ctx* c, const char* label, double* total_ms, long long* launches) {
  if (!c || !label) return YQ_ERR_ARG;
  TRY(sync(c));
  const int i = g_prof.find(label);
  if (i < 0) return fail(c, YQ_ERR_ARG, std::string("label not enabled: ") + label);
  ProfEntry& e = g_prof.entries[i];
  for (auto& p : e.pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) e.total_ms += ms;
    ++e.launches;
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  e.pending.clear();
  if (total_ms) *total_ms = e.total_ms;
  if (launches) *launches = e.launches;
  e.total_ms = 0.0;
  e.launches = 0;
  return YQ_OK;
}

YQ_API int yq_set_graph_mode(yq_ctx* c, int use_graphs) {
  if (!c) return YQ_ERR_ARG;
  if (use_graphs < 0 || use_graphs > 256 || (use_graphs > 1 && use_graphs % 2))
    return fail(c, YQ_ERR_ARG, "graph mode: 0, 1 or an even chunk length <= 256");
  const int chunk = use_graphs > 1 ? use_graphs : 8;
  c->mixed.use_graphs = use_graphs != 0;
  c->mixed.chunk_iters = chunk;
  if (c->group.chunk_iters != chunk && c->group.gexec) {  // a new capture
    (void)hipGraphExecDestroy(c->group.gexec);
    c->group.gexec = nullptr;
  }
  c->group.use_graphs = use_graphs != 0;
  c->group.chunk_iters = chunk;
  for (yq_ctx* sub : c->subs) {
    sub->mixed.use_graphs = use_graphs != 0;
    sub->mixed.chunk_iters = chunk;
  }
  return YQ_OK;
}

YQ_API int yq_set_tiles(yq_ctx* c, int col_tiles, int chunk) {
  if (!c) return YQ_ERR_ARG;
  const bool off = col_tiles == 0 && chunk == 0;
  if (!off && !((col_tiles == 2 || col_tiles == 4 || col_tiles == 8) &&
                (chunk == 128 || chunk == 256 || chunk == 512)))
    return fail(c, YQ_ERR_ARG, "tiles: col_tiles in {2, 4, 8}, chunk in {128, 256, 512}");
  for (yq_ctx* x : c->subs) {
    x->mixed.pin_ct = col_tiles;
    x->mixed.pin_chunk = chunk;
  }
  c->mixed.pin_ct = col_tiles;
  c->mixed.pin_chunk = chunk;
  return YQ_OK;
}

YQ_API int yq_set_autotune(yq_ctx* c, int on) {
  if (!c) return YQ_ERR_ARG;
  c->mixed.autotune = on != 0;
  for (yq_ctx* sub : c->subs) sub->mixed.autotune = on != 0;
  return YQ_OK;
}

YQ_API int yq_sr_set_options(yq_ctx* c, int precond, int warm_start) {
  if (!c) return YQ_ERR_ARG;
  if (precond != 0 && precond != 1) return fail(c, YQ_ERR_ARG, "precond must be 0 or 1");
  c->precond = precond;
  c->warm_start = warm_start ? 1 : 0;
  if (!c->warm_start) c->have_xprev = false;
  return YQ_OK;
}

// ---------------------------------------------------------------- runner
}  // extern "C"

namespace {
// constructFromSamples + solve + SGD update of RunRBM.hpp:176-212 on the
// current batch (E_loc already computed); *solve_rc keeps a
// YQ_ERR_NOT_CONVERGED of the solve while the update still happens.
int step_apply(yq_ctx* c, const yq_step_cfg* cfg, double* nv);
int step_update(yq_ctx* c, const yq_step_cfg* cfg, double* nv, double* err, int* it,
                int* solve_rc) {
  TRY(sr_build(c, cfg->solver, cfg->precision));
  const size_t len = srlen(c);
  TRY(ensure(c, c->vx, sizeof(double) * len));
  int rc = solve(c, cfg->lambda, cfg->cg_tol, cfg->cg_max_iter, c->F.as<double>(),
                 c->vx.as<double>(), err, it);
  if (rc != YQ_OK && rc != YQ_ERR_NOT_CONVERGED) return rc;
  *solve_rc = rc;
  return step_apply(c, cfg, nv);
}
// params -= lr x for the solution x in vx (|x| clipped to grad_clip), then the
// derived parameter copies; nv = |x|
int step_apply(yq_ctx* c, const yq_step_cfg* cfg, double* nv) {
  const size_t len = srlen(c);
  double vv;
  TRY(dot(c, c->vx.as<double>(), c->vx.as<double>(), len, &vv));
  *nv = sqrt(vv);
  double scale = -cfg->lr;
  if (cfg->grad_clip != 0.0 && *nv > cfg->grad_clip) scale *= cfg->grad_clip / *nv;
  launch_axpy(scale, c->vx.as<double>(), c->params.as<double>(), vlen(c), c->st);
  LAUNCHCHK(c);
  if (c->ap) {  // AmplitudePhase::updateParams (:153-158): the tail moves the phase network
    std::vector<double> tail((size_t)c->ff_npar);
    HIPCHK(c, hipMemcpyAsync(tail.data(), c->vx.as<double>() + c->dim,
                             sizeof(double) * tail.size(), hipMemcpyDeviceToHost, c->st));
    TRY(sync(c));
    for (size_t q = 0; q < tail.size(); ++q) c->ff_host[q] += scale * tail[q];
    TRY(ff_upload(c));
  }
  return refresh_params(c);
}
}  // namespace

extern "C" {

YQ_API int yq_sr_step(yq_ctx* c, const yq_step_cfg* cfg, yq_step_out* out) {
  TRY(need_rbm(c));
  if (!cfg || !out) return fail(c, YQ_ERR_ARG, "null cfg/out");
  if (!c->have_sampler) return fail(c, YQ_ERR_STATE, "call yq_sampler_config first");
  hipEvent_t e0, e1, e2;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventCreate(&e2));
  auto cleanup = [&]() {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
  };
  int rc = YQ_OK;
  do {
    if ((rc = yq_sr_set_options(c, cfg->precond, cfg->warm_start)) != YQ_OK) break;
    if (cfg->randomize && (rc = yq_sampler_randomize(c, cfg->nup)) != YQ_OK) break;
    if ((rc = hipEventRecord(e0, c->st) == hipSuccess ? YQ_OK : YQ_ERR_HIP)) break;
    int ns = 0;
    if ((rc = yq_sample(c, cfg->n_sweeps, cfg->n_discard, &ns)) != YQ_OK) break;
    if ((rc = hipEventRecord(e1, c->st) == hipSuccess ? YQ_OK : YQ_ERR_HIP)) break;
    if ((rc = compute_eloc(c)) != YQ_OK) break;
    int it = 0, solve_rc = YQ_OK;
    double err = 0.0, nv = 0.0;
    if ((rc = step_update(c, cfg, &nv, &err, &it, &solve_rc)) != YQ_OK) break;
    if ((rc = hipEventRecord(e2, c->st) == hipSuccess ? YQ_OK : YQ_ERR_HIP)) break;
    if ((rc = sync(c)) != YQ_OK) break;
    float t01 = 0, t12 = 0;
    (void)hipEventElapsedTime(&t01, e0, e1);
    (void)hipEventElapsedTime(&t12, e1, e2);
    out->energy = c->e_mean;
    out->energy_var = c->e_var;
    out->update_norm = nv;
    out->cg_err = err;
    out->cg_iters = it;
    out->n_samples = ns;
    out->sample_ms = t01;
    out->solve_ms = t12;
    if (!std::isfinite(nv) || !std::isfinite(c->e_mean)) {
      rc = fail(c, YQ_ERR_NAN, "non-finite energy or update");
      break;
    }
    rc = solve_rc;
  } while (0);
  cleanup();
  return rc;
}

// ---------------------------------------------------------------- many machines
}  // extern "C"

namespace {
int need_multi(yq_ctx* c) {
  if (!c) return YQ_ERR_ARG;
  if (c->subs.empty()) return fail(c, YQ_ERR_STATE, "call yq_multi_create first");
  return YQ_OK;
}
// status of a sub-context call, with its message moved to the owner
int from_sub(yq_ctx* c, yq_ctx* sub, int rc, int idx) {
  if (rc == YQ_OK) return rc;
  return fail(c, rc, "machine " + std::to_string(idx) + ": " + sub->err);
}
void drop_subs(yq_ctx* c) {
  for (yq_ctx* sub : c->subs) yq_ctx_destroy(sub);
  c->subs.clear();
}
// machines ordered by decreasing cost of a sweep (n * sweep factor * m)
std::vector<int> cost_order(const yq_ctx* c) {
  std::vector<int> ord(c->subs.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
  auto cost = [&](int i) {
    const yq_ctx* s = c->subs[i];
    return (long long)s->n * s->sweep_factor * s->m;
  };
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cost(a) > cost(b); });
  return ord;
}
int multi_sample(yq_ctx* c, int n_sweeps, int n_discard) {
  const int count = (int)c->subs.size();
  std::vector<SampleArgs> args(count);
  for (int i = 0; i < count; ++i)
    TRY(from_sub(c, c->subs[i], sample_args(c->subs[i], n_sweeps, n_discard, &args[i]), i));
  TRY(ensure(c, c->m_sargs, sizeof(SampleArgs) * count));
  HIPCHK(c, hipMemcpyAsync(c->m_sargs.p, args.data(), sizeof(SampleArgs) * count,
                           hipMemcpyHostToDevice, c->st));
  launch_sample_multi(c->m_sargs.as<SampleArgs>(), c->m_items.as<int2>(), c->m_nitems,
                      c->m_npad, c->multi_kind, c->multi_prec == YQ_MIXED, 0, c->st);
  LAUNCHCHK(c);
  TRY(sync(c));
  for (yq_ctx* sub : c->subs) batch_ready(sub);
  return YQ_OK;
}
int multi_eloc(yq_ctx* c) {
  const int count = (int)c->subs.size();
  std::vector<ElocArgs> args(count);
  size_t lds = 0;
  int max_blocks = 0, max_ns = 0;
  for (int i = 0; i < count; ++i) {
    yq_ctx* sub = c->subs[i];
    TRY(from_sub(c, sub, eloc_args(sub, &args[i]), i));
    lds = std::max(lds, eloc_stage_lds_bytes(args[i]));
    const size_t pairs = (size_t)args[i].ns * args[i].B;
    max_blocks = std::max(max_blocks, (int)((pairs + 255) / 256));
    max_ns = std::max(max_ns, args[i].ns);
  }
  if (lds > 64 * 1024) return fail(c, YQ_ERR_UNSUPPORTED, "E_loc tile exceeds LDS");
  TRY(ensure(c, c->m_eargs, sizeof(ElocArgs) * count));
  TRY(ensure(c, c->m_stats, sizeof(double) * 2 * count));
  HIPCHK(c, hipMemcpyAsync(c->m_eargs.p, args.data(), sizeof(ElocArgs) * count,
                           hipMemcpyHostToDevice, c->st));
  launch_eloc_multi(c->m_eargs.as<ElocArgs>(), count, max_blocks, max_ns, lds,
                    c->multi_prec == YQ_MIXED, c->m_stats.as<double>(), c->st);
  LAUNCHCHK(c);
  std::vector<double> h(2 * count);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->m_stats.p, sizeof(double) * h.size(),
                           hipMemcpyDeviceToHost, c->st));
  TRY(sync(c));
  for (int i = 0; i < count; ++i) {
    c->subs[i]->e_mean = h[2 * i];
    c->subs[i]->e_var = h[2 * i + 1];
    c->subs[i]->have_eloc = true;
  }
  return YQ_OK;
}
}  // namespace

extern "C" {

YQ_API int yq_multi_create(yq_ctx* c, int count, const yq_multi_desc* d, int sweeper,
                           int precision) {
  if (!c) return YQ_ERR_ARG;
  if (count < 1 || !d) return fail(c, YQ_ERR_ARG, "need count >= 1 and descriptors");
  if (sweeper != YQ_LOCAL && sweeper != YQ_EXCHANGE) return fail(c, YQ_ERR_ARG, "unknown sweeper");
  if (precision != YQ_FP64 && precision != YQ_MIXED) return fail(c, YQ_ERR_ARG, "unknown precision");
  for (int i = 0; i < count; ++i)
    if (d[i].m > kWaveHidden)
      return fail(c, YQ_ERR_UNSUPPORTED,
                  "the many-machine launch keeps one wave per chain: m <= 512 (machine " +
                      std::to_string(i) + ")");
  drop_subs(c);
  for (int i = 0; i < count; ++i) {
    yq_ctx* sub = nullptr;
    int rc = yq_ctx_create(c->device, &sub);
    if (rc != YQ_OK) {
      drop_subs(c);
      return fail(c, rc, "cannot create the context of machine " + std::to_string(i));
    }
    c->subs.push_back(sub);
    sub->mixed.use_graphs = c->mixed.use_graphs;
    sub->mixed.chunk_iters = c->mixed.chunk_iters;
    sub->mixed.autotune = c->mixed.autotune;
    rc = yq_rbm_create(sub, YQ_DTYPE_F64, d[i].n, d[i].m, d[i].use_bias);
    if (rc == YQ_OK) rc = yq_ham_set(sub, &d[i].ham);
    if (rc == YQ_OK)
      rc = yq_sampler_config(sub, sweeper, 1, d[i].chains, d[i].sweep_factor, precision,
                             d[i].seed);
    if (rc != YQ_OK) {
      rc = from_sub(c, sub, rc, i);
      drop_subs(c);
      return rc;
    }
  }
  c->multi_kind = sweeper;
  c->multi_prec = precision;
  // (machine, chain) items, the costliest machines first
  std::vector<int2> items;
  int npad = 16;
  for (int i : cost_order(c)) {
    const yq_ctx* sub = c->subs[i];
    npad = std::max(npad, (sub->n + 15) & ~15);
    for (int ch = 0; ch < sub->chains; ++ch) items.push_back(make_int2(i, ch));
  }
  c->m_nitems = (int)items.size();
  c->m_npad = npad;
  TRY(ensure(c, c->m_items, sizeof(int2) * items.size()));
  HIPCHK(c, hipMemcpyAsync(c->m_items.p, items.data(), sizeof(int2) * items.size(),
                           hipMemcpyHostToDevice, c->st));
  return sync(c);
}

YQ_API int yq_multi_set_params(yq_ctx* c, int idx, const double* p) {
  TRY(need_multi(c));
  if (idx < 0 || idx >= (int)c->subs.size()) return fail(c, YQ_ERR_ARG, "machine index");
  return from_sub(c, c->subs[idx], yq_rbm_set_params(c->subs[idx], p), idx);
}

YQ_API int yq_multi_get_params(yq_ctx* c, int idx, double* p) {
  TRY(need_multi(c));
  if (idx < 0 || idx >= (int)c->subs.size()) return fail(c, YQ_ERR_ARG, "machine index");
  return from_sub(c, c->subs[idx], yq_rbm_get_params(c->subs[idx], p), idx);
}

YQ_API int yq_multi_init_random(yq_ctx* c, uint64_t seed, double sigma) {
  TRY(need_multi(c));
  for (size_t i = 0; i < c->subs.size(); ++i)
    TRY(from_sub(c, c->subs[i], yq_rbm_init_random(c->subs[i], seed + i, sigma), (int)i));
  return YQ_OK;
}

YQ_API int yq_multi_sample(yq_ctx* c, int n_sweeps, int n_discard) {
  TRY(need_multi(c));
  return multi_sample(c, n_sweeps, n_discard);
}

YQ_API int yq_multi_samples_get(yq_ctx* c, int idx, int8_t* sigma, double* theta) {
  TRY(need_multi(c));
  if (idx < 0 || idx >= (int)c->subs.size()) return fail(c, YQ_ERR_ARG, "machine index");
  return from_sub(c, c->subs[idx], yq_samples_get(c->subs[idx], sigma, theta), idx);
}

YQ_API int yq_multi_sampler_state(yq_ctx* c, int idx, int8_t* sigma, uint64_t* ctr,
                                  uint64_t* acc) {
  TRY(need_multi(c));
  if (idx < 0 || idx >= (int)c->subs.size()) return fail(c, YQ_ERR_ARG, "machine index");
  return from_sub(c, c->subs[idx], yq_sampler_get_state(c->subs[idx], sigma, ctr, acc), idx);
}

YQ_API int yq_multi_eloc(yq_ctx* c, double* mean, double* var) {
  TRY(need_multi(c));
  TRY(multi_eloc(c));
  for (size_t i = 0; i < c->subs.size(); ++i) {
    if (mean) mean[i] = c->subs[i]->e_mean;
    if (var) var[i] = c->subs[i]->e_var;
  }
  return YQ_OK;
}

}  // extern "C"

namespace {
// fn(i) for i in [0, count) on up to n_threads host threads (each with the
// context's device current); the statuses land in rcs
void parallel_for(yq_ctx* c, int count, int n_threads, std::vector<int>& rcs,
                  const std::function<int(int)>& fn) {
  rcs.assign(count, YQ_OK);
  std::atomic<int> next(0);
  auto work = [&]() {
    (void)hipSetDevice(c->device);
    for (int i = next++; i < count; i = next++) rcs[i] = fn(i);
  };
  const int nt = std::max(1, std::min(count, n_threads > 0 ? n_threads : 16));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  (void)hipSetDevice(c->device);
}
// first failing status (machine index in idx), or YQ_OK
int first_fail(const std::vector<int>& rcs, int* idx) {
  for (size_t i = 0; i < rcs.size(); ++i)
    if (rcs[i] != YQ_OK) {
      *idx = (int)i;
      return rcs[i];
    }
  return YQ_OK;
}

// The SR solve of every machine (real RBMs, mixed factored solver, cold
// start): each machine runs the FP64 refinement of solve_mixed (the same
// rounds, inner tolerances and budgets), and the float inner CG of all the
// machines in a round runs grouped (mixed_cg_group), so each machine gets the
// x that solve_mixed would give it.  Per machine: err, iters and rc (YQ_OK or
// YQ_ERR_NOT_CONVERGED) in the outputs.
int solve_mixed_group(yq_ctx* c, const yq_step_cfg* cfg, int n_threads,
                      std::vector<double>& errs, std::vector<int>& iters,
                      std::vector<int>& rcs) {
  const int count = (int)c->subs.size();
  const double shift = cfg->lambda, tol = cfg->cg_tol;
  const int max_iter = cfg->cg_max_iter;
  if (tol <= 0 || max_iter <= 0) return fail(c, YQ_ERR_ARG, "tol and max_iter must be > 0");
  if (!(shift >= 0.0)) return fail(c, YQ_ERR_ARG, "shift must be >= 0");
  std::vector<double> bn2(count, 0.0), rel(count, 1.0);
  std::vector<int> total(count, 0);
  std::vector<char> done(count, 0);
  std::vector<int> st;
  int bad = 0, rc;
  // start: r = b, x = 0 (or x = 0 and done when b = 0)
  parallel_for(c, count, n_threads, st, [&](int i) -> int {
    yq_ctx* s = c->subs[i];
    const size_t dim = s->dim;
    TRY(ensure(s, s->vr, sizeof(double) * dim));
    TRY(ensure(s, s->vAp, sizeof(double) * dim));
    TRY(ensure(s, s->vp, sizeof(double) * dim));
    if (s->precond) TRY(ensure_diag(s));
    const double* b = s->F.as<double>();
    TRY(dot(s, b, b, dim, &bn2[i]));
    HIPCHK(s, hipMemsetAsync(s->vx.p, 0, sizeof(double) * dim, s->st));
    if (bn2[i] == 0.0) {
      done[i] = 1;
      rel[i] = 0.0;
    } else {
      HIPCHK(s, hipMemcpyAsync(s->vr.p, b, sizeof(double) * dim, hipMemcpyDeviceToDevice,
                               s->st));
    }
    return sync(s);
  });
  if ((rc = first_fail(st, &bad)) != YQ_OK) return from_sub(c, c->subs[bad], rc, bad);
  std::vector<yq::GroupJob> jobs(count);
  for (int outer = 0; outer < 30; ++outer) {
    bool any = false;
    for (int i = 0; i < count; ++i) {
      yq_ctx* s = c->subs[i];
      yq::GroupJob& j = jobs[i];
      j.w = &s->mixed;
      j.omean64 = s->omean.as<double>();
      j.diag64 = s->precond ? s->diagS.as<double>() : nullptr;
      j.b = s->vr.as<double>();
      j.x = s->vp.as<double>();
      j.active = !done[i] && total[i] < max_iter && !(rel[i] < tol);
      if (!j.active) continue;
      any = true;
      const double t = std::max(tol / rel[i], 1e-4);
      j.tol = t < 1.0 ? t : 0.5;
      j.max_iter = max_iter - total[i];
    }
    if (!any) break;
    rc = mixed_cg_group(c->group, jobs, shift, c->st, c->err);
    if (rc != YQ_OK) return fail(c, rc, c->err);
    // x += d and the FP64 residual of each machine that iterated
    parallel_for(c, count, n_threads, st, [&](int i) -> int {
      const yq::GroupJob& j = jobs[i];
      if (!j.active) return YQ_OK;
      yq_ctx* s = c->subs[i];
      const size_t dim = s->dim;
      total[i] += j.iters;
      double* x = s->vx.as<double>();
      double* r = s->vr.as<double>();
      double* Ax = s->vAp.as<double>();
      launch_axpy(1.0, s->vp.as<double>(), x, dim, s->st);
      LAUNCHCHK(s);
      TRY(apply64(s, shift, x, Ax));
      HIPCHK(s, hipMemcpyAsync(r, s->F.p, sizeof(double) * dim, hipMemcpyDeviceToDevice, s->st));
      launch_axpy(-1.0, Ax, r, dim, s->st);
      LAUNCHCHK(s);
      double rn2;
      TRY(dot(s, r, r, dim, &rn2));
      rel[i] = sqrt(rn2 / bn2[i]);
      if (rel[i] < tol) done[i] = 1;
      return YQ_OK;
    });
    if ((rc = first_fail(st, &bad)) != YQ_OK) return from_sub(c, c->subs[bad], rc, bad);
  }
  errs.assign(count, 0.0);
  iters.assign(count, 0);
  rcs.assign(count, YQ_OK);
  for (int i = 0; i < count; ++i) {
    errs[i] = rel[i];
    iters[i] = total[i];
    rcs[i] = rel[i] < tol ? YQ_OK : YQ_ERR_NOT_CONVERGED;
  }
  return YQ_OK;
}

// yq_multi_sr_step after sampling and E_loc with the grouped SR solve: build
// (per machine, host threads), grouped solve, update (host threads)
int multi_step_grouped(yq_ctx* c, const yq_step_cfg* cfg, yq_step_out* out, int n_threads,
                       float smp) {
  const int count = (int)c->subs.size();
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> st;
  int bad = 0, rc;
  parallel_for(c, count, n_threads, st, [&](int i) -> int {
    yq_ctx* s = c->subs[i];
    TRY(sr_build(s, cfg->solver, cfg->precision));
    TRY(ensure(s, s->vx, sizeof(double) * srlen(s)));
    return sync(s);
  });
  if ((rc = first_fail(st, &bad)) != YQ_OK) return from_sub(c, c->subs[bad], rc, bad);
  std::vector<double> errs;
  std::vector<int> iters, srcs;
  TRY(solve_mixed_group(c, cfg, n_threads, errs, iters, srcs));
  std::vector<double> nvs(count, 0.0);
  parallel_for(c, count, n_threads, st, [&](int i) -> int {
    yq_ctx* s = c->subs[i];
    // the solve's bookkeeping, as solve() keeps it
    TRY(ensure(s, s->xprev, sizeof(double) * srlen(s)));
    HIPCHK(s, hipMemcpyAsync(s->xprev.p, s->vx.p, sizeof(double) * srlen(s),
                             hipMemcpyDeviceToDevice, s->st));
    s->have_xprev = true;
    TRY(step_apply(s, cfg, &nvs[i]));
    TRY(sync(s));
    if (!std::isfinite(nvs[i]) || !std::isfinite(s->e_mean))
      return fail(s, YQ_ERR_NAN, "non-finite energy or update");
    return YQ_OK;
  });
  if ((rc = first_fail(st, &bad)) != YQ_OK) return from_sub(c, c->subs[bad], rc, bad);
  const float slv =
      std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  int worst = YQ_OK;
  for (int i = 0; i < count; ++i) {
    yq_ctx* s = c->subs[i];
    yq_step_out& o = out[i];
    o.energy = s->e_mean;
    o.energy_var = s->e_var;
    o.update_norm = nvs[i];
    o.cg_err = errs[i];
    o.cg_iters = iters[i];
    o.n_samples = s->ns;
    o.sample_ms = smp;
    o.solve_ms = slv;
    if (srcs[i] == YQ_ERR_NOT_CONVERGED) worst = srcs[i];
  }
  if (worst != YQ_OK) return fail(c, worst, "an iterative solve reached max_iter");
  return YQ_OK;
}
}  // namespace

extern "C" {

YQ_API int yq_multi_set_grouped(yq_ctx* c, int on) {
  TRY(need_multi(c));
  c->grouped = on != 0;
  return YQ_OK;
}

YQ_API int yq_multi_sr_step(yq_ctx* c, const yq_step_cfg* cfg, yq_step_out* out,
                            int n_threads) {
  TRY(need_multi(c));
  if (!cfg || !out) return fail(c, YQ_ERR_ARG, "null cfg/out");
  const int count = (int)c->subs.size();
  for (int i = 0; i < count; ++i) {
    yq_ctx* sub = c->subs[i];
    TRY(from_sub(c, sub, yq_sr_set_options(sub, cfg->precond, cfg->warm_start), i));
    if (cfg->randomize) TRY(from_sub(c, sub, yq_sampler_randomize(sub, cfg->nup), i));
  }
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  int rc = hipEventRecord(e0, c->st) == hipSuccess ? YQ_OK : YQ_ERR_HIP;
  if (rc == YQ_OK) rc = multi_sample(c, cfg->n_sweeps, cfg->n_discard);
  if (rc == YQ_OK) rc = hipEventRecord(e1, c->st) == hipSuccess ? YQ_OK : YQ_ERR_HIP;
  if (rc == YQ_OK) rc = sync(c);
  float smp = 0.f;
  if (rc == YQ_OK) (void)hipEventElapsedTime(&smp, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  TRY(rc);
  TRY(multi_eloc(c));
  if (c->grouped && cfg->solver == YQ_SOLVE_CG_FACTORED && cfg->precision == YQ_MIXED &&
      !cfg->warm_start)
    return multi_step_grouped(c, cfg, out, n_threads, smp);
  // per-machine SR on worker threads, costliest machines first
  const std::vector<int> ord = cost_order(c);
  std::vector<int> rcs(count, YQ_OK);
  std::atomic<int> next(0);
  auto work = [&]() {
    (void)hipSetDevice(c->device);
    for (int q = next++; q < count; q = next++) {
      const int i = ord[q];
      yq_ctx* sub = c->subs[i];
      hipEvent_t s0 = nullptr, s1 = nullptr;
      int r = hipEventCreate(&s0) == hipSuccess && hipEventCreate(&s1) == hipSuccess
                  ? YQ_OK : YQ_ERR_HIP;
      int it = 0, solve_rc = YQ_OK;
      double err = 0.0, nv = 0.0;
      if (r == YQ_OK) r = hipEventRecord(s0, sub->st) == hipSuccess ? YQ_OK : YQ_ERR_HIP;
      if (r == YQ_OK) r = step_update(sub, cfg, &nv, &err, &it, &solve_rc);
      if (r == YQ_OK) r = hipEventRecord(s1, sub->st) == hipSuccess ? YQ_OK : YQ_ERR_HIP;
      if (r == YQ_OK) r = sync(sub);
      float slv = 0.f;
      if (r == YQ_OK) (void)hipEventElapsedTime(&slv, s0, s1);
      if (s0) (void)hipEventDestroy(s0);
      if (s1) (void)hipEventDestroy(s1);
      yq_step_out& o = out[i];
      o.energy = sub->e_mean;
      o.energy_var = sub->e_var;
      o.update_norm = nv;
      o.cg_err = err;
      o.cg_iters = it;
      o.n_samples = sub->ns;
      o.sample_ms = smp;
      o.solve_ms = slv;
      if (r == YQ_OK && (!std::isfinite(nv) || !std::isfinite(sub->e_mean)))
        r = fail(sub, YQ_ERR_NAN, "non-finite energy or update");
      rcs[i] = r == YQ_OK ? solve_rc : r;
    }
  };
  const int nt = std::max(1, std::min(count, n_threads > 0 ? n_threads : 16));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  (void)hipSetDevice(c->device);
  int worst = YQ_OK;
  for (int i = 0; i < count; ++i)
    if (rcs[i] != YQ_OK && rcs[i] != YQ_ERR_NOT_CONVERGED)
      return from_sub(c, c->subs[i], rcs[i], i);
    else if (rcs[i] == YQ_ERR_NOT_CONVERGED)
      worst = rcs[i];
  if (worst != YQ_OK) return fail(c, worst, "an iterative solve reached max_iter");
  return YQ_OK;
}

}  // extern "C"
