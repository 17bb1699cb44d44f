// Batched Metropolis sampler for a real RBM on gfx950.
//
// One wavefront owns one Markov chain.  Hidden unit j lives in lane j % 64,
// register slot j / 64, so theta_j and logcosh(theta_j) stay in VGPRs for the
// whole launch; the spins of the chain sit in a 1-byte-per-site LDS strip.
// Past m = 512 a chain takes a whole block of G = 64 NW threads (unit j in
// thread j % G, slot j / G, still at most 8 slots; chain_group), and the
// log-ratio's tree over G lanes crosses the waves through LDS (group_sum).
// A proposal costs one Philox4x32-10 block (counter = the chain's proposal
// index), one coalesced read of the flipped column(s) of W (L2 resident),
// ceil(m/64) logcosh evaluations per lane, and a 6-level xor-butterfly that
// leaves the same log-ratio in every lane; acceptance is therefore uniform
// across the wave and the rank-1 theta update is a register select.
//
// Semantics: LocalSweeper.hpp:32-53 (single flips, k ~ U{0..n-1},
// p = min(1, exp(2 beta Re dlogpsi)), accept iff u < p) and
// SwapSweeper.hpp:28-56 (k, l ~ U{0..n-1} independently, skip if parallel),
// scheduled as SamplerMT::sample, SamplerMT.hpp:180-204 (n_discard sweeps,
// then n_sweeps sweeps each followed by recording every chain).
// Arithmetic: yq_math.h; the log-ratio is the lane-strided sum followed by a
// stride-32..1 tree, which the CPU replay reproduces exactly.
#include <algorithm>

#include "yq_internal.h"
#include "yq_math.h"

namespace yq {
namespace {

// theta_j = (sum_i +/-W_ji, index order, fp64) + b_j, then rounded to R, and
// lc_j = logcosh(theta_j), for the hidden units of thread gl of a G-thread
// chain group
template <typename R, int RMAX, int G = kWave>
__device__ __forceinline__ void refresh_theta(const SampleArgs& A, const int8_t* sig, int gl,
                                              R (&th)[RMAX], R (&lc)[RMAX]) {
  const int n = A.n, m = A.m;
  double accd[RMAX];
#pragma unroll
  for (int r = 0; r < RMAX; ++r) accd[r] = 0.0;
  for (int i = 0; i < n; ++i) {
    const bool up = sig[i] > 0;
    const double* wc = A.W64 + (size_t)i * m;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = gl + G * r;
      if (j < m) {
        const double w = wc[j];
        accd[r] = up ? accd[r] + w : accd[r] - w;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    const int j = gl + G * r;
    if (j < m) {
      th[r] = (R)(accd[r] + A.b64[j]);
      lc[r] = Math<R>::logcosh(th[r]);
    } else {
      th[r] = (R)0;
      lc[r] = (R)0;
    }
  }
}

// One sweep (A.nprop proposals) of one chain at inverse temperature beta;
// returns the accepted count.  ev is the chain's proposal counter.  gl is
// the thread's index in its G-thread chain group; red and par are
// group_sum's LDS exchange and parity (unused when G = 64).  With G > 64
// every wave of the block stores each accepted flip itself (the same byte),
// so every wave reads the spins it wrote; the barrier inside group_sum
// orders each proposal's reads before any wave's flip of that proposal.
template <typename R, int KIND, int RMAX, int G = kWave>
__device__ __forceinline__ uint64_t sweep_chain(const SampleArgs& A, int chain, int8_t* sig,
                                                int gl, R (&th)[RMAX], R (&lc)[RMAX],
                                                R (&tt)[RMAX], R (&lt)[RMAX], uint64_t& ev,
                                                R beta, R* red = nullptr, int* par = nullptr) {
  const int n = A.n, m = A.m;
  const R* Wr = static_cast<const R*>(A.Wr);
  const R* ar = static_cast<const R*>(A.ar);
  uint64_t nacc = 0;
  // the chain index and proposal counter are the same in every lane: read
  // them from lane 0 so that the Philox block runs on the scalar unit
  const uint32_t stream = __builtin_amdgcn_readfirstlane((uint32_t)chain);
  for (int p = 0; p < A.nprop; ++p) {
    const uint64_t evu =
        ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ev >> 32)) << 32) |
        (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)ev);
    ev = evu + 1;
    const u32x4 x = draw(A.seed, evu, stream, kDomSweep);
    const int k = __builtin_amdgcn_readfirstlane((int)index_from(x.x, (uint32_t)n));
    const int8_t sk = sig[k];
    const R twok = sk > 0 ? (R)2 : (R)-2;
    const R* wk = Wr + (size_t)k * m;
    R cst = ((R)-2 * ar[k]) * (R)sk;
    int l = 0;
    int8_t sl = 0;
    R twol = (R)0;
    const R* wl = wk;
    if (KIND == 1) {
      l = __builtin_amdgcn_readfirstlane((int)index_from(x.y, (uint32_t)n));
      sl = sig[l];
      if (sk == sl) continue;  // SwapSweeper.hpp:44-47, uniform branch
      twol = sl > 0 ? (R)2 : (R)-2;
      wl = Wr + (size_t)l * m;
      cst = cst - ((R)2 * ar[l]) * (R)sl;
    }
    R part = (R)0;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = gl + G * r;
      if (j < m) {
        R t = th[r] - twok * wk[j];
        if (KIND == 1) t = t - twol * wl[j];
        tt[r] = t;
        lt[r] = Math<R>::logcosh(t);
        part = part + (lt[r] - lc[r]);
      }
    }
    if constexpr (G > kWave) {
      part = group_sum<R, G>(part, red, *par);
    } else {
      int p0 = 0;
      part = group_sum<R, G>(part, nullptr, p0);
    }
    const R delta = cst + part;
    bool ok;
    if (sizeof(R) == 4)
      ok = accept_f((float)beta, (float)delta, u01_float(x.z));
    else
      ok = accept_d((double)beta, (double)delta, u01_double(x.z, x.w));
    if (ok) {
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int j = gl + G * r;
        if (j < m) {
          th[r] = tt[r];
          lc[r] = lt[r];
        }
      }
      sig[k] = (int8_t)-sk;  // every thread stores the same byte
      if (KIND == 1) sig[l] = (int8_t)-sl;
      ++nacc;
    }
  }
  return nacc;
}

// sample row `row`: spins and theta (mixed precision: theta refreshed from
// the fp64 parameters first, fresh is then set)
template <typename R, int RMAX, int G = kWave>
__device__ __forceinline__ void record_row(const SampleArgs& A, size_t row, const int8_t* sig,
                                           int gl, R (&th)[RMAX], R (&lc)[RMAX],
                                           bool& fresh) {
  const int n = A.n, m = A.m;
  for (int i = gl; i < n; i += G) A.out_sigma[row * n + i] = sig[i];
  // mixed precision with the MFMA theta pass: the pass that follows the
  // launch records theta refreshed from FP64 (one sweep per launch)
  if (A.mixed && A.theta_in) return;
  if (A.mixed) {
    refresh_theta<R, RMAX, G>(A, sig, gl, th, lc);
    fresh = true;
  }
  R* ot = static_cast<R*>(A.out_theta) + row * m;
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    const int j = gl + G * r;
    if (j < m) ot[j] = th[r];
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One chain of one machine, run by a G-thread chain group (one wavefront,
// or the whole block when G > 64): sig is the group's LDS spin strip, gl the
// thread's index in the group, red group_sum's LDS exchange.  RMAX >=
// ceil(m/G); larger values only add predicated-off slots, so the arithmetic
// (and the replay) does not depend on it.
template <typename R, int KIND, int RMAX, int G = kWave>
__device__ void run_chain(const SampleArgs& A, int chain, int8_t* sig, int gl,
                          R* red = nullptr) {
  const int n = A.n;
  for (int i = gl; i < n; i += G) sig[i] = A.sigma[(size_t)chain * n + i];
  if (G > kWave)
    __syncthreads();
  else
    wave_sync();
  int par = 0;
  R th[RMAX], lc[RMAX], tt[RMAX], lt[RMAX];
  uint64_t ev = A.ctr[chain];
  uint64_t nacc = 0;
  const R beta = (R)A.beta;
  if (A.theta_in) {  // from the MFMA theta pass (the same bits as refresh_theta)
    const R* ti = static_cast<const R*>(A.theta_in) + (size_t)chain * A.m;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = gl + G * r;
      th[r] = j < A.m ? ti[j] : (R)0;
      lc[r] = j < A.m ? Math<R>::logcosh(th[r]) : (R)0;
    }
  } else {
    refresh_theta<R, RMAX, G>(A, sig, gl, th, lc);
  }
  bool fresh = true;
  const int total = A.n_discard + A.n_sweeps;
  for (int sw = 0; sw < total; ++sw) {
    if (A.mixed && !fresh) refresh_theta<R, RMAX, G>(A, sig, gl, th, lc);
    fresh = false;
    nacc += sweep_chain<R, KIND, RMAX, G>(A, chain, sig, gl, th, lc, tt, lt, ev, beta, red, &par);
    if (sw >= A.n_discard)
      record_row<R, RMAX, G>(A, (size_t)(sw - A.n_discard) * A.C + chain, sig, gl, th, lc,
                             fresh);
  }
  for (int i = gl; i < n; i += G) A.sigma[(size_t)chain * n + i] = sig[i];
  if (gl == 0) {
    A.ctr[chain] = ev;
    A.acc[chain] += nacc;
  }
  if (G == kWave) wave_sync();  // the strip is reused by the wave's next chain
}

// Parallel tempering (SamplerMT.hpp:44-60, 112-204 with nTmps = T > 1): one
// block per chain group g.  Chain g * T + t is temperature slot t at
// beta_t = (T - t) / T with its own Philox stream and counter.  The block's
// NW waves (the largest of 4, 2 dividing T) take slots w, w + NW, ... in turn;
// a slot's state (spins and theta) lives in LDS storage idx[t], so a replica
// swap is an exchange of two storage indices.  After every sweep of all slots
// the swaps run as in mixChains (:120-153): pairs (0,1), (2,3), ..., then
// (1,2), ..., (T-3, T-2); the pair (t, t+1) swaps iff
// u < exp((beta_{t+1} - beta_t) 2 (lp_t - lp_{t+1})), lp = Re log psi =
// a . sigma + sum_j logcosh(theta_j) (lane-strided sums and an xor butterfly),
// u from slot t's stream at its proposal counter after the sweep (a separate
// Philox domain).  Slot 0 (beta = 1) is recorded after every sweep past the
// discard phase (appendData), row (sweep * C/T + g).
template <typename R, int KIND, int RMAX>
__global__ void __launch_bounds__(256) k_sample_pt(const SampleArgs A) {
  extern __shared__ unsigned char lds_raw[];
  const int T = A.ntemps, n = A.n, m = A.m, npad = A.npad;
  const int NW = (T % 4 == 0) ? 4 : 2;
  // dynamic LDS (pt_lds_bytes): lp, counters, accept counts and storage index
  // per slot, spin strips and theta storage
  R* lpv = reinterpret_cast<R*>(lds_raw);                           // 16
  uint64_t* evs = reinterpret_cast<uint64_t*>(lds_raw + 128);       // 16
  uint64_t* nac = reinterpret_cast<uint64_t*>(lds_raw + 256);       // 16
  int* idx = reinterpret_cast<int*>(lds_raw + 384);                 // 16
  int8_t* strips = reinterpret_cast<int8_t*>(lds_raw + 448);        // T x npad
  R* ths = reinterpret_cast<R*>(lds_raw + 448 + ((T * npad + 15) & ~15));  // T x m
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int G = kWave;  // one wave per temperature slot
  const int gl = lane;
  const int g = blockIdx.x;
  const int groups = A.C / T;
  const R* ar = static_cast<const R*>(A.ar);
  R th[RMAX], lc[RMAX], tt[RMAX], lt[RMAX];
  // initial states: slot t in storage t, theta refreshed from fp64
  for (int t = w; t < T; t += NW) {
    const int chain = g * T + t;
    int8_t* sig = strips + t * npad;
    for (int i = lane; i < n; i += kWave) sig[i] = A.sigma[(size_t)chain * n + i];
    wave_sync();
    refresh_theta<R, RMAX>(A, sig, lane, th, lc);
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = gl + G * r;
      if (j < m) ths[t * m + j] = th[r];
    }
    if (lane == 0) {
      evs[t] = A.ctr[chain];
      nac[t] = 0;
      idx[t] = t;
    }
  }
  __syncthreads();
  const int total = A.n_discard + A.n_sweeps;
  for (int sw = 0; sw < total; ++sw) {
    for (int t = w; t < T; t += NW) {
      const int chain = g * T + t;
      const int st = idx[t];
      int8_t* sig = strips + st * npad;
      if (A.mixed && sw > 0) {
        refresh_theta<R, RMAX>(A, sig, lane, th, lc);
      } else {
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
          const int j = gl + G * r;
          th[r] = j < m ? ths[st * m + j] : (R)0;
          lc[r] = j < m ? Math<R>::logcosh(th[r]) : (R)0;
        }
      }
      uint64_t ev = evs[t];
      const R beta = (R)((double)(T - t) / T);
      const uint64_t acc = sweep_chain<R, KIND, RMAX>(A, chain, sig, lane, th, lc, tt, lt, ev,
                                                      beta);
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int j = gl + G * r;
        if (j < m) ths[st * m + j] = th[r];
      }
      // lp = a . sigma + sum_j logcosh(theta_j)
      R sa = (R)0, sl = (R)0;
      for (int i = lane; i < n; i += kWave) sa = sa + ar[i] * (R)sig[i];
#pragma unroll
      for (int r = 0; r < RMAX; ++r)
        if (gl + G * r < m) sl = sl + lc[r];
#pragma unroll
      for (int s = 32; s >= 1; s >>= 1) {
        sa = sa + __shfl_xor(sa, s);
        sl = sl + __shfl_xor(sl, s);
      }
      if (lane == 0) {
        lpv[t] = sa + sl;
        evs[t] = ev;
        nac[t] += acc;
      }
      wave_sync();
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // the two swap passes, pair by pair in index order
      for (int pass = 0; pass < 2; ++pass)
        for (int t = pass; t + 1 < T && (pass == 0 || t + 1 <= T - 2); t += 2) {
          const u32x4 x = draw(A.seed, evs[t], (uint32_t)(g * T + t), kDomMix);
          const double dbeta = (double)(T - t - 1) / T - (double)(T - t) / T;
          const R dl = lpv[t] - lpv[t + 1];
          const bool d = sizeof(R) == 4 ? accept_f((float)dbeta, (float)dl, u01_float(x.z))
                                        : accept_d(dbeta, (double)dl, u01_double(x.z, x.w));
          if (d) {
            const int i0 = idx[t];
            idx[t] = idx[t + 1];
            idx[t + 1] = i0;
            const R l0 = lpv[t];
            lpv[t] = lpv[t + 1];
            lpv[t + 1] = l0;
          }
        }
    }
    __syncthreads();
    if (w == 0 && sw >= A.n_discard) {  // slot 0
      const int st = idx[0];
      const int8_t* sig = strips + st * npad;
      const size_t row = (size_t)(sw - A.n_discard) * groups + g;
      for (int i = lane; i < n; i += kWave) A.out_sigma[row * n + i] = sig[i];
      if (A.mixed) refresh_theta<R, RMAX>(A, sig, lane, th, lc);
      R* ot = static_cast<R*>(A.out_theta) + row * m;
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int j = gl + G * r;
        if (j < m) ot[j] = A.mixed ? th[r] : ths[st * m + j];
      }
    }
  }
  __syncthreads();
  for (int t = w; t < T; t += NW) {
    const int chain = g * T + t;
    const int8_t* sig = strips + idx[t] * npad;
    for (int i = lane; i < n; i += kWave) A.sigma[(size_t)chain * n + i] = sig[i];
    if (lane == 0) {
      A.ctr[chain] = evs[t];
      A.acc[chain] += nac[t];
    }
  }
}

template <typename R, int KIND, int RMAX>
__global__ void __launch_bounds__(256) k_sample(const SampleArgs A) {
  extern __shared__ int8_t lds_sig[];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int chain = blockIdx.x * 4 + wave;
  if (chain >= A.C) return;  // wave-uniform; no block barrier follows
  run_chain<R, KIND, RMAX>(A, chain, lds_sig + wave * A.npad, lane);
}

// m > 512: one chain per block of G threads.  LDS: the spin strip (npad
// bytes), then group_sum's exchange (2 G values of R).
template <typename R, int KIND, int RMAX, int G>
__global__ void __launch_bounds__(G) k_sample_wide(const SampleArgs A) {
  extern __shared__ __align__(16) unsigned char lds_w[];
  R* red = reinterpret_cast<R*>(lds_w + ((A.npad + 15) & ~15));
  run_chain<R, KIND, RMAX, G>(A, blockIdx.x, reinterpret_cast<int8_t*>(lds_w), threadIdx.x, red);
}

// Many machines in one persistent launch: every wave walks the list of
// (machine, chain) items with a grid-sized stride.  The list is ordered by
// decreasing cost (n * m * sweep factor), so the strided walk hands every
// wave a similar amount of work; a chain's result does not depend on which
// wave runs it.
template <typename R, int KIND>
__global__ void __launch_bounds__(256)
    k_sample_multi(const SampleArgs* __restrict__ inst, const int2* __restrict__ items,
                   int n_items, int npad) {
  extern __shared__ int8_t lds_sig[];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * 4;
  for (int q = blockIdx.x * 4 + wave; q < n_items; q += nw) {
    const int2 it = items[q];
    const SampleArgs A = inst[it.x];
    run_chain<R, KIND, kWaveHidden / kWave>(A, it.y, lds_sig + wave * npad, lane);
  }
}

template <typename R, int KIND, int G>
void launch_wide(const SampleArgs& a, hipStream_t st) {
  const size_t lds = (size_t)((a.npad + 15) & ~15) + 2 * G * sizeof(R);
  // 6 slots when they suffice, else 8 (at most a third of the slots idle)
  if (a.m <= 6 * G)
    hipLaunchKernelGGL((k_sample_wide<R, KIND, 6, G>), dim3(a.C), dim3(G), lds, st, a);
  else
    hipLaunchKernelGGL((k_sample_wide<R, KIND, 8, G>), dim3(a.C), dim3(G), lds, st, a);
}

template <typename R, int KIND>
void launch_r(const SampleArgs& a, hipStream_t st) {
  switch (chain_group(a.m)) {
    case kWave: break;
    case 128: launch_wide<R, KIND, 128>(a, st); return;
    case 256: launch_wide<R, KIND, 256>(a, st); return;
    case 512: launch_wide<R, KIND, 512>(a, st); return;
    default: launch_wide<R, KIND, 1024>(a, st); return;
  }
  const dim3 grid((a.C + 3) / 4), block(256);
  const size_t lds = (size_t)4 * a.npad;
  // exactly ceil(m/64) register slots, so no slot is computed for nothing
  switch ((a.m + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL((k_sample<R, KIND, 1>), grid, block, lds, st, a); break;
    case 2: hipLaunchKernelGGL((k_sample<R, KIND, 2>), grid, block, lds, st, a); break;
    case 3: hipLaunchKernelGGL((k_sample<R, KIND, 3>), grid, block, lds, st, a); break;
    case 4: hipLaunchKernelGGL((k_sample<R, KIND, 4>), grid, block, lds, st, a); break;
    case 5: hipLaunchKernelGGL((k_sample<R, KIND, 5>), grid, block, lds, st, a); break;
    case 6: hipLaunchKernelGGL((k_sample<R, KIND, 6>), grid, block, lds, st, a); break;
    case 7: hipLaunchKernelGGL((k_sample<R, KIND, 7>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((k_sample<R, KIND, 8>), grid, block, lds, st, a); break;
  }
}

// theta = W sigma + b for `rows` configurations on v_mfma_f64_16x16x4f64
// (lane l holds A[row l%16][k l/16], B[k l/16][col l%16], and returns
// D[(l/16) + 4q][l%16]).  A = the spins (+/-1, 0 past n), B = W (W_ji at
// W64[i m + j]), k = the spin index i.  Every product is exact and the f64
// MFMA adds its four k terms in order, one rounding each, onto the running
// sum, so the result is bit for bit the sequential FP64 sum over i of the
// other theta paths (refresh_theta, the oracle's replay), then + b_j and the
// rounding to O.  Block: 4 waves x 16 configurations, 64 hidden units (4
// column tiles per wave); the loads of several k-steps are in flight
// together, with clamped addresses instead of branches.
typedef double double4v __attribute__((ext_vector_type(4)));

template <typename O>
__global__ void __launch_bounds__(256)
    k_theta_mfma(const double* __restrict__ W64, const double* __restrict__ b64,
                 const int8_t* __restrict__ sigma, int rows, int n, int m,
                 O* __restrict__ out) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int li = l & 15, lg = l >> 4;
  const int s0 = (blockIdx.x * 4 + wave) * 16;
  const int j0 = blockIdx.y * 64;
  if (s0 >= rows) return;  // wave-uniform; no barrier in this kernel
  const int sa = min(s0 + li, rows - 1);  // this lane's A row (clamped)
  const int8_t* srow = sigma + (size_t)sa * n;
  int jc[4];
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) jc[ct] = min(j0 + ct * 16 + li, m - 1);
  double4v acc[4];
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) acc[ct] = (double4v){0.0, 0.0, 0.0, 0.0};
  const int ksteps = (n + 3) / 4;
  constexpr int kU = 5;  // k-steps whose loads are issued together
  for (int k0 = 0; k0 < ksteps; k0 += kU) {
    double a[kU], b[kU][4];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = (k0 + u) * 4 + lg;
      const int ic = min(i, n - 1);
      const int8_t g = srow[ic];
      a[u] = (i < n && k0 + u < ksteps) ? (g > 0 ? 1.0 : -1.0) : 0.0;
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) b[u][ct] = W64[(size_t)ic * m + jc[ct]];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (k0 + u < ksteps)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
          acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u][ct], acc[ct], 0, 0, 0);
  }
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) {
    const int j = j0 + ct * 16 + li;
    if (j >= m) continue;
    const double bj = b64[j];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int s = s0 + lg + 4 * q;
      if (s < rows) out[(size_t)s * m + j] = (O)(acc[ct][q] + bj);
    }
  }
}

void theta_pass(const double* W64, const double* b64, const int8_t* sigma, int rows, int n,
                int m, void* out, int out_float, hipStream_t st) {
  const dim3 grid((unsigned)((rows + 63) / 64), (unsigned)((m + 63) / 64)), block(256);
  ProfGuard g("theta_mfma", st);
  if (out_float)
    hipLaunchKernelGGL(k_theta_mfma<float>, grid, block, 0, st, W64, b64, sigma, rows, n, m,
                       static_cast<float*>(out));
  else
    hipLaunchKernelGGL(k_theta_mfma<double>, grid, block, 0, st, W64, b64, sigma, rows, n, m,
                       static_cast<double*>(out));
}

// logRatio(vector) of RBMState.hpp:108-126 for ns configurations, one wave
// each: log psi(sigma with the listed sites flipped) - log psi(sigma).  Row s
// of sites lists up to k sites (entries < 0 are skipped); theta (ns x m,
// interleaved complex when CPLX) is W sigma + b of the unflipped
// configuration.  Per hidden unit t_j = theta_j - sum over the list of
// 2 sigma_e W_je (list order); the wave sums logcosh(t_j) - logcosh(theta_j)
// lane-strided with an xor butterfly, then adds the visible terms
// -2 a_e sigma_e in list order.
template <bool CPLX>
__global__ void __launch_bounds__(256)
    k_log_ratio_vec(const double* __restrict__ W, const double* __restrict__ a,
                    const int8_t* __restrict__ sigma, const double* __restrict__ theta,
                    const int* __restrict__ sites, int ns, int n, int m, int k,
                    double* __restrict__ out) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= ns) return;  // wave-uniform; no barrier in this kernel
  const int8_t* sg = sigma + (size_t)s * n;
  const int* v = sites + (size_t)s * k;
  double pr = 0.0, pi = 0.0;
  for (int j = lane; j < m; j += kWave) {
    if (CPLX) {
      const double tr0 = theta[2 * ((size_t)s * m + j)], ti0 = theta[2 * ((size_t)s * m + j) + 1];
      double tr = tr0, ti = ti0;
      for (int e = 0; e < k; ++e) {
        const int q = v[e];
        if (q < 0) continue;
        const double two = sg[q] > 0 ? 2.0 : -2.0;
        tr -= two * W[2 * ((size_t)q * m + j)];
        ti -= two * W[2 * ((size_t)q * m + j) + 1];
      }
      double r1, i1, r0, i0;
      logcosh_c_d(tr, ti, &r1, &i1);
      logcosh_c_d(tr0, ti0, &r0, &i0);
      pr += r1 - r0;
      pi += i1 - i0;
    } else {
      const double t0 = theta[(size_t)s * m + j];
      double t = t0;
      for (int e = 0; e < k; ++e) {
        const int q = v[e];
        if (q < 0) continue;
        t -= (sg[q] > 0 ? 2.0 : -2.0) * W[(size_t)q * m + j];
      }
      pr += logcosh_d(t) - logcosh_d(t0);
    }
  }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    pr += __shfl_xor(pr, w);
    if (CPLX) pi += __shfl_xor(pi, w);
  }
  if (lane == 0) {
    double vr = 0.0, vi = 0.0;
    for (int e = 0; e < k; ++e) {
      const int q = v[e];
      if (q < 0) continue;
      const double two = sg[q] > 0 ? 2.0 : -2.0;
      vr -= two * a[CPLX ? 2 * q : q];
      if (CPLX) vi -= two * a[2 * q + 1];
    }
    if (CPLX) {
      out[2 * s] = vr + pr;
      out[2 * s + 1] = vi + pi;
    } else {
      out[s] = vr + pr;
    }
  }
}

}  // namespace

void launch_log_ratio_vec(const double* W, const double* a, const int8_t* sigma,
                          const double* theta, const int* sites, int ns, int n, int m, int k,
                          int cplx, double* out, hipStream_t st) {
  const dim3 grid((unsigned)((ns + 3) / 4)), block(256);
  if (cplx)
    hipLaunchKernelGGL(k_log_ratio_vec<true>, grid, block, 0, st, W, a, sigma, theta, sites, ns,
                       n, m, k, out);
  else
    hipLaunchKernelGGL(k_log_ratio_vec<false>, grid, block, 0, st, W, a, sigma, theta, sites, ns,
                       n, m, k, out);
}

void launch_sample_multi(const SampleArgs* inst_dev, const int2* items, int n_items,
                         int npad, int kind, int mixed, int blocks, hipStream_t st) {
  ProfGuard g("sample_multi", st);
  const size_t lds = (size_t)4 * npad;
  if (blocks <= 0) {  // persistent grid: what the chip keeps resident, at most one wave per item
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (mixed)
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, kind == 1 ? (const void*)k_sample_multi<float, 1> : (const void*)k_sample_multi<float, 0>,
          256, lds);
    else
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, kind == 1 ? (const void*)k_sample_multi<double, 1> : (const void*)k_sample_multi<double, 0>,
          256, lds);
    blocks = std::max(1, cus * std::max(per_cu, 1));
    blocks = std::min(blocks, (n_items + 3) / 4);
  }
  if (mixed) {
    if (kind == 1)
      hipLaunchKernelGGL((k_sample_multi<float, 1>), dim3(blocks), dim3(256), lds, st, inst_dev,
                         items, n_items, npad);
    else
      hipLaunchKernelGGL((k_sample_multi<float, 0>), dim3(blocks), dim3(256), lds, st, inst_dev,
                         items, n_items, npad);
  } else {
    if (kind == 1)
      hipLaunchKernelGGL((k_sample_multi<double, 1>), dim3(blocks), dim3(256), lds, st, inst_dev,
                         items, n_items, npad);
    else
      hipLaunchKernelGGL((k_sample_multi<double, 0>), dim3(blocks), dim3(256), lds, st, inst_dev,
                         items, n_items, npad);
  }
}

size_t pt_lds_bytes(const SampleArgs& a) {
  // complex machines keep Re and Im of theta
  const size_t rs = (a.mixed ? sizeof(float) : sizeof(double)) * (a.cplx ? 2 : 1);
  return 448 + (((size_t)a.ntemps * a.npad + 15) & ~(size_t)15) + (size_t)a.ntemps * a.m * rs;
}

template <typename R, int KIND>
void launch_pt(const SampleArgs& a, hipStream_t st) {
  const dim3 grid(a.C / a.ntemps), block(kWave * (a.ntemps % 4 == 0 ? 4 : 2));
  const size_t lds = pt_lds_bytes(a);
  switch ((a.m + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL((k_sample_pt<R, KIND, 1>), grid, block, lds, st, a); break;
    case 2: hipLaunchKernelGGL((k_sample_pt<R, KIND, 2>), grid, block, lds, st, a); break;
    case 3: hipLaunchKernelGGL((k_sample_pt<R, KIND, 3>), grid, block, lds, st, a); break;
    case 4: hipLaunchKernelGGL((k_sample_pt<R, KIND, 4>), grid, block, lds, st, a); break;
    case 5: hipLaunchKernelGGL((k_sample_pt<R, KIND, 5>), grid, block, lds, st, a); break;
    case 6: hipLaunchKernelGGL((k_sample_pt<R, KIND, 6>), grid, block, lds, st, a); break;
    case 7: hipLaunchKernelGGL((k_sample_pt<R, KIND, 7>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((k_sample_pt<R, KIND, 8>), grid, block, lds, st, a); break;
  }
}

void launch_sample(const SampleArgs& a, hipStream_t st) {
  ProfGuard g("sample", st);
  if (a.ntemps > 1) {
    if (a.mixed) {
      if (a.kind == 1) launch_pt<float, 1>(a, st);
      else launch_pt<float, 0>(a, st);
    } else {
      if (a.kind == 1) launch_pt<double, 1>(a, st);
      else launch_pt<double, 0>(a, st);
    }
    return;
  }
  auto run = [&](const SampleArgs& x) {
    if (x.mixed) {
      if (x.kind == 1) launch_r<float, 1>(x, st);
      else launch_r<float, 0>(x, st);
    } else {
      if (x.kind == 1) launch_r<double, 1>(x, st);
      else launch_r<double, 0>(x, st);
    }
  };
  if (!a.theta_buf) {
    run(a);
    return;
  }
  const int fl = a.mixed ? 1 : 0;
  const size_t rs = a.mixed ? sizeof(float) : sizeof(double);
  // theta of the chains' current spins
  theta_pass(a.W64, a.b64, a.sigma, a.C, a.n, a.m, a.theta_buf, fl, st);
  if (!a.mixed) {  // FP64 chains keep their theta by rank-1 updates: one launch
    SampleArgs x = a;
    x.theta_in = a.theta_buf;
    run(x);
    return;
  }
  // float chains restart every sweep from theta refreshed from FP64: one
  // launch per sweep, each followed by the theta pass of the spins it left
  // (the recorded rows, or the chain state after a discard sweep), which is
  // both the record and the next sweep's start
  const void* tin = a.theta_buf;
  for (int sw = 0; sw < a.n_discard + a.n_sweeps; ++sw) {
    SampleArgs x = a;
    x.theta_in = tin;
    if (sw < a.n_discard) {
      x.n_discard = 1;
      x.n_sweeps = 0;
      run(x);
      theta_pass(a.W64, a.b64, a.sigma, a.C, a.n, a.m, a.theta_buf, fl, st);
      tin = a.theta_buf;
    } else {
      const size_t r = (size_t)(sw - a.n_discard) * a.C;
      x.n_discard = 0;
      x.n_sweeps = 1;
      x.out_sigma = a.out_sigma + r * a.n;
      x.out_theta = static_cast<char*>(a.out_theta) + r * a.m * rs;
      run(x);
      theta_pass(a.W64, a.b64, x.out_sigma, a.C, a.n, a.m, x.out_theta, fl, st);
      tin = x.out_theta;
    }
  }
}

void launch_theta_batch(const double* W64, const double* b64,
                        const int8_t* sigma, int ns, int n, int m, void* out,
                        int out_float, hipStream_t st) {
  theta_pass(W64, b64, sigma, ns, n, m, out, out_float, st);
}

}  // namespace yq
