// Batched Metropolis sampler for a real RBM on gfx950.
//
// One wavefront owns one Markov chain.  Hidden unit j lives in lane j % 64,
// register slot j / 64, so theta_j and logcosh(theta_j) stay in VGPRs for the
// whole launch; the spins of the chain sit in a 1-byte-per-site LDS strip.
// A proposal costs one Philox4x32-10 block (counter = the chain's proposal
// index), one coalesced read of the flipped column(s) of W (L2 resident),
// ceil(m/64) logcosh evaluations per lane, and a 6-level xor-butterfly that
// leaves the same log-ratio in every lane; acceptance is therefore uniform
// across the wave and the rank-1 theta update is a register select.
//
// Semantics: LocalSweeper.hpp:32-53 (single flips, k ~ U{0..n-1},
// p = min(1, exp(2 beta Re dlogpsi)), accept iff u < p) and
// SwapSweeper.hpp:28-56 (k, l ~ U{0..n-1} independently, skip if parallel),
// scheduled as SamplerMT::sample, SamplerMT.hpp:293-317 (n_discard sweeps,
// then n_sweeps sweeps each followed by recording every chain).
// Arithmetic: yq_math.h; the log-ratio is the lane-strided sum followed by a
// stride-32..1 tree, which the CPU replay reproduces exactly.
#include <algorithm>

#include "yq_internal.h"
#include "yq_math.h"

namespace yq {
namespace {

// One chain of one machine, run by one wavefront: sig is the wave's LDS spin
// strip.  RMAX >= ceil(m/64); larger values only add predicated-off slots,
// so the arithmetic (and the replay) does not depend on it.
template <typename R, int KIND, int RMAX>
__device__ void run_chain(const SampleArgs& A, int chain, int8_t* sig, int lane) {
  const int n = A.n, m = A.m;
  for (int i = lane; i < n; i += kWave) sig[i] = A.sigma[(size_t)chain * n + i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const R* Wr = static_cast<const R*>(A.Wr);
  const R* ar = static_cast<const R*>(A.ar);
  R th[RMAX], lc[RMAX], tt[RMAX], lt[RMAX];

  // theta_j = (sum_i +/-W_ji, index order, fp64) + b_j, then rounded to R
  auto refresh = [&]() {
    double accd[RMAX];
#pragma unroll
    for (int r = 0; r < RMAX; ++r) accd[r] = 0.0;
    for (int i = 0; i < n; ++i) {
      const bool up = sig[i] > 0;
      const double* wc = A.W64 + (size_t)i * m;
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int j = lane + kWave * r;
        if (j < m) {
          const double w = wc[j];
          accd[r] = up ? accd[r] + w : accd[r] - w;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = lane + kWave * r;
      if (j < m) {
        th[r] = (R)(accd[r] + A.b64[j]);
        lc[r] = Math<R>::logcosh(th[r]);
      } else {
        th[r] = (R)0;
        lc[r] = (R)0;
      }
    }
  };

  uint64_t ev = A.ctr[chain];
  uint64_t nacc = 0;
  const R beta = (R)A.beta;
  refresh();
  bool fresh = true;
  const int total = A.n_discard + A.n_sweeps;
  for (int sw = 0; sw < total; ++sw) {
    if (A.mixed && !fresh) refresh();
    fresh = false;
    for (int p = 0; p < A.nprop; ++p) {
      const u32x4 x = draw(A.seed, ev++, (uint32_t)chain, kDomSweep);
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
        const int j = lane + kWave * r;
        if (j < m) {
          R t = th[r] - twok * wk[j];
          if (KIND == 1) t = t - twol * wl[j];
          tt[r] = t;
          lt[r] = Math<R>::logcosh(t);
          part = part + (lt[r] - lc[r]);
        }
      }
#pragma unroll
      for (int s = 32; s >= 1; s >>= 1) part = part + __shfl_xor(part, s);
      const R delta = cst + part;
      bool ok;
      if (sizeof(R) == 4)
        ok = accept_f((float)beta, (float)delta, u01_float(x.z));
      else
        ok = accept_d((double)beta, (double)delta, u01_double(x.z, x.w));
      if (ok) {
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
          const int j = lane + kWave * r;
          if (j < m) {
            th[r] = tt[r];
            lc[r] = lt[r];
          }
        }
        sig[k] = (int8_t)-sk;  // every lane stores the same byte
        if (KIND == 1) sig[l] = (int8_t)-sl;
        ++nacc;
      }
    }
    if (sw >= A.n_discard) {
      const size_t row = (size_t)(sw - A.n_discard) * A.C + chain;
      for (int i = lane; i < n; i += kWave) A.out_sigma[row * n + i] = sig[i];
      if (A.mixed) {
        refresh();
        fresh = true;
      }
      R* ot = static_cast<R*>(A.out_theta) + row * m;
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int j = lane + kWave * r;
        if (j < m) ot[j] = th[r];
      }
    }
  }
  for (int i = lane; i < n; i += kWave) A.sigma[(size_t)chain * n + i] = sig[i];
  if (lane == 0) {
    A.ctr[chain] = ev;
    A.acc[chain] += nacc;
  }
  // the strip is reused by the wave's next chain
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
    run_chain<R, KIND, kMaxHidden / kWave>(A, it.y, lds_sig + wave * npad, lane);
  }
}

template <typename R, int KIND>
void launch_r(const SampleArgs& a, hipStream_t st) {
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

// theta for a batch, one thread per (sample, hidden unit)
template <typename O>
__global__ void __launch_bounds__(256)
    k_theta_batch(const double* __restrict__ W64, const double* __restrict__ b64,
                  const int8_t* __restrict__ sigma, int ns, int n, int m,
                  O* __restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ns * m) return;
  const int s = (int)(idx / m), j = (int)(idx % m);
  const int8_t* sg = sigma + (size_t)s * n;
  double acc = 0.0;
  for (int i = 0; i < n; ++i) {
    const double w = W64[(size_t)i * m + j];
    acc = sg[i] > 0 ? acc + w : acc - w;
  }
  out[idx] = (O)(acc + b64[j]);
}

}  // namespace

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

void launch_sample(const SampleArgs& a, hipStream_t st) {
  ProfGuard g("sample", st);
  if (a.mixed) {
    if (a.kind == 1) launch_r<float, 1>(a, st);
    else launch_r<float, 0>(a, st);
  } else {
    if (a.kind == 1) launch_r<double, 1>(a, st);
    else launch_r<double, 0>(a, st);
  }
}

void launch_theta_batch(const double* W64, const double* b64,
                        const int8_t* sigma, int ns, int n, int m, void* out,
                        int out_float, hipStream_t st) {
  const size_t total = (size_t)ns * m;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  ProfGuard g("theta_batch", st);
  if (out_float)
    hipLaunchKernelGGL(k_theta_batch<float>, grid, block, 0, st, W64, b64,
                       sigma, ns, n, m, static_cast<float*>(out));
  else
    hipLaunchKernelGGL(k_theta_batch<double>, grid, block, 0, st, W64, b64,
                       sigma, ns, n, m, static_cast<double*>(out));
}

}  // namespace yq
