// Complex RBM (RBM<std::complex<double>>) on gfx950: sampler, local
// energies and the dense SR building blocks.
//
// Parameters are complex doubles interleaved (re, im) in the [W | a | b]
// layout of Machines/RBM.hpp:271-281.  The kernels read them as planes: the
// real and imaginary parts of W by visible unit (i*m + j) and transposed
// (j*n + i), and Re a / Im a, in double and float (launch_cplx_planes).
//
// Sampler: one wavefront per chain as in sampler.hip, with Re and Im of
// theta in VGPRs.  Acceptance needs only Re dlogpsi (RBMState.hpp:42-45,
// logRatioRe), so each hidden unit costs one Re logcosh (yq_math.h
// logcosh_re, replayed bit for bit on the CPU); the lane-strided sum and the
// stride-32..1 butterfly are those of the real sampler.
//
// Local energies: one thread per (sample, bond) with the full complex
// log-ratio, ratio = exp(Re) (cos Im + i sin Im).
//
// SR: O is held conjugated, Oh = conj(O) (ns x dim complex, row-major), so
// that with Oh read as a dim x ns column-major matrix A:
//   S   = A A^H / ns          (the hand-written SYRK on the rows read as
//                              2 dim reals, then Re / Im S from its 2 x 2
//                              blocks: herk_mfma in api.hip)
//   F   = A E_c / ns          (column sums below)
//   O_c v = A^H v,  O_c^H y = A y   (zgemv C and N)
// which are the adjoint products of SRMat.hpp:89-103, 142-149, 170-189.
#include "yq_internal.h"
#include "yq_math.h"

namespace yq {
namespace {

constexpr int kChunkC = 256;  // samples per partial-sum chunk (as sr.hip)

__global__ void k_cplx_planes(const double* __restrict__ p, int n, int m, CplxPlanes P) {
  const size_t nm = (size_t)n * m;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nm) {
    const size_t i = q / m, j = q % m;
    const double wr = p[2 * q], wi = p[2 * q + 1];
    P.Wr64[q] = wr;
    P.Wi64[q] = wi;
    P.Wr32[q] = (float)wr;
    P.Wi32[q] = (float)wi;
    P.WTr64[j * n + i] = wr;
    P.WTi64[j * n + i] = wi;
    P.WTr32[j * n + i] = (float)wr;
    P.WTi32[j * n + i] = (float)wi;
  } else if (q < nm + n) {
    const size_t i = q - nm;
    const double ar = p[2 * q], ai = p[2 * q + 1];
    P.ar64[i] = ar;
    P.ai64[i] = ai;
    P.ar32[i] = (float)ar;
    P.ai32[i] = (float)ai;
  }
}

template <typename R>
__device__ inline R lc_re(R x, R y);
template <>
__device__ inline double lc_re<double>(double x, double y) {
  return logcosh_re_d(x, y);
}
template <>
__device__ inline float lc_re<float>(float x, float y) {
  return logcosh_re_f(x, y);
}

// theta_j = (sum_i +/-W_ji, index order, fp64, Re and Im apart) + b_j, then
// rounded to R, and lc_j = Re logcosh(theta_j), for the hidden units of
// thread gl of a G-thread chain group
template <typename R, int RMAX, int G = kWave>
__device__ __forceinline__ void refresh_c(const SampleArgs& A, const int8_t* sig, int gl,
                                          R (&tr)[RMAX], R (&ti)[RMAX], R (&lc)[RMAX]) {
  const int n = A.n, m = A.m;
  double sr[RMAX], si[RMAX];
#pragma unroll
  for (int r = 0; r < RMAX; ++r) sr[r] = si[r] = 0.0;
  for (int i = 0; i < n; ++i) {
    const bool up = sig[i] > 0;
    const double* wc = A.W64 + 2 * (size_t)i * m;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = gl + G * r;
      if (j < m) {
        const double wr = wc[2 * j], wi = wc[2 * j + 1];
        sr[r] = up ? sr[r] + wr : sr[r] - wr;
        si[r] = up ? si[r] + wi : si[r] - wi;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    const int j = gl + G * r;
    if (j < m) {
      tr[r] = (R)(sr[r] + A.b64[2 * j]);
      ti[r] = (R)(si[r] + A.b64[2 * j + 1]);
      lc[r] = lc_re<R>(tr[r], ti[r]);
    } else {
      tr[r] = ti[r] = lc[r] = (R)0;
    }
  }
}

// One sweep (A.nprop proposals) of one chain at inverse temperature beta;
// returns the accepted count.  ev is the chain's proposal counter; gl, G,
// red and par as in the real sweep (sampler.hip): a G > 64 group is the
// whole block, every wave stores each accepted flip itself.
template <typename R, int KIND, int RMAX, int G = kWave>
__device__ __forceinline__ uint64_t sweep_chain_c(const SampleArgs& A, int chain, int8_t* sig,
                                                  int gl, R (&tr)[RMAX], R (&ti)[RMAX],
                                                  R (&lc)[RMAX], R (&ur)[RMAX], R (&ui)[RMAX],
                                                  R (&lt)[RMAX], uint64_t& ev, R beta,
                                                  R* red = nullptr, int* par = nullptr) {
  const int n = A.n, m = A.m;
  const R* Wr = static_cast<const R*>(A.Wr);
  const R* Wi = static_cast<const R*>(A.Wi);
  const R* ar = static_cast<const R*>(A.ar);
  uint64_t nacc = 0;
  for (int p = 0; p < A.nprop; ++p) {
    // wave-uniform counter and stream (lane 0's): Philox on the scalar unit
    const uint64_t evu =
        ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ev >> 32)) << 32) |
        (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)ev);
    ev = evu + 1;
    const u32x4 x = draw(A.seed, evu, __builtin_amdgcn_readfirstlane((uint32_t)chain),
                         kDomSweep);
    const int k = __builtin_amdgcn_readfirstlane((int)index_from(x.x, (uint32_t)n));
    const int8_t sk = sig[k];
    const R twok = sk > 0 ? (R)2 : (R)-2;
    const R* wkr = Wr + (size_t)k * m;
    const R* wki = Wi + (size_t)k * m;
    R cst = ((R)-2 * ar[k]) * (R)sk;
    int l = 0;
    int8_t sl = 0;
    R twol = (R)0;
    const R* wlr = wkr;
    const R* wli = wki;
    if (KIND == 1) {
      l = __builtin_amdgcn_readfirstlane((int)index_from(x.y, (uint32_t)n));
      sl = sig[l];
      if (sk == sl) continue;  // SwapSweeper.hpp:44-47, uniform branch
      twol = sl > 0 ? (R)2 : (R)-2;
      wlr = Wr + (size_t)l * m;
      wli = Wi + (size_t)l * m;
      cst = cst - ((R)2 * ar[l]) * (R)sl;
    }
    R part = (R)0;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = gl + G * r;
      if (j < m) {
        R vr = tr[r] - twok * wkr[j];
        R vi = ti[r] - twok * wki[j];
        if (KIND == 1) {
          vr = vr - twol * wlr[j];
          vi = vi - twol * wli[j];
        }
        ur[r] = vr;
        ui[r] = vi;
        lt[r] = lc_re<R>(vr, vi);
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
          tr[r] = ur[r];
          ti[r] = ui[r];
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

__device__ __forceinline__ void wave_sync_c() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One chain run by a G-thread chain group (one wavefront, or the whole block
// when G > 64): sig is the group's LDS spin strip, gl the thread's index in
// the group, red group_sum's LDS exchange.
template <typename R, int KIND, int RMAX, int G = kWave>
__device__ void run_chain_c(const SampleArgs& A, int chain, int8_t* sig, int gl,
                            R* red = nullptr) {
  const int n = A.n, m = A.m;
  for (int i = gl; i < n; i += G) sig[i] = A.sigma[(size_t)chain * n + i];
  if (G > kWave)
    __syncthreads();
  else
    wave_sync_c();
  int par = 0;
  R tr[RMAX], ti[RMAX], lc[RMAX], ur[RMAX], ui[RMAX], lt[RMAX];
  uint64_t ev = A.ctr[chain];
  uint64_t nacc = 0;
  const R beta = (R)A.beta;
  refresh_c<R, RMAX, G>(A, sig, gl, tr, ti, lc);
  bool fresh = true;
  const int total = A.n_discard + A.n_sweeps;
  for (int sw = 0; sw < total; ++sw) {
    if (A.mixed && !fresh) refresh_c<R, RMAX, G>(A, sig, gl, tr, ti, lc);
    fresh = false;
    nacc += sweep_chain_c<R, KIND, RMAX, G>(A, chain, sig, gl, tr, ti, lc, ur, ui, lt, ev, beta,
                                            red, &par);
    if (sw >= A.n_discard) {
      const size_t row = (size_t)(sw - A.n_discard) * A.C + chain;
      for (int i = gl; i < n; i += G) A.out_sigma[row * n + i] = sig[i];
      if (A.mixed) {
        refresh_c<R, RMAX, G>(A, sig, gl, tr, ti, lc);
        fresh = true;
      }
      R* ot = static_cast<R*>(A.out_theta) + 2 * row * m;
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int j = gl + G * r;
        if (j < m) {
          ot[2 * j] = tr[r];
          ot[2 * j + 1] = ti[r];
        }
      }
    }
  }
  for (int i = gl; i < n; i += G) A.sigma[(size_t)chain * n + i] = sig[i];
  if (gl == 0) {
    A.ctr[chain] = ev;
    A.acc[chain] += nacc;
  }
}

template <typename R, int KIND, int RMAX>
__global__ void __launch_bounds__(256) k_sample_c(const SampleArgs A) {
  extern __shared__ int8_t lds_sig[];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int chain = blockIdx.x * 4 + wave;
  if (chain >= A.C) return;  // wave-uniform; no block barrier follows
  run_chain_c<R, KIND, RMAX>(A, chain, lds_sig + wave * A.npad, lane);
}

// m > 512: one chain per block of G threads (LDS: the spin strip, then
// group_sum's exchange of 2 G values)
template <typename R, int KIND, int RMAX, int G>
__global__ void __launch_bounds__(G) k_sample_c_wide(const SampleArgs A) {
  extern __shared__ __align__(16) unsigned char lds_w[];
  R* red = reinterpret_cast<R*>(lds_w + ((A.npad + 15) & ~15));
  run_chain_c<R, KIND, RMAX, G>(A, blockIdx.x, reinterpret_cast<int8_t*>(lds_w), threadIdx.x,
                                red);
}

// Parallel tempering of complex machines (SamplerMT.hpp:44-60, 112-204 with
// nTmps = T > 1): the scheme of the real kernel (sampler.hip, k_sample_pt)
// with complex theta storage.  One block per chain group g; chain g * T + t
// is temperature slot t at beta_t = (T - t) / T; a slot's spins and theta
// (Re and Im) live in LDS storage idx[t]; after every sweep of all slots the
// swap passes of mixChains (:120-153) accept with
// u < exp((beta_{t+1} - beta_t) 2 (lp_t - lp_{t+1})), lp = Re log psi =
// Re a . sigma + sum_j Re logcosh(theta_j); slot 0 is recorded after every
// sweep past the discard phase.
template <typename R, int KIND, int RMAX>
__global__ void __launch_bounds__(256) k_sample_pt_c(const SampleArgs A) {
  extern __shared__ unsigned char lds_raw[];
  const int T = A.ntemps, n = A.n, m = A.m, npad = A.npad;
  const int NW = (T % 4 == 0) ? 4 : 2;
  R* lpv = reinterpret_cast<R*>(lds_raw);                      // 16
  uint64_t* evs = reinterpret_cast<uint64_t*>(lds_raw + 128);  // 16
  uint64_t* nac = reinterpret_cast<uint64_t*>(lds_raw + 256);  // 16
  int* idx = reinterpret_cast<int*>(lds_raw + 384);            // 16
  int8_t* strips = reinterpret_cast<int8_t*>(lds_raw + 448);   // T x npad
  R* thr = reinterpret_cast<R*>(lds_raw + 448 + ((T * npad + 15) & ~15));  // T x m
  R* thi = thr + (size_t)T * m;                                            // T x m
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int G = kWave;  // one wave per temperature slot
  const int gl = lane;
  const int g = blockIdx.x;
  const int groups = A.C / T;
  const R* ar = static_cast<const R*>(A.ar);
  R tr[RMAX], ti[RMAX], lc[RMAX], ur[RMAX], ui[RMAX], lt[RMAX];
  auto keep = [&](int st) {
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int j = gl + G * r;
      if (j < m) {
        thr[st * m + j] = tr[r];
        thi[st * m + j] = ti[r];
      }
    }
  };
  for (int t = w; t < T; t += NW) {
    const int chain = g * T + t;
    int8_t* sig = strips + t * npad;
    for (int i = lane; i < n; i += kWave) sig[i] = A.sigma[(size_t)chain * n + i];
    wave_sync_c();
    refresh_c<R, RMAX>(A, sig, lane, tr, ti, lc);
    keep(t);
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
        refresh_c<R, RMAX>(A, sig, lane, tr, ti, lc);
      } else {
#pragma unroll
        for (int r = 0; r < RMAX; ++r) {
          const int j = gl + G * r;
          tr[r] = j < m ? thr[st * m + j] : (R)0;
          ti[r] = j < m ? thi[st * m + j] : (R)0;
          lc[r] = j < m ? lc_re<R>(tr[r], ti[r]) : (R)0;
        }
      }
      uint64_t ev = evs[t];
      const R beta = (R)((double)(T - t) / T);
      const uint64_t acc =
          sweep_chain_c<R, KIND, RMAX>(A, chain, sig, lane, tr, ti, lc, ur, ui, lt, ev, beta);
      keep(st);
      R sa = (R)0, sl = (R)0;
      for (int i = lane; i < n; i += kWave) sa = sa + ar[i] * (R)sig[i];
#pragma unroll
      for (int r = 0; r < RMAX; ++r)
        if (lane + kWave * r < m) sl = sl + lc[r];
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
      wave_sync_c();
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
      if (A.mixed) refresh_c<R, RMAX>(A, sig, lane, tr, ti, lc);
      R* ot = static_cast<R*>(A.out_theta) + 2 * row * m;
#pragma unroll
      for (int r = 0; r < RMAX; ++r) {
        const int j = gl + G * r;
        if (j < m) {
          ot[2 * j] = A.mixed ? tr[r] : thr[st * m + j];
          ot[2 * j + 1] = A.mixed ? ti[r] : thi[st * m + j];
        }
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

template <typename R, int KIND>
void launch_pt_c(const SampleArgs& a, hipStream_t st) {
  const dim3 grid(a.C / a.ntemps), block(kWave * (a.ntemps % 4 == 0 ? 4 : 2));
  const size_t lds = pt_lds_bytes(a);
  switch ((a.m + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 1>), grid, block, lds, st, a); break;
    case 2: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 2>), grid, block, lds, st, a); break;
    case 3: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 3>), grid, block, lds, st, a); break;
    case 4: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 4>), grid, block, lds, st, a); break;
    case 5: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 5>), grid, block, lds, st, a); break;
    case 6: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 6>), grid, block, lds, st, a); break;
    case 7: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 7>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((k_sample_pt_c<R, KIND, 8>), grid, block, lds, st, a); break;
  }
}

template <typename R, int KIND, int G>
void launch_rc_wide(const SampleArgs& a, hipStream_t st) {
  const size_t lds = (size_t)((a.npad + 15) & ~15) + 2 * G * sizeof(R);
  if (a.m <= 6 * G)
    hipLaunchKernelGGL((k_sample_c_wide<R, KIND, 6, G>), dim3(a.C), dim3(G), lds, st, a);
  else
    hipLaunchKernelGGL((k_sample_c_wide<R, KIND, 8, G>), dim3(a.C), dim3(G), lds, st, a);
}

template <typename R, int KIND>
void launch_rc(const SampleArgs& a, hipStream_t st) {
  switch (chain_group(a.m)) {
    case kWave: break;
    case 128: launch_rc_wide<R, KIND, 128>(a, st); return;
    case 256: launch_rc_wide<R, KIND, 256>(a, st); return;
    case 512: launch_rc_wide<R, KIND, 512>(a, st); return;
    default: launch_rc_wide<R, KIND, 1024>(a, st); return;
  }
  const dim3 grid((a.C + 3) / 4), block(256);
  const size_t lds = (size_t)4 * a.npad;
  switch ((a.m + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL((k_sample_c<R, KIND, 1>), grid, block, lds, st, a); break;
    case 2: hipLaunchKernelGGL((k_sample_c<R, KIND, 2>), grid, block, lds, st, a); break;
    case 3: hipLaunchKernelGGL((k_sample_c<R, KIND, 3>), grid, block, lds, st, a); break;
    case 4: hipLaunchKernelGGL((k_sample_c<R, KIND, 4>), grid, block, lds, st, a); break;
    case 5: hipLaunchKernelGGL((k_sample_c<R, KIND, 5>), grid, block, lds, st, a); break;
    case 6: hipLaunchKernelGGL((k_sample_c<R, KIND, 6>), grid, block, lds, st, a); break;
    case 7: hipLaunchKernelGGL((k_sample_c<R, KIND, 7>), grid, block, lds, st, a); break;
    default: hipLaunchKernelGGL((k_sample_c<R, KIND, 8>), grid, block, lds, st, a); break;
  }
}

// theta for a batch, one thread per (sample, hidden unit); out interleaved
template <typename O>
__global__ void __launch_bounds__(256)
    k_theta_batch_c(const double* __restrict__ W64, const double* __restrict__ b64,
                    const int8_t* __restrict__ sigma, int ns, int n, int m, O* __restrict__ out) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ns * m) return;
  const int s = (int)(idx / m), j = (int)(idx % m);
  const int8_t* sg = sigma + (size_t)s * n;
  double sr = 0.0, si = 0.0;
  for (int i = 0; i < n; ++i) {
    const double wr = W64[2 * ((size_t)i * m + j)], wi = W64[2 * ((size_t)i * m + j) + 1];
    sr = sg[i] > 0 ? sr + wr : sr - wr;
    si = sg[i] > 0 ? si + wi : si - wi;
  }
  out[2 * idx] = (O)(sr + b64[2 * j]);
  out[2 * idx + 1] = (O)(si + b64[2 * j + 1]);
}

// ---------------------------------------------------------------- E_loc
__device__ inline void lc_c(double x, double y, double* re, double* im) {
  logcosh_c_d(x, y, re, im);
}
__device__ inline void lc_c(float x, float y, float* re, float* im) {
  logcosh_c_f(x, y, re, im);
}
__device__ inline void sc_c(double x, double* s, double* c) { sincos_d(x, s, c); }
__device__ inline void sc_c(float x, float* s, float* c) { sincos_f(x, s, c); }

// One thread per (sample, bond): the complex log-ratio of RBMState.hpp:48-63
// / 78-93 in the reference's j order; ratio = exp of it (complex double).
// STAGED: theta and logcosh(theta) of the block's samples in LDS; else (large
// m) theta is read from global memory and logcosh(theta_j) recomputed (the
// same bits).
template <typename R, bool STAGED>
__global__ void __launch_bounds__(256) k_eloc_ratio_c(const ElocArgs A, int span) {
  extern __shared__ unsigned char lds_raw[];
  const size_t sm = STAGED ? (size_t)span * A.m : 0;
  R* thr = reinterpret_cast<R*>(lds_raw);
  R* thi = thr + sm;
  R* lcr = thi + sm;
  R* lci = lcr + sm;
  int8_t* sg = reinterpret_cast<int8_t*>(lci + sm);
  const int n = A.n, m = A.m, B = A.B;
  const size_t first = (size_t)blockIdx.x * 256;
  const size_t total = (size_t)A.ns * B;
  const int s_lo = (int)(first / B);
  const size_t last = first + 255 < total ? first + 255 : total - 1;
  const int s_hi = (int)(last / B);
  const int cnt = s_hi - s_lo + 1;
  const R* theta = static_cast<const R*>(A.theta);
  if (STAGED)
    for (int q = threadIdx.x; q < cnt * m; q += 256) {
      const R x = theta[2 * ((size_t)s_lo * m + q)], y = theta[2 * ((size_t)s_lo * m + q) + 1];
      thr[q] = x;
      thi[q] = y;
      lc_c(x, y, &lcr[q], &lci[q]);
    }
  for (int q = threadIdx.x; q < cnt * n; q += 256) sg[q] = A.sigma[(size_t)s_lo * n + q];
  __syncthreads();
  const size_t idx = first + threadIdx.x;
  if (idx >= total) return;
  const int s = (int)(idx / B), b = (int)(idx % B);
  const int ls = s - s_lo;
  const int8_t* sgs = sg + (size_t)ls * n;
  int k, l;
  bond_sites(A.ham.kind, n, b, k, l);
  double2* ratio = reinterpret_cast<double2*>(A.ratio);
  if (l >= 0 && sgs[k] == sgs[l]) {  // prefactor (1 - zz) vanishes
    ratio[idx] = make_double2(0.0, 0.0);
    return;
  }
  const R* WTr = static_cast<const R*>(A.WT);
  const R* WTi = static_cast<const R*>(A.WTi);
  const R* ar = static_cast<const R*>(A.ar);
  const R* ai = static_cast<const R*>(A.ai);
  const int8_t sk = sgs[k];
  const R twok = sk > 0 ? (R)2 : (R)-2;
  R rr = ((R)-2 * ar[k]) * (R)sk, ri = ((R)-2 * ai[k]) * (R)sk;
  const R* tsr = thr + (size_t)ls * m;
  const R* tsi = thi + (size_t)ls * m;
  const R* lsr = lcr + (size_t)ls * m;
  const R* lsi = lci + (size_t)ls * m;
  int8_t sl = 0;
  R twol = (R)0;
  if (l >= 0) {
    sl = sgs[l];
    twol = sl > 0 ? (R)2 : (R)-2;
    rr = rr - ((R)2 * ar[l]) * (R)sl;
    ri = ri - ((R)2 * ai[l]) * (R)sl;
  }
  const R* tg = theta + 2 * (size_t)s * m;  // unstaged: this sample's theta
  for (int j = 0; j < m; ++j) {
    const R t0r = STAGED ? tsr[j] : tg[2 * j], t0i = STAGED ? tsi[j] : tg[2 * j + 1];
    R x = t0r - twok * WTr[(size_t)j * n + k];
    R y = t0i - twok * WTi[(size_t)j * n + k];
    if (l >= 0) {
      x = x - twol * WTr[(size_t)j * n + l];
      y = y - twol * WTi[(size_t)j * n + l];
    }
    R lr, li;
    lc_c(x, y, &lr, &li);
    R l0r, l0i;
    if (STAGED) {
      l0r = lsr[j];
      l0i = lsi[j];
    } else {
      lc_c(t0r, t0i, &l0r, &l0i);
    }
    rr = rr + (lr - l0r);
    ri = ri + (li - l0i);
  }
  R sn, cs;
  sc_c(ri, &sn, &cs);
  const R mag = Math<R>::exp(rr);
  ratio[idx] = make_double2((double)(mag * cs), (double)(mag * sn));
}

// the functor's accumulation (TFIsing.hpp:28-39, XXZ.hpp:37-49,
// J1J2.hpp:35-55) with complex ratios
__global__ void __launch_bounds__(256) k_eloc_combine_c(const ElocArgs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.ns) return;
  const int n = A.n;
  const int8_t* sg = A.sigma + (size_t)s * n;
  const double2* ratio = reinterpret_cast<const double2*>(A.ratio) + (size_t)s * A.B;
  const yq_ham_desc& H = A.ham;
  double er = 0.0, ei = 0.0;
  if (H.kind == YQ_TFI) {
    for (int i = 0; i < n; ++i) {
      er += H.J * (double)sg[i] * (double)sg[(i + 1) % n];
      er += H.h * ratio[i].x;
      ei += H.h * ratio[i].y;
    }
  } else if (H.kind == YQ_XXZ) {
    const double sign = H.sign_rule ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) {
      const int zz = (int)sg[i] * (int)sg[(i + 1) % n];
      er += H.J * H.delta * zz;
      const double c = sign * H.J * (1 - zz);
      er += c * ratio[i].x;
      ei += c * ratio[i].y;
    }
  } else {
    const int sign = H.sign_rule ? -1 : 1;
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)sg[i] * (int)sg[(i + 1) % n]);
      er += -H.J * yy;
      const double c = H.J * sign * (1.0 + yy);
      er += c * ratio[i].x;
      ei += c * ratio[i].y;
    }
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)sg[i] * (int)sg[(i + 2) % n]);
      er += -H.J2 * yy;
      const double c = H.J2 * (1.0 + yy);
      er += c * ratio[n + i].x;
      ei += c * ratio[n + i].y;
    }
  }
  reinterpret_cast<double2*>(A.eloc)[s] = make_double2(er, ei);
}

// E = Re <E_loc>, Var = <(Re E_loc)^2> - E^2 (Energy.hpp:53-58), one block
__global__ void __launch_bounds__(1024) k_eloc_stats_c(const double* e, int ns, double* out) {
  __shared__ double s1[1024], s2[1024];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < ns; i += 1024) {
    const double v = e[2 * (size_t)i];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = 512; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mean = s1[0] / ns;
    out[0] = mean;
    out[1] = s2[0] / ns - mean * mean;
  }
}

// ---------------------------------------------------------------- SR
template <typename I>
__global__ void k_tanh_c(const I* __restrict__ th, double* __restrict__ T, size_t cnt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cnt) return;
  double re, im;
  tanh_c_d((double)th[2 * q], (double)th[2 * q + 1], &re, &im);
  T[2 * q] = re;
  T[2 * q + 1] = im;
}

// Oh[s][c] = conj(O_sc) with O_s = logDeriv (RBM.hpp:390-406)
__global__ void k_build_Oh(const int8_t* __restrict__ sigma, const double* __restrict__ T, int ns,
                           int n, int m, int use_bias, double* __restrict__ Oh) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t total = (size_t)ns * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t s = q / dim, c = q % dim;
    double re, im;
    if (c < nm) {
      const int i = (int)(c / m), j = (int)(c % m);
      const double sg = sigma[s * n + i] > 0 ? 1.0 : -1.0;
      re = sg * T[2 * (s * m + j)];
      im = sg * T[2 * (s * m + j) + 1];
    } else if (c < nm + n) {
      re = sigma[s * n + (c - nm)] > 0 ? 1.0 : -1.0;
      im = 0.0;
    } else {
      const size_t j = c - nm - n;
      re = T[2 * (s * m + j)];
      im = T[2 * (s * m + j) + 1];
    }
    Oh[2 * q] = re;
    Oh[2 * q + 1] = -im;
  }
}

// Oh -= omean_h (in place); partial[k][c] = sum_{s in chunk k} Oh_sc e_s
// conj(O)_c = rs_s (conj(O) - om) (rs_s = 1 when rs is null) and partial
// sums of rs_s conj(O)_c e_s
__global__ void k_center_force_c(double* __restrict__ Oh, const double* __restrict__ om,
                                 const double* __restrict__ e, int ns, int dim,
                                 const double* __restrict__ rs, double* __restrict__ partial) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  const int k = blockIdx.y;
  const int s0 = k * kChunkC, s1 = min(ns, s0 + kChunkC);
  const double mr = om[2 * c], mi = om[2 * c + 1];
  double fr = 0.0, fi = 0.0;
  for (int s = s0; s < s1; ++s) {
    double* o = Oh + 2 * ((size_t)s * dim + c);
    const double f = rs ? rs[s] : 1.0;
    const double vr = (o[0] - mr) * f, vi = (o[1] - mi) * f;
    o[0] = vr;
    o[1] = vi;
    const double er = f * e[2 * s], ei = f * e[2 * s + 1];
    fr += vr * er - vi * ei;
    fi += vr * ei + vi * er;
  }
  partial[2 * ((size_t)k * dim + c)] = fr;
  partial[2 * ((size_t)k * dim + c) + 1] = fi;
}

// partial[k][c] = sum_{s in chunk k} |Oh_sc|^2
__global__ void k_absq_colsum(const double* __restrict__ Oh, int ns, int dim,
                              double* __restrict__ partial) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  const int k = blockIdx.y;
  const int s0 = k * kChunkC, s1 = min(ns, s0 + kChunkC);
  double acc = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double* o = Oh + 2 * ((size_t)s * dim + c);
    acc += o[0] * o[0] + o[1] * o[1];
  }
  partial[(size_t)k * dim + c] = acc;
}

__global__ void k_dup2(const double* __restrict__ x, double* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[2 * i] = y[2 * i + 1] = x[i];
}

// M = S + shift I for a column-major complex dim x dim matrix
__global__ void k_copy_shift_c(const double* __restrict__ S, double* __restrict__ M, int dim,
                               double shift) {
  const size_t total = (size_t)dim * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t r = q % dim, c = q / dim;
    M[2 * q] = S[2 * q] + (r == c ? shift : 0.0);
    M[2 * q + 1] = S[2 * q + 1];
  }
}

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_cplx_planes(const double* p, int n, int m, const CplxPlanes& P, hipStream_t st) {
  hipLaunchKernelGGL(k_cplx_planes, dim3(nblk((size_t)n * m + n)), dim3(256), 0, st, p, n, m, P);
}

void launch_sample_c(const SampleArgs& a, hipStream_t st) {
  ProfGuard g("sample_c", st);
  if (a.ntemps > 1) {
    if (a.mixed) {
      if (a.kind == 1) launch_pt_c<float, 1>(a, st);
      else launch_pt_c<float, 0>(a, st);
    } else {
      if (a.kind == 1) launch_pt_c<double, 1>(a, st);
      else launch_pt_c<double, 0>(a, st);
    }
    return;
  }
  if (a.mixed) {
    if (a.kind == 1) launch_rc<float, 1>(a, st);
    else launch_rc<float, 0>(a, st);
  } else {
    if (a.kind == 1) launch_rc<double, 1>(a, st);
    else launch_rc<double, 0>(a, st);
  }
}

void launch_theta_batch_c(const double* W64, const double* b64, const int8_t* sigma, int ns,
                          int n, int m, void* out, int out_float, hipStream_t st) {
  const size_t total = (size_t)ns * m;
  if (out_float)
    hipLaunchKernelGGL(k_theta_batch_c<float>, dim3(nblk(total)), dim3(256), 0, st, W64, b64,
                       sigma, ns, n, m, static_cast<float*>(out));
  else
    hipLaunchKernelGGL(k_theta_batch_c<double>, dim3(nblk(total)), dim3(256), 0, st, W64, b64,
                       sigma, ns, n, m, static_cast<double*>(out));
}

// theta staging (Re, Im and their logcosh) of a block's samples fits the
// LDS budget
static bool eloc_c_staged(const ElocArgs& a) {
  const int span = 256 / a.B + 2;
  const size_t rs = a.use_float ? sizeof(float) : sizeof(double);
  return (size_t)span * a.m * 4 * rs <= kElocStageBytes;
}

size_t eloc_c_lds_bytes(const ElocArgs& a) {
  const int span = 256 / a.B + 2;
  const size_t rs = a.use_float ? sizeof(float) : sizeof(double);
  const size_t stage = eloc_c_staged(a) ? (size_t)span * a.m * 4 * rs : 0;
  return stage + (size_t)span * a.n + 16;
}

void launch_eloc_c(const ElocArgs& a, hipStream_t st) {
  const size_t total = (size_t)a.ns * a.B;
  const int span = 256 / a.B + 2;
  const size_t lds = eloc_c_lds_bytes(a);
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  {
    ProfGuard g("eloc_ratio_c", st);
    const bool staged = eloc_c_staged(a);
    if (a.use_float) {
      if (staged) hipLaunchKernelGGL((k_eloc_ratio_c<float, true>), grid, block, lds, st, a, span);
      else hipLaunchKernelGGL((k_eloc_ratio_c<float, false>), grid, block, lds, st, a, span);
    } else {
      if (staged) hipLaunchKernelGGL((k_eloc_ratio_c<double, true>), grid, block, lds, st, a, span);
      else hipLaunchKernelGGL((k_eloc_ratio_c<double, false>), grid, block, lds, st, a, span);
    }
  }
  launch_eloc_combine_c(a, st);
}

void launch_eloc_combine_c(const ElocArgs& a, hipStream_t st) {
  ProfGuard g("eloc_combine_c", st);
  hipLaunchKernelGGL(k_eloc_combine_c, dim3((a.ns + 255) / 256), dim3(256), 0, st, a);
}

void launch_eloc_stats_c(const double* eloc, int ns, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_eloc_stats_c, dim3(1), dim3(1024), 0, st, eloc, ns, out);
}

// ---- factored complex SR operator (FP64), composed of real factored products
namespace {
__global__ void k_split_c(const double* __restrict__ z, size_t n, double* __restrict__ re,
                          double* __restrict__ im) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  re[i] = z[2 * i];
  im[i] = z[2 * i + 1];
}
// y = (P1 - P2 - Re ov, P3 + P4 - Im ov) * scale, as planes and interleaved
__global__ void k_cfwd_combine(const double* __restrict__ P1, const double* __restrict__ P2,
                               const double* __restrict__ P3, const double* __restrict__ P4,
                               const double* __restrict__ ov, size_t n, double scale,
                               double* __restrict__ yr, double* __restrict__ yi,
                               double* __restrict__ yz) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = ((P1[i] - P2[i]) - ov[0]) * scale, b = ((P3[i] + P4[i]) - ov[1]) * scale;
  yr[i] = a;
  yi[i] = b;
  yz[2 * i] = a;
  yz[2 * i + 1] = b;
}
// out = (B1 + B2, B3 - B4) - a * sy + shift v, interleaved; a and v
// interleaved, sy one complex scalar
__global__ void k_cbwd_combine(const double* __restrict__ B1, const double* __restrict__ B2,
                               const double* __restrict__ B3, const double* __restrict__ B4,
                               const double* __restrict__ a, const double* __restrict__ sy,
                               const double* __restrict__ v, double shift, size_t n,
                               double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ar = a[2 * i], ai = a[2 * i + 1];
  const double cr = ar * sy[0] - ai * sy[1], ci = ar * sy[1] + ai * sy[0];
  out[2 * i] = ((B1[i] + B2[i]) - cr) + shift * v[2 * i];
  out[2 * i + 1] = ((B3[i] - B4[i]) - ci) + shift * v[2 * i + 1];
}
// out[0..1] = sum_i conj(a_i) b_i (a, b interleaved), or sum_i a_i when b is
// null; one block, fixed order: thread-strided sums, then a tree over the
// block
__global__ void __launch_bounds__(1024) k_cdot1(const double* __restrict__ a,
                                                const double* __restrict__ b, size_t n,
                                                double* __restrict__ out) {
  __shared__ double sr[1024], si[1024];
  double xr = 0.0, xi = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 1024) {
    const double ar = a[2 * i], ai = a[2 * i + 1];
    if (b) {
      const double br = b[2 * i], bi = b[2 * i + 1];
      xr += ar * br + ai * bi;
      xi += ar * bi - ai * br;
    } else {
      xr += ar;
      xi += ai;
    }
  }
  sr[threadIdx.x] = xr;
  si[threadIdx.x] = xi;
  __syncthreads();
  for (int w = 512; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sr[0];
    out[1] = si[0];
  }
}
// (yr, yi) = (e - esum / n) / n as planes: the centred energies over n
// (e interleaved, esum their complex sum)
__global__ void k_center_c(const double* __restrict__ e, const double* __restrict__ esum,
                           size_t n, double* __restrict__ yr, double* __restrict__ yi) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dn = (double)n;
  yr[i] = (e[2 * i] - esum[0] / dn) / dn;
  yi[i] = (e[2 * i + 1] - esum[1] / dn) / dn;
}
// z = (re, -im): the conjugate, interleaved
__global__ void k_conj_merge(const double* __restrict__ re, const double* __restrict__ im,
                             size_t n, double* __restrict__ z) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  z[2 * i] = re[i];
  z[2 * i + 1] = -im[i];
}
}  // namespace

void launch_split_c(const double* z, size_t n, double* re, double* im, hipStream_t st) {
  hipLaunchKernelGGL(k_split_c, dim3(nblk(n)), dim3(256), 0, st, z, n, re, im);
}
void launch_cfwd_combine(const double* P1, const double* P2, const double* P3, const double* P4,
                         const double* ov, size_t n, double scale, double* yr, double* yi,
                         double* yz, hipStream_t st) {
  hipLaunchKernelGGL(k_cfwd_combine, dim3(nblk(n)), dim3(256), 0, st, P1, P2, P3, P4, ov, n,
                     scale, yr, yi, yz);
}
void launch_cbwd_combine(const double* B1, const double* B2, const double* B3, const double* B4,
                         const double* a, const double* sy, const double* v, double shift,
                         size_t n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_cbwd_combine, dim3(nblk(n)), dim3(256), 0, st, B1, B2, B3, B4, a, sy, v,
                     shift, n, out);
}
void launch_cdot1(const double* a, const double* b, size_t n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_cdot1, dim3(1), dim3(1024), 0, st, a, b, n, out);
}
void launch_center_c(const double* e, const double* esum, size_t n, double* yr, double* yi,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_center_c, dim3(nblk(n)), dim3(256), 0, st, e, esum, n, yr, yi);
}
void launch_conj_merge(const double* re, const double* im, size_t n, double* z, hipStream_t st) {
  hipLaunchKernelGGL(k_conj_merge, dim3(nblk(n)), dim3(256), 0, st, re, im, n, z);
}

void launch_tanh_c(const void* theta, int in_float, double* T, size_t count, hipStream_t st) {
  if (in_float)
    hipLaunchKernelGGL(k_tanh_c<float>, dim3(nblk(count)), dim3(256), 0, st,
                       static_cast<const float*>(theta), T, count);
  else
    hipLaunchKernelGGL(k_tanh_c<double>, dim3(nblk(count)), dim3(256), 0, st,
                       static_cast<const double*>(theta), T, count);
}

void launch_build_Oh(const int8_t* sigma, const double* T, int ns, int n, int m, int use_bias,
                     double* Oh, hipStream_t st) {
  ProfGuard g("build_Oh", st);
  hipLaunchKernelGGL(k_build_Oh, dim3(4096), dim3(256), 0, st, sigma, T, ns, n, m, use_bias, Oh);
}

void launch_center_force_c(double* Oh, const double* om, const double* e, int ns, int dim,
                           double* partial, hipStream_t st, const double* rs) {
  hipLaunchKernelGGL(k_center_force_c, dim3(nblk(dim), (ns + kChunkC - 1) / kChunkC), dim3(256),
                     0, st, Oh, om, e, ns, dim, rs, partial);
}

void launch_absq_colsum(const double* Oh, int ns, int dim, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(k_absq_colsum, dim3(nblk(dim), (ns + kChunkC - 1) / kChunkC), dim3(256), 0,
                     st, Oh, ns, dim, partial);
}

void launch_dup2(const double* x, double* y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_dup2, dim3(nblk(n)), dim3(256), 0, st, x, y, n);
}

void launch_copy_shift_c(const double* S, double* M, int dim, double shift, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_shift_c, dim3(8192), dim3(256), 0, st, S, M, dim, shift);
}

}  // namespace yq
