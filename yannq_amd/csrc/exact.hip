// Exact sums over a basis of the Hilbert space (GroundState/SRMatExact.hpp:
// 97-149, Samplers/ExactSampler.hpp:49-77, Machines/RBM.hpp:411-451):
// every configuration of the full basis (toSigma order, Utilities/Utility.hpp:
// 91-99) or of a fixed-magnetisation sector (Basis/BasisJz.hpp order) becomes
// one row of the batch, weighted by p_s = |psi_s|^2 / sum |psi|^2.  With
// those weights the batch's E_loc mean and variance and the SR quantities are
// the exact expectation values SRMatExact::constructExact computes:
//   E = sum_s p_s E_loc(s)           (= <psi|H|psi> / <psi|psi>)
//   Var = sum_s p_s |E_loc(s)|^2 - E^2
//   ō = sum_s p_s O_s,  F = sum_s p_s conj(O_s - ō) E_loc(s),
//   S = sum_s p_s conj(O_s - ō) (O_s - ō)^T
// since conj(psi_s) (H psi)_s = |psi_s|^2 E_loc(s).  The weights are formed
// in the log domain (shifted by the largest Re log psi) so no amplitude
// overflows.  ExactSampler draws configurations from p by inverse CDF.
#include "yq_internal.h"
#include "yq_math.h"

namespace yq {
namespace {

constexpr int kScanTile = 2048;  // elements per block of the prefix scan (256 x 8)

__global__ void k_basis_spins(const uint32_t* __restrict__ basis, int B, int n,
                              int8_t* __restrict__ sigma) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)B * n) return;
  const size_t s = q / n;
  const int i = (int)(q % n);
  sigma[q] = ((basis[s] >> i) & 1u) ? (int8_t)-1 : (int8_t)1;  // 1 - 2 bit_i
}

// log psi(sigma) = a . sigma + sum_j logcosh(theta_j) (RBM.hpp:367-378);
// complex: (re, im) pairs throughout
__global__ void k_log_psi(const int8_t* __restrict__ sigma, const double* __restrict__ theta,
                          const double* __restrict__ a, int B, int n, int m, int cplx,
                          double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const int8_t* sg = sigma + (size_t)s * n;
  if (!cplx) {
    double v = 0.0;
    if (a)
      for (int i = 0; i < n; ++i) v += a[i] * (double)sg[i];
    const double* th = theta + (size_t)s * m;
    for (int j = 0; j < m; ++j) v += logcosh_d(th[j]);
    out[s] = v;
  } else {
    double re = 0.0, im = 0.0;
    if (a)
      for (int i = 0; i < n; ++i) {
        re += a[2 * i] * (double)sg[i];
        im += a[2 * i + 1] * (double)sg[i];
      }
    const double* th = theta + 2 * (size_t)s * m;
    for (int j = 0; j < m; ++j) {
      double lr, li;
      logcosh_c_d(th[2 * j], th[2 * j + 1], &lr, &li);
      re += lr;
      im += li;
    }
    out[2 * (size_t)s] = re;
    out[2 * (size_t)s + 1] = im;
  }
}

// fixed-order block reductions (256 threads): max or sum of v over the block
__device__ double block_reduce(double v, bool is_max) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) red[t] = is_max ? fmax(red[t], red[t + w]) : red[t] + red[t + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// per-block max of Re log psi
__global__ void __launch_bounds__(256) k_max_re(const double* __restrict__ lp, int B, int cplx,
                                                double* __restrict__ part) {
  double v = -HUGE_VAL;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < B; s += gridDim.x * 256)
    v = fmax(v, lp[cplx ? 2 * (size_t)s : (size_t)s]);
  const double r = block_reduce(v, true);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}
// w_s = exp(pw (Re log psi_s - max)) (|psi|^2: pw = 2; an AmplitudePhase
// machine's |sqrt(RBM)|^2 = RBM: pw = 1); per-block sums of w
__global__ void __launch_bounds__(256) k_unnorm_weights(const double* __restrict__ lp, int B,
                                                        int cplx, double pw,
                                                        const double* __restrict__ mx,
                                                        int nmx, double* __restrict__ w,
                                                        double* __restrict__ part) {
  __shared__ double top;
  if (threadIdx.x == 0) {
    double v = -HUGE_VAL;
    for (int i = 0; i < nmx; ++i) v = fmax(v, mx[i]);
    top = v;
  }
  __syncthreads();
  double acc = 0.0;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < B; s += gridDim.x * 256) {
    const double v = exp(pw * (lp[cplx ? 2 * (size_t)s : (size_t)s] - top));
    w[s] = v;
    acc += v;
  }
  const double r = block_reduce(acc, false);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}
// w_s /= sum w; rs_s = sqrt(B w_s) (the row scale of the weighted SR build)
__global__ void __launch_bounds__(256) k_normalize(double* __restrict__ w, double* __restrict__ rs,
                                                   int B, const double* __restrict__ part,
                                                   int nparts) {
  __shared__ double tot;
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int i = 0; i < nparts; ++i) v += part[i];
    tot = v;
  }
  __syncthreads();
  for (int s = blockIdx.x * 256 + threadIdx.x; s < B; s += gridDim.x * 256) {
    const double v = w[s] / tot;
    w[s] = v;
    rs[s] = sqrt((double)B * v);
  }
}

// E = sum w Re e; Var = sum w |e|^2 - E^2 (one block, fixed order)
__global__ void __launch_bounds__(256) k_eloc_stats_w(const double* __restrict__ e,
                                                      const double* __restrict__ w, int ns,
                                                      int cplx, double* __restrict__ out) {
  double a = 0.0, b = 0.0;
  for (int s = threadIdx.x; s < ns; s += 256) {
    const double re = cplx ? e[2 * (size_t)s] : e[s];
    const double im = cplx ? e[2 * (size_t)s + 1] : 0.0;
    a += w[s] * re;
    b += w[s] * (re * re + im * im);
  }
  const double sa = block_reduce(a, false);
  const double sb = block_reduce(b, false);
  if (threadIdx.x == 0) {
    out[0] = sa;
    out[1] = sb - sa * sa;
  }
}

// inclusive prefix sums: tile sums of 2048 elements (thread: 8 consecutive)
__global__ void __launch_bounds__(256) k_scan_tiles(const double* __restrict__ x, int B,
                                                    double* __restrict__ out,
                                                    double* __restrict__ tile_sum) {
  __shared__ double ex[256];
  const int t = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)t * 8;
  double v[8], run = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = base + k < (size_t)B ? x[base + k] : 0.0;
    run += v[k];
  }
  ex[t] = run;
  __syncthreads();
  // exclusive scan over the 256 thread sums (one thread, fixed order)
  if (t == 0) {
    double acc = 0.0;
    for (int i = 0; i < 256; ++i) {
      const double c = ex[i];
      ex[i] = acc;
      acc += c;
    }
    tile_sum[blockIdx.x] = acc;
  }
  __syncthreads();
  double acc = ex[t];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    acc += v[k];
    if (base + k < (size_t)B) out[base + k] = acc;
  }
}
// exclusive scan of the tile sums (one thread) and the offsets added to the tiles
__global__ void k_scan_tile_sums(double* __restrict__ tile_sum, int ntiles) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double acc = 0.0;
    for (int i = 0; i < ntiles; ++i) {
      const double c = tile_sum[i];
      tile_sum[i] = acc;
      acc += c;
    }
  }
}
__global__ void k_scan_add(double* __restrict__ out, int B, const double* __restrict__ offs) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < (size_t)B) out[q] += offs[q / kScanTile];
}

// ExactSampler::sample (ExactSampler.hpp:49-77): draw k takes u in [0, 1) from
// Philox (seed, event k, exact-sampler domain), r = u * total, and the first
// index whose inclusive prefix sum is >= r (std::lower_bound); its spins go
// to row k
__global__ void k_exact_draw(const double* __restrict__ cdf, int B, const uint32_t* __restrict__ basis,
                             int n, int count, uint64_t seed, int8_t* __restrict__ sigma,
                             int* __restrict__ idx_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const u32x4 r4 = draw(seed, (uint64_t)k, 0u, kDomExact);
  const double r = u01_double(r4.x, r4.y) * cdf[B - 1];
  int lo = 0, hi = B - 1;  // first cdf >= r; r < total keeps it in range
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (cdf[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  if (idx_out) idx_out[k] = lo;
  const uint32_t v = basis[lo];
  for (int i = 0; i < n; ++i) sigma[(size_t)k * n + i] = ((v >> i) & 1u) ? (int8_t)-1 : (int8_t)1;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int exact_partials(int B) { return (int)std::min<size_t>(blocks(B), 1024); }
int exact_scan_tiles(int B) { return (B + kScanTile - 1) / kScanTile; }

void launch_basis_spins(const uint32_t* basis, int B, int n, int8_t* sigma, hipStream_t st) {
  hipLaunchKernelGGL(k_basis_spins, dim3(blocks((size_t)B * n)), dim3(256), 0, st, basis, B, n,
                     sigma);
}

void launch_exact_weights(const int8_t* sigma, const double* theta, const double* a, int B, int n,
                          int m, int cplx, double* logpsi, double* w, double* rs, double* part,
                          hipStream_t st, double pw) {
  hipLaunchKernelGGL(k_log_psi, dim3(blocks(B)), dim3(256), 0, st, sigma, theta, a, B, n, m, cplx,
                     logpsi);
  const int np = exact_partials(B);
  hipLaunchKernelGGL(k_max_re, dim3(np), dim3(256), 0, st, logpsi, B, cplx, part);
  hipLaunchKernelGGL(k_unnorm_weights, dim3(np), dim3(256), 0, st, logpsi, B, cplx, pw, part, np,
                     w, part + np);
  hipLaunchKernelGGL(k_normalize, dim3(np), dim3(256), 0, st, w, rs, B, part + np, np);
}

void launch_log_psi(const int8_t* sigma, const double* theta, const double* a, int B, int n, int m,
                    int cplx, double* logpsi, hipStream_t st) {
  hipLaunchKernelGGL(k_log_psi, dim3(blocks(B)), dim3(256), 0, st, sigma, theta, a, B, n, m, cplx,
                     logpsi);
}

void launch_eloc_stats_w(const double* e, const double* w, int ns, int cplx, double* out,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_eloc_stats_w, dim3(1), dim3(256), 0, st, e, w, ns, cplx, out);
}

void launch_exact_cdf(const double* w, int B, double* cdf, double* tile_sums, hipStream_t st) {
  const int nt = exact_scan_tiles(B);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, st, w, B, cdf, tile_sums);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(64), 0, st, tile_sums, nt);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks(B)), dim3(256), 0, st, cdf, B, tile_sums);
}

void launch_exact_draw(const double* cdf, int B, const uint32_t* basis, int n, int count,
                       uint64_t seed, int8_t* sigma, int* idx, hipStream_t st) {
  hipLaunchKernelGGL(k_exact_draw, dim3(blocks(count)), dim3(256), 0, st, cdf, B, basis, n, count,
                     seed, sigma, idx);
}

}  // namespace yq
