// Local energies E_loc(sigma) = sum_b c_b(sigma) psi(sigma^b)/psi(sigma) + diag
// for the TFI, XXZ and J1-J2 chains (Hamiltonians/TFIsing.hpp:28-39,
// XXZ.hpp:37-49, J1J2.hpp:35-55) as batched flip-ratio kernels.
//
// Pass 1 (k_eloc_ratio): one thread per (sample, bond).  A block of 256
// consecutive (sample, bond) pairs stages theta_s and logcosh(theta_s) of the
// few samples it touches in LDS, then each thread walks the hidden units in
// index order exactly like RBMStateObj::logRatio (RBMState.hpp:54-62 /
// 82-91), reading W^T rows that are contiguous across the bonds of a sample.
// When the staged theta of a block's samples would not fit in 48 KB of LDS
// (large m), the unstaged form reads theta from global memory and recomputes
// logcosh(theta_j) next to logcosh(t_j): the same operations on the same
// values, so the same bits.
// Pass 2 (k_eloc_combine): one thread per sample accumulates the diagonal and
// ratio terms in the functor's order, in double.
#include "yq_internal.h"
#include "yq_math.h"

namespace yq {

int num_bonds(const yq_ham_desc& h) { return h.kind == YQ_J1J2 ? 2 * h.n : h.n; }

namespace {

// One 256-thread block of (sample, bond) pairs starting at pair `first`.
// STAGED: theta and logcosh(theta) of the block's samples in LDS; else only
// the spins are staged.
template <typename R, bool STAGED = true>
__device__ void eloc_ratio_block(const ElocArgs& A, int span, size_t first,
                                 unsigned char* lds_raw) {
  R* th = reinterpret_cast<R*>(lds_raw);
  R* lc = th + (STAGED ? (size_t)span * A.m : 0);
  int8_t* sg = reinterpret_cast<int8_t*>(lc + (STAGED ? (size_t)span * A.m : 0));
  const int n = A.n, m = A.m, B = A.B;
  const size_t total = (size_t)A.ns * B;
  const int s_lo = (int)(first / B);
  const size_t last = first + 255 < total ? first + 255 : total - 1;
  const int s_hi = (int)(last / B);
  const int cnt = s_hi - s_lo + 1;
  const R* theta = static_cast<const R*>(A.theta);
  if (STAGED)
    for (int q = threadIdx.x; q < cnt * m; q += 256) {
      const R t = theta[(size_t)s_lo * m + q];
      th[q] = t;
      lc[q] = Math<R>::logcosh(t);
    }
  for (int q = threadIdx.x; q < cnt * n; q += 256)
    sg[q] = A.sigma[(size_t)s_lo * n + q];
  __syncthreads();
  const size_t idx = first + threadIdx.x;
  if (idx >= total) return;
  const int s = (int)(idx / B), b = (int)(idx % B);
  const int ls = s - s_lo;
  const int8_t* sgs = sg + (size_t)ls * n;
  int k, l;
  bond_sites(A.ham.kind, n, b, k, l);
  if (l >= 0 && sgs[k] == sgs[l]) {  // prefactor (1 - zz) vanishes
    if (A.ap_ffs)
      reinterpret_cast<double2*>(A.ratio)[idx] = make_double2(0.0, 0.0);
    else
      A.ratio[idx] = 0.0;
    return;
  }
  const R* WT = static_cast<const R*>(A.WT);
  const R* ar = static_cast<const R*>(A.ar);
  const int8_t sk = sgs[k];
  const R twok = sk > 0 ? (R)2 : (R)-2;
  R res = ((R)-2 * ar[k]) * (R)sk;
  const R* ths = STAGED ? th + (size_t)ls * m : theta + (size_t)s * m;
  const R* lcs = lc + (size_t)ls * m;
  auto lc_of = [&](int j) { return STAGED ? lcs[j] : Math<R>::logcosh(ths[j]); };
  if (l < 0) {
    for (int j = 0; j < m; ++j) {
      const R t = ths[j] - twok * WT[(size_t)j * n + k];
      res = res + (Math<R>::logcosh(t) - lc_of(j));
    }
  } else {
    const int8_t sl = sgs[l];
    const R twol = sl > 0 ? (R)2 : (R)-2;
    res = res - ((R)2 * ar[l]) * (R)sl;
    for (int j = 0; j < m; ++j) {
      R t = ths[j] - twok * WT[(size_t)j * n + k];
      t = t - twol * WT[(size_t)j * n + l];
      res = res + (Math<R>::logcosh(t) - lc_of(j));
    }
  }
  if (A.ap_ffs) {
    // AmplitudePhaseState.hpp:57-120: exp(logRatio / 2 + i pi (FF(flipped) - FF(sigma)))
    const double mag = exp((double)res / 2.0);
    const double ph = M_PI * (A.ap_ffb[idx] - A.ap_ffs[s]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    reinterpret_cast<double2*>(A.ratio)[idx] = make_double2(mag * cs, mag * sn);
    return;
  }
  A.ratio[idx] = (double)Math<R>::exp(res);
}

template <typename R, bool STAGED>
__global__ void __launch_bounds__(256) k_eloc_ratio(const ElocArgs A, int span) {
  extern __shared__ unsigned char lds_raw[];
  eloc_ratio_block<R, STAGED>(A, span, (size_t)blockIdx.x * 256, lds_raw);
}

// many machines: blockIdx.y selects the machine (its own span of samples per block)
template <typename R>
__global__ void __launch_bounds__(256) k_eloc_ratio_multi(const ElocArgs* __restrict__ inst) {
  extern __shared__ unsigned char lds_raw[];
  const ElocArgs A = inst[blockIdx.y];
  const size_t first = (size_t)blockIdx.x * 256;
  if (first >= (size_t)A.ns * A.B) return;  // block-uniform
  eloc_ratio_block<R>(A, 256 / A.B + 2, first, lds_raw);
}

__device__ void eloc_combine_one(const ElocArgs& A, int s) {
  const int n = A.n;
  const int8_t* sg = A.sigma + (size_t)s * n;
  const double* ratio = A.ratio + (size_t)s * A.B;
  const yq_ham_desc& H = A.ham;
  double e = 0.0;
  if (H.kind == YQ_TFI) {
    for (int i = 0; i < n; ++i) {
      e += H.J * (double)sg[i] * (double)sg[(i + 1) % n];
      e += H.h * ratio[i];
    }
  } else if (H.kind == YQ_XXZ) {
    const double sign = H.sign_rule ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) {
      const int zz = (int)sg[i] * (int)sg[(i + 1) % n];
      e += H.J * H.delta * zz;
      e += sign * H.J * (1 - zz) * ratio[i];
    }
  } else {
    const int sign = H.sign_rule ? -1 : 1;
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)sg[i] * (int)sg[(i + 1) % n]);
      e += -H.J * yy;
      e += H.J * sign * (1.0 + yy) * ratio[i];
    }
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)sg[i] * (int)sg[(i + 2) % n]);
      e += -H.J2 * yy;
      e += H.J2 * (1.0 + yy) * ratio[n + i];
    }
  }
  A.eloc[s] = e;
}

__global__ void __launch_bounds__(256) k_eloc_combine(const ElocArgs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < A.ns) eloc_combine_one(A, s);
}
__global__ void __launch_bounds__(256) k_eloc_combine_multi(const ElocArgs* __restrict__ inst) {
  const ElocArgs A = inst[blockIdx.y];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < A.ns) eloc_combine_one(A, s);
}

// mean and variance in one block, fixed reduction order
__device__ void eloc_stats_block(const double* e, int ns, double* out) {
  __shared__ double s1[1024], s2[1024];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < ns; i += 1024) {
    a += e[i];
    b += e[i] * e[i];
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
__global__ void __launch_bounds__(1024) k_eloc_stats(const double* e, int ns, double* out) {
  eloc_stats_block(e, ns, out);
}
// one block per machine: out[2 i], out[2 i + 1]
__global__ void __launch_bounds__(1024) k_eloc_stats_multi(const ElocArgs* __restrict__ inst,
                                                           double* out) {
  const ElocArgs A = inst[blockIdx.x];
  eloc_stats_block(A.eloc, A.ns, out + 2 * blockIdx.x);
}

}  // namespace

// theta staging of a block's samples fits the LDS budget
bool eloc_staged(const ElocArgs& a) {
  const int span = 256 / a.B + 2;
  const size_t rs = a.use_float ? sizeof(float) : sizeof(double);
  return (size_t)span * a.m * 2 * rs <= kElocStageBytes;
}

void launch_eloc_ratio(const ElocArgs& a, hipStream_t st) {
  const size_t total = (size_t)a.ns * a.B;
  const int span = 256 / a.B + 2;
  const size_t lds = eloc_lds_bytes(a);
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  ProfGuard g("eloc_ratio", st);
  const bool staged = eloc_staged(a);
  if (a.use_float) {
    if (staged) hipLaunchKernelGGL((k_eloc_ratio<float, true>), grid, block, lds, st, a, span);
    else hipLaunchKernelGGL((k_eloc_ratio<float, false>), grid, block, lds, st, a, span);
  } else {
    if (staged) hipLaunchKernelGGL((k_eloc_ratio<double, true>), grid, block, lds, st, a, span);
    else hipLaunchKernelGGL((k_eloc_ratio<double, false>), grid, block, lds, st, a, span);
  }
}

void launch_eloc(const ElocArgs& a, hipStream_t st) {
  launch_eloc_ratio(a, st);
  ProfGuard g("eloc_combine", st);
  hipLaunchKernelGGL(k_eloc_combine, dim3((a.ns + 255) / 256), dim3(256), 0, st, a);
}

size_t eloc_stage_lds_bytes(const ElocArgs& a) {
  const int span = 256 / a.B + 2;
  const size_t rs = a.use_float ? sizeof(float) : sizeof(double);
  return (size_t)span * a.m * 2 * rs + (size_t)span * a.n + 16;
}

size_t eloc_lds_bytes(const ElocArgs& a) {
  const int span = 256 / a.B + 2;
  const size_t rs = a.use_float ? sizeof(float) : sizeof(double);
  const size_t stage = eloc_staged(a) ? (size_t)span * a.m * 2 * rs : 0;
  return stage + (size_t)span * a.n + 16;
}

void launch_eloc_multi(const ElocArgs* inst_dev, int count, int max_blocks, int max_ns,
                       size_t lds, int use_float, double* stats, hipStream_t st) {
  {
    ProfGuard g("eloc_ratio_multi", st);
    const dim3 grid((unsigned)max_blocks, (unsigned)count);
    if (use_float)
      hipLaunchKernelGGL(k_eloc_ratio_multi<float>, grid, dim3(256), lds, st, inst_dev);
    else
      hipLaunchKernelGGL(k_eloc_ratio_multi<double>, grid, dim3(256), lds, st, inst_dev);
  }
  hipLaunchKernelGGL(k_eloc_combine_multi, dim3((max_ns + 255) / 256, count), dim3(256), 0, st,
                     inst_dev);
  hipLaunchKernelGGL(k_eloc_stats_multi, dim3(count), dim3(1024), 0, st, inst_dev, stats);
}

void launch_eloc_stats(const double* eloc, int ns, double* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_eloc_stats, dim3(1), dim3(1024), 0, st, eloc, ns, out);
}

}  // namespace yq
