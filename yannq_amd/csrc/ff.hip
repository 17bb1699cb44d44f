// FeedForward machine of SURVEY config C4: Conv1D(1 -> C channels, width K,
// periodic, stride 1) -> LnCosh -> FullyConnected(C*n -> H) -> Tanh -> Mean,
// without biases (the layer defaults, layers/Conv1D.hpp:57-61,
// layers/FullyConnected.hpp:57-58).  log psi(sigma) is the network output
// and O(sigma) = d log psi / d params is the hand-written reverse sweep of
// FeedForward::backward (FeedForward.hpp:186-205).
//
// One 256-thread block owns kSB samples at a time; every activation lives in
// LDS only.  Modes:
//   kFwd  log psi (optionally storing z, a1 and a2 for a later VJP);
//   kJac  the full Jacobian rows (ns x npar, the FC block is the outer
//         product a1 (x) d2 written with coalesced stores -- HBM-write bound);
//   kVjp  grad = sum_s r_s O_s without materialising O: per-thread register
//         accumulators over a grid-stride loop of sample groups, then a
//         fixed-order reduction of the per-block partials.  Activations are
//         read back when stored by kFwd and recomputed otherwise.
// Parameter layout [conv kernel (index ki + oc*K) | FC weight W(i, o) at
// i + o*C*n], FeedForward.hpp's concatenation of the layers' getParams().
#include <mutex>
#include <vector>

#include "yq_internal.h"
#include "yq_math.h"

namespace yq {
namespace {

// samples per group: 8 in float, 4 in double (LDS per block <= 64 KB at C4)
template <typename R>
constexpr int sb() {
  return sizeof(R) == 8 ? 4 : 8;
}

enum { kFwd = 0, kJac = 1, kVjp = 2 };

// periodic site index j mod n for j in [-n, 2n) (the kernel width is at
// most n), without an integer division
__device__ inline int wrap(int j, int n) {
  j += j < 0 ? n : 0;
  return j >= n ? j - n : j;
}

template <typename R>
__device__ inline R tanh_r(R x) {
  return tanh(x);
}

template <typename R, int MODE, int EMAX>
__global__ void __launch_bounds__(256) k_ff(const FFArgs A) {
  constexpr int kSB = sb<R>();
  extern __shared__ unsigned char smem[];
  const int n = A.n, C = A.C, K = A.K, H = A.H, CN = C * n;
  const int npar = K * C + CN * H;
  R* sg = reinterpret_cast<R*>(smem);  // kSB x n
  R* z = sg + kSB * n;                 // kSB x CN
  R* a1 = z + kSB * CN;                // kSB x CN
  R* dz = a1 + kSB * CN;               // kSB x CN
  R* a2 = dz + kSB * CN;               // kSB x H
  R* d2 = a2 + kSB * H;                // kSB x H
  R* rr = d2 + kSB * H;                // kSB
  const R* ker = static_cast<const R*>(A.params);
  const R* W = ker + K * C;
  const int tid = threadIdx.x;
  const int groups = (A.ns + kSB - 1) / kSB;
  R acc[EMAX > 0 ? EMAX : 1];
  if (MODE == kVjp)
#pragma unroll
    for (int e = 0; e < (EMAX > 0 ? EMAX : 1); ++e) acc[e] = (R)0;

  for (int g = blockIdx.x; g < groups; g += gridDim.x) {
    const int s0 = g * kSB;
    const int cnt = min(kSB, A.ns - s0);
    if (A.flip_B > 0) {
      for (int q = tid; q < cnt * n; q += 256) {
        const int cfg = s0 + q / n, i = q % n;
        int k, l;
        bond_sites(A.flip_kind, n, cfg % A.flip_B, k, l);
        const bool up = A.sigma[(size_t)(cfg / A.flip_B) * n + i] > 0;
        sg[q] = (up != (i == k || i == l)) ? (R)1 : (R)-1;
      }
    } else {
      for (int q = tid; q < cnt * n; q += 256)
        sg[q] = A.sigma[(size_t)s0 * n + q] > 0 ? (R)1 : (R)-1;
    }
    if (MODE == kVjp)
      for (int q = tid; q < cnt; q += 256) rr[q] = static_cast<const R*>(A.r)[s0 + q];
    const R* act_in = static_cast<const R*>(A.act_in);
    const int AW = 2 * CN + H;  // stored activations per sample: z, a1, a2
    if (MODE == kVjp && act_in) {
      __syncthreads();
      for (int q = tid; q < cnt * CN; q += 256) {
        const int s = q / CN, c = q % CN;
        z[q] = act_in[(size_t)(s0 + s) * AW + c];
        a1[q] = act_in[(size_t)(s0 + s) * AW + CN + c];
      }
      for (int q = tid; q < cnt * H; q += 256) {
        const int s = q / H, o = q % H;
        a2[q] = act_in[(size_t)(s0 + s) * AW + 2 * CN + o];
      }
      __syncthreads();
    } else {
      __syncthreads();
      // Conv1D::forward (layers/Conv1D.hpp:155-163) then LnCosh
      for (int q = tid; q < cnt * CN; q += 256) {
        const int s = q / CN, c = q % CN, oc = c / n, r = c % n;
        R v = (R)0;
        for (int ki = 0; ki < K; ++ki) v += ker[ki + oc * K] * sg[s * n + wrap(r + ki - K / 2, n)];
        z[q] = v;
        a1[q] = Math<R>::logcosh(v);
      }
      __syncthreads();
      // FullyConnected::forward (W^T a1) then Tanh.  Neighbouring threads take
      // neighbouring outputs o and read the transposed weight WT[i][o] row by
      // row (coalesced); each WT value serves all samples of the group.  When
      // H divides the block, the P = 256 / H thread groups split i into P
      // contiguous ranges whose partial sums (kept in dz, free until the
      // backward pass) are added in range order.
      const R* WT = static_cast<const R*>(A.wt);
      const int P = 256 / H;
      if (256 % H == 0 && CN >= 256) {
        const int o = tid % H, pr = tid / H;
        const int len = (CN + P - 1) / P, i0 = pr * len, i1 = min(CN, i0 + len);
        R v[kSB];
#pragma unroll
        for (int s = 0; s < kSB; ++s) v[s] = (R)0;
        for (int i = i0; i < i1; ++i) {
          const R w = WT[(size_t)i * H + o];
#pragma unroll
          for (int s = 0; s < kSB; ++s) v[s] += w * a1[s * CN + i];
        }
#pragma unroll
        for (int s = 0; s < kSB; ++s) dz[(pr * kSB + s) * H + o] = v[s];
        __syncthreads();
        for (int q = tid; q < cnt * H; q += 256) {
          const int s = q / H, oo = q % H;
          R t = (R)0;
          for (int p2 = 0; p2 < P; ++p2) t += dz[(p2 * kSB + s) * H + oo];
          a2[q] = tanh_r(t);
        }
      } else {
        for (int q = tid; q < cnt * H; q += 256) {
          const int s = q / H, o = q % H;
          const R* x = a1 + s * CN;
          R t = (R)0;
          for (int i = 0; i < CN; ++i) t += WT[(size_t)i * H + o] * x[i];
          a2[q] = tanh_r(t);
        }
      }
      __syncthreads();
    }
    if (MODE == kFwd) {
      // Mean::forward: sum / H, one thread per sample
      if (tid < cnt) {
        R s = (R)0;
        for (int o = 0; o < H; ++o) s += a2[tid * H + o];
        static_cast<R*>(A.out)[s0 + tid] = s / (R)H;
      }
      R* act = static_cast<R*>(A.act_out);
      if (act) {
        for (int q = tid; q < cnt * CN; q += 256) {
          const int s = q / CN, c = q % CN;
          act[(size_t)(s0 + s) * AW + c] = z[q];
          act[(size_t)(s0 + s) * AW + CN + c] = a1[q];
        }
        for (int q = tid; q < cnt * H; q += 256) {
          const int s = q / H, o = q % H;
          act[(size_t)(s0 + s) * AW + 2 * CN + o] = a2[q];
        }
      }
      __syncthreads();
      continue;
    }
    // Mean::backprop (1/H) then Tanh::ApplyJacobian (1 - a2^2)
    for (int q = tid; q < cnt * H; q += 256) d2[q] = ((R)1 / (R)H) * ((R)1 - a2[q] * a2[q]);
    if (MODE == kJac && tid < cnt) {
      R s = (R)0;
      for (int o = 0; o < H; ++o) s += a2[tid * H + o];
      static_cast<R*>(A.out)[s0 + tid] = s / (R)H;
    }
    __syncthreads();
    // kJac: the FullyConnected rows a1 (x) d2 (all but K*C of every row's
    // entries, the bulk of the kernel's HBM writes) go out first, so that
    // their stores drain while the backward pass below computes
    R* const jac = static_cast<R*>(A.jac);
    const int KC = K * C;
    if (MODE == kJac) {
      const int o0 = tid / CN, i0 = tid % CN;
      constexpr int VV = 16 / sizeof(R);
      const bool vec = npar % VV == 0 && KC % VV == 0 && CN % VV == 0 &&
                       reinterpret_cast<uintptr_t>(jac) % 16 == 0;
      const int o0v = tid * VV / CN, i0v = tid * VV % CN;
      for (int s = 0; s < cnt; ++s) {
        // FullyConnected: a1 (x) d2 at KC + o * CN + i, with (o, i) advanced
        // by the block stride instead of a division per element; 16-byte
        // stores when the row layout allows them
        const R* a1s = a1 + s * CN;
        const R* d2s = d2 + s * H;
        R* fc = jac + (size_t)(s0 + s) * npar + KC;
        if (vec) {
          constexpr int V = 16 / sizeof(R);
          int o = o0v, i = i0v;
          for (int q = tid * V; q < CN * H; q += 256 * V) {
            const R dv = d2s[o];
            if (V == 4) {
              float4 t;
              t.x = (float)(a1s[i] * dv);
              t.y = (float)(a1s[i + 1] * dv);
              t.z = (float)(a1s[i + 2] * dv);
              t.w = (float)(a1s[i + 3] * dv);
              *reinterpret_cast<float4*>(fc + q) = t;
            } else {
              double2 t;
              t.x = (double)(a1s[i] * dv);
              t.y = (double)(a1s[i + 1] * dv);
              *reinterpret_cast<double2*>(fc + q) = t;
            }
            i += 256 * V;
            while (i >= CN) {
              i -= CN;
              ++o;
            }
          }
        } else {
          int o = o0, i = i0;
          for (int q = tid; q < CN * H; q += 256) {
            fc[q] = a1s[i] * d2s[o];
            i += 256;
            while (i >= CN) {
              i -= CN;
              ++o;
            }
          }
        }
      }
    }
    // FullyConnected::backprop din = W d2, then LnCosh Jacobian tanh(z): one
    // thread per input i for all samples of the group, so every W value is
    // read once (coalesced over i) and the samples' sums run side by side
    for (int i = tid; i < CN; i += 256) {
      R v[kSB];
#pragma unroll
      for (int s = 0; s < kSB; ++s) v[s] = (R)0;
      for (int o = 0; o < H; ++o) {
        const R w = W[i + (size_t)o * CN];
#pragma unroll
        for (int s = 0; s < kSB; ++s) v[s] += w * d2[s * H + o];
      }
#pragma unroll
      for (int s = 0; s < kSB; ++s)
        if (s < cnt) dz[s * CN + i] = v[s] * tanh_r(z[s * CN + i]);
    }
    __syncthreads();
    if (MODE == kJac) {
      // Conv1D::backprop weight part for every (sample, kernel entry) of the
      // group, spread over the block, written straight into the rows
      for (int q = tid; q < cnt * KC; q += 256) {
        const int s = q / KC, c = q % KC, ki = c % K, oc = c / K;
        R v = (R)0;
        for (int r = 0; r < n; ++r)
          v += sg[s * n + wrap(r + ki - K / 2, n)] * dz[s * CN + r + oc * n];
        jac[(size_t)(s0 + s) * npar + c] = v;
      }
    } else {  // kVjp: acc[e] += sum_s r_s O_s[c], c = tid + 256 e
#pragma unroll
      for (int e = 0; e < EMAX; ++e) {
        const int c = tid + 256 * e;
        if (c < npar) {
          R v = (R)0;
          for (int s = 0; s < cnt; ++s) {
            R o;
            if (c < K * C) {
              const int ki = c % K, oc = c / K;
              o = (R)0;
              for (int r = 0; r < n; ++r)
                o += sg[s * n + wrap(r + ki - K / 2, n)] * dz[s * CN + r + oc * n];
            } else {
              const int q = c - K * C;
              o = a1[s * CN + q % CN] * d2[s * H + q / CN];
            }
            v += rr[s] * o;
          }
          acc[e] += v;
        }
      }
    }
    __syncthreads();
  }
  if (MODE == kVjp) {
    R* part = static_cast<R*>(A.partial) + (size_t)blockIdx.x * npar;
#pragma unroll
    for (int e = 0; e < EMAX; ++e) {
      const int c = tid + 256 * e;
      if (c < npar) part[c] = acc[e];
    }
  }
}

template <typename R>
__global__ void k_ff_transpose(const R* __restrict__ W, int CN, int H, R* __restrict__ WT) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;  // q = i * H + o
  if (q < CN * H) WT[q] = W[(size_t)(q % H) * CN + q / H];
}

template <typename R>
__global__ void k_ff_reduce(const R* __restrict__ part, int blocks, int npar, R* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= npar) return;
  R v = (R)0;
  for (int b = 0; b < blocks; ++b) v += part[(size_t)b * npar + c];
  out[c] = v;
}

template <typename R>
size_t smem_bytes(const FFArgs& a) {
  constexpr int kSB = sb<R>();
  const int CN = a.C * a.n;
  return sizeof(R) * ((size_t)kSB * a.n + 3 * (size_t)kSB * CN + 2 * (size_t)kSB * a.H + kSB);
}

template <typename R>
int launch_r(const FFArgs& a, int mode, hipStream_t st) {
  const size_t lds = smem_bytes<R>(a);
  const int groups = (a.ns + sb<R>() - 1) / sb<R>();
  const int npar = a.K * a.C + a.C * a.n * a.H;
  if (mode == kFwd) {
    ProfGuard g("ff_forward", st);
    hipLaunchKernelGGL((k_ff<R, kFwd, 0>), dim3(groups), dim3(256), lds, st, a);
  } else if (mode == kJac) {
    ProfGuard g("ff_jacobian", st);
    hipLaunchKernelGGL((k_ff<R, kJac, 0>), dim3(groups), dim3(256), lds, st, a);
  } else {
    const int blocks = a.vjp_blocks > 0 ? (a.vjp_blocks < groups ? a.vjp_blocks : groups)
                                        : (groups < 512 ? groups : 512);
    ProfGuard g("ff_vjp", st);
    const int e = (npar + 255) / 256;
    if (e <= 32)
      hipLaunchKernelGGL((k_ff<R, kVjp, 32>), dim3(blocks), dim3(256), lds, st, a);
    else if (e <= 64)
      hipLaunchKernelGGL((k_ff<R, kVjp, 64>), dim3(blocks), dim3(256), lds, st, a);
    else if (e <= 96)
      hipLaunchKernelGGL((k_ff<R, kVjp, 96>), dim3(blocks), dim3(256), lds, st, a);
    else
      return YQ_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(k_ff_reduce<R>, dim3((npar + 255) / 256), dim3(256), 0, st,
                       static_cast<const R*>(a.partial), blocks, npar, static_cast<R*>(a.grad));
  }
  return YQ_OK;
}

// Per-block VJP partials of yq_ff_vjp_dev_noalloc: a fixed array instead of a
// context allocation, for callers that cannot allocate device memory.  Every
// call on a device shares that device's copy; each launch writes every
// partial it reads (rows 0 .. blocks-1, all npar entries) before the
// reduction reads them, so nothing has to start at zero.
static __device__ __align__(256) unsigned char scratch[kFFScratchBytes];

}  // namespace

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device ordinal
}  // namespace

int ff_scratch_init(int device) {
  if (device < 0) return YQ_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[(size_t)device]) return YQ_OK;
  // the array's host-side name is not a device address: take the address the
  // runtime gave this device's copy
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return YQ_ERR_HIP;
  }
  g_scratch_ptr[(size_t)device] = p;
  return YQ_OK;
}

void* ff_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && (size_t)device < g_scratch_ptr.size() ? g_scratch_ptr[(size_t)device]
                                                              : nullptr;
}

int ff_vjp_blocks(int ns, int use_double) {
  const int groups = (ns + (use_double ? 3 : 7)) / (use_double ? 4 : 8);
  return groups < 512 ? groups : 512;
}

size_t ff_smem_bytes(const FFArgs& a, int use_double) {
  return use_double ? smem_bytes<double>(a) : smem_bytes<float>(a);
}

void launch_ff_transpose(const FFArgs& a, int use_double, hipStream_t st) {
  const int CN = a.C * a.n, cnt = CN * a.H, KC = a.K * a.C;
  const dim3 grid((cnt + 255) / 256);
  if (use_double)
    hipLaunchKernelGGL(k_ff_transpose<double>, grid, dim3(256), 0, st,
                       static_cast<const double*>(a.params) + KC, CN, a.H,
                       static_cast<double*>(const_cast<void*>(a.wt)));
  else
    hipLaunchKernelGGL(k_ff_transpose<float>, grid, dim3(256), 0, st,
                       static_cast<const float*>(a.params) + KC, CN, a.H,
                       static_cast<float*>(const_cast<void*>(a.wt)));
}

int launch_ff(const FFArgs& a, int mode, int use_double, hipStream_t st) {
  if (a.layers) return launch_ffs(a, mode, use_double, st);
  return use_double ? launch_r<double>(a, mode, st) : launch_r<float>(a, mode, st);
}

}  // namespace yq
