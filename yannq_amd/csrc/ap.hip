// AmplitudePhase machine (Machines/AmplitudePhase.hpp): psi(sigma) =
// sqrt(RBM(sigma)) exp(i pi FF(sigma)).  The heavy passes are the RBM and
// FeedForward kernels themselves (E_loc ratios with the phase network's
// outputs folded in, eloc.hip; FF forward over each sample's neighbours and
// the FF Jacobian, ff.hip); the kernels here only recombine their outputs.
#include "yq_internal.h"

namespace yq {
namespace {

__global__ void k_ap_log_psi(const double* __restrict__ lp, const double* __restrict__ ff, int ns,
                             double2* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ns) out[s] = make_double2(lp[s] / 2.0, M_PI * ff[s]);  // AmplitudePhase.hpp:127-138
}

__global__ void k_split_complex(const double2* __restrict__ z, int ns, double* __restrict__ re,
                                double* __restrict__ im) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const double2 v = z[s];
  re[s] = v.x;
  im[s] = v.y;
}

}  // namespace

void launch_ap_log_psi(const double* lp, const double* ff, int ns, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_ap_log_psi, dim3((ns + 255) / 256), dim3(256), 0, st, lp, ff, ns,
                     reinterpret_cast<double2*>(out));
}

void launch_split_complex(const double* z, int ns, double* re, double* im, hipStream_t st) {
  hipLaunchKernelGGL(k_split_complex, dim3((ns + 255) / 256), dim3(256), 0, st,
                     reinterpret_cast<const double2*>(z), ns, re, im);
}

}  // namespace yq
